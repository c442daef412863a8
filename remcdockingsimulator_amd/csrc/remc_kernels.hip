// remc_kernels.hip -- gfx950 kernels of the REMC hot path.
//
//   k_mc<BLOCK,CACHE,SPEC>  batched persistent MC: 1-8 workgroups per replica
//                      run nsteps steps (proposal, dE of the moved body against
//                      all other beads, FP64 Metropolis, commit) without
//                      leaving the GPU, small moves in speculative windows.
//                      Replaces Replica::MCSearch's per-step
//                      MoleculeDataToDevice/EonDevice round trip
//                      (Replica.cpp:368-478).
//   k_mc_fb<BLOCK,CACHE>  force-biased (smart) MC steps: 1-8 workgroups per
//                      replica (large moves split), one-bead crowder moves in
//                      windows.
//   k_adjoint_*        gradient of one replica's energy: per-body force and
//                      torque, d/d(lambda, e0, xi).
//   k_energy_full      batched full energy, every replica in one launch,
//                      256-bead tiles, half-window tile schedule (each
//                      unordered tile pair once), Kahan FP32 per thread,
//                      FP64 per tile; k_sum_partials adds tiles in fixed order.
//   k_energy_plain     the reference-shaped kernel (E_TiledKernel,
//                      CudaFunctions.cu:455-685): (G,G) grid of B-bead tiles,
//                      lower triangle culled, FP32 partial per tile, diagonal
//                      tile halved.  The host adds the partials in FP64.
//   k_exchange         temperature-swap replica exchange per system
//                      (main.cpp:1427-1453).
//
// Pair term (Replica.cpp:584-600,694-733; CudaFunctions.cu:430-452,536-649):
// minimum image, r = |d|; crowder pairs c*(6/r)^6 for r < 15; else
// E_ij*(...) contact term with the repulsive flip plus Debye-Hueckel.  The
// table etab already holds 4*lambda*(T_ij - e0)*lj_cf*kbt_cf, the DH
// coefficient dh_cf*kbt_cf, so a pair's value is directly kcal/mol.
// x^6 is formed by multiplication; exp(-r/xi) = exp2(r*k_dh) with the
// hardware exp2; 1/r from the hardware reciprocal square root.
#include <algorithm>
#include <type_traits>

#include "remc_internal.h"
#include "remc_math.h"

namespace remc {

// A load from LDS through a pointer the compiler cannot place (it derives
// from the dynamic LDS block, or meets a global pointer in a select, which
// becomes a flat load): the address-space cast makes it an LDS read.
template <typename T>
__device__ __forceinline__ T lds_ld(const T *ptr)
{
    return *(const __attribute__((address_space(3))) T *)ptr;
}

__device__ __forceinline__ float min_image(float d, float box, float inv_box)
{
    return __builtin_fmaf(-box, __builtin_rintf(d * inv_box), d);
}

// Energy of one non-crowder pair from squared distance.  eij4 = 4 E_ij (the
// table holds the factor 4, so LJ = E (4 x6 (x6 - 1)) costs one FMA: scaling by
// 4 is exact, the result is the same bits as forming 4 x6 (x6 - 1) and E).
// SKIP_UNCHARGED: the Debye-Hueckel term is formed under a branch on qq != 0.
// Four residue types of twenty carry a charge, so where the charge product
// is the same for a whole wave (one partner bead per wave in the full energy,
// one moved bead per wave in the MC pair loops) most waves skip the exp2 and
// its multiplies.  The skipped term is +-0, so the sum is the same number;
// coincident beads still give NaN from the LJ part (r2 = 0: rinv = inf,
// x6 = inf, inf*inf - inf) and hence the 1e30 cap.  The plain per-replica
// kernel forms every term.
template <bool SKIP_UNCHARGED = true>
__device__ __forceinline__ float pair_lj_dh(float r2, float eij4, float sij2, float qq, const DevFF &ff)
{
    float rinv = __builtin_amdgcn_rsqf(r2);
    float s2 = sij2 * (rinv * rinv);
    float x6 = s2 * s2 * s2;
    float base = __builtin_fmaf(x6, x6, -x6);           // x6 (x6 - 1)
    bool flip = (eij4 > 0.0f) & (x6 > 0.5f);            // r < 2^(1/6) sigma
    float e = eij4 * (flip ? base + 0.5f : -base);      // E (LJ + 2) when flipped, else -E LJ
    if (!SKIP_UNCHARGED || qq != 0.0f) {
        float r = r2 * rinv;
        // keeps the exp2 inside the branch (otherwise it is hoisted above it
        // and the branch becomes a select)
        if (SKIP_UNCHARGED) asm volatile("" : "+v"(r));
        e = e + qq * __builtin_amdgcn_exp2f(r * ff.k_dh) * rinv;
    }
    // coincident beads (r2 = 0: rinv = inf, r = 0 * inf) give NaN; the cap
    // (IEEE minNum returns the number) turns it, like any overflow, into 1e30
    // so compensated sums stay finite
    return fminf(e, 1e30f);
}

// pair_lj_dh of an uncharged pair (qq = 0) without the Debye-Hueckel branch,
// with (sigma/r)^2 from one reciprocal instead of a squared reciprocal root
// (the two differ in the last bits; a difference en - eo is always formed
// from one of the two forms).
__device__ __forceinline__ float pair_lj(float r2, float eij4, float sij2)
{
    float s2 = sij2 * __builtin_amdgcn_rcpf(r2);
    float x6 = s2 * s2 * s2;
    float base = __builtin_fmaf(x6, x6, -x6);
    bool flip = (eij4 > 0.0f) & (x6 > 0.5f);
    float e = eij4 * (flip ? base + 0.5f : -base);
    return fminf(e, 1e30f);
}

// pair_lj_dh of every pair, charged or not, with the Debye-Hueckel exponent
// scale given (k_dh times the length unit of r2).
__device__ __forceinline__ float pair_lj_dh_k(float r2, float eij4, float sij2, float qq, float kdh)
{
    float rinv = __builtin_amdgcn_rsqf(r2);
    float s2 = sij2 * (rinv * rinv);
    float x6 = s2 * s2 * s2;
    float base = __builtin_fmaf(x6, x6, -x6);
    bool flip = (eij4 > 0.0f) & (x6 > 0.5f);
    float e = eij4 * (flip ? base + 0.5f : -base);
    e = e + qq * __builtin_amdgcn_exp2f(r2 * rinv * kdh) * rinv;
    return fminf(e, 1e30f);
}

// Crowder term C (sigma_c / r)^6 inside the cutoff as crow_k (1/r^2)^3.  r2 = 0
// (coincident beads) gives inf, capped like the pair term; no clamp is needed
// since no zero multiplies the infinity.
__device__ __forceinline__ float pair_crowder(float r2, const DevFF &ff)
{
    const float rc = __builtin_amdgcn_rcpf(r2);
    const float rc3 = rc * rc * rc;
    return r2 < ff.crow_cut2 ? fminf(ff.crow_k * rc3, 1e30f) : 0.0f;
}

// pair_crowder with the squared cutoff and the constant given (in the length
// unit of r2).
__device__ __forceinline__ float pair_crowder_u(float r2, float cut2, float ck)
{
    const float rc = __builtin_amdgcn_rcpf(r2);
    const float rc3 = rc * rc * rc;
    return r2 < cut2 ? fminf(ck * rc3, 1e30f) : 0.0f;
}

// Squared screening radius of a one-bead crowder move from o to n: a partner
// bead farther than the cutoff plus the move's length (+0.01 A for the FP32
// rounding of the distances) from o is beyond the cutoff of both positions,
// so both of its crowder terms are exactly 0.
__device__ __forceinline__ float crowder_screen2(float4 o, float4 n, const DevSys &S, const DevFF &ff)
{
    const float dx = min_image(n.x - o.x, S.box, S.inv_box), dy = min_image(n.y - o.y, S.box, S.inv_box),
                dz = min_image(n.z - o.z, S.box, S.inv_box);
    const float lim = ff.crow_cut + sqrtf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz))) + 0.01f;
    return lim * lim;
}

// Kahan step, kept as written (no reassociation is enabled for this file).
__device__ __forceinline__ void kahan_add(float &s, float &c, float v)
{
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// --------------------------------------------------------------------------
// batched persistent MC
//
// Grid = coop * n_rep workgroups; workgroup b serves replica r = b % n_rep as
// member c = b / n_rep of its cooperative group (so with n_rep a multiple of 8
// the members of one replica share an XCD and its L2).  Member c keeps a
// private copy c of the replica's dynamic state and evaluates dE over the
// partner-bead slice [c*N/coop, (c+1)*N/coop).  Members publish their partial
// dE per split step (per window: one per slot) in a 64-bit slot (value with
// its low 20 bits replaced by the step
// tag, written with one atomic store) and every member adds the coop partials
// in member order, so all members take the identical decision and update
// their copies identically.  With coop = 1 no slot is used.  A member that
// waits longer than a bounded number of polls sets the error flag; every
// member that sees the flag leaves the step loop, so the grid always drains.
// --------------------------------------------------------------------------
constexpr unsigned long long SLOT_TAG_MASK = 0xFFFFFull;   // 0xFFFFF = empty slot
constexpr int SLOT_SPIN_LIMIT = 1 << 22;

__device__ __forceinline__ unsigned long long slot_pack(double v, unsigned long long tag)
{
    return ((unsigned long long)__double_as_longlong(v) & ~SLOT_TAG_MASK) | tag;
}

__device__ __forceinline__ double slot_value(unsigned long long w)
{
    return __longlong_as_double((long long)(w & ~SLOT_TAG_MASK));
}

// floor(a / b) and ceil(a / b) for 0 <= a < 2^22, 0 < b < 2^22: the float
// quotient is within one of the true one, and one correction step each way
// makes it exact.
__device__ __forceinline__ int div_floor_small(int a, int b)
{
    int q = (int)((float)a * __builtin_amdgcn_rcpf((float)b));
    if (q * b > a) --q;
    if ((q + 1) * b <= a) ++q;
    return q;
}

__device__ __forceinline__ int div_ceil_small(int a, int b) { return div_floor_small(a + b - 1, b); }

// floor(c n / d) for 0 <= c <= d and n >= 0 in 32-bit arithmetic (the bound
// of member c's share of n partners over d members): with n = q d + r it is
// c q + floor(c r / d), where c q <= n and c r < d^2
__device__ __forceinline__ int share_bound(int c, int n, int d)
{
    const unsigned q = (unsigned)n / (unsigned)d, r = (unsigned)n - q * (unsigned)d;
    return (int)((unsigned)c * q + ((unsigned)c * r) / (unsigned)d);
}

// The split choice (G, kper) of the last two move shapes (moved beads,
// partner beads): a move of either shape takes it without the search.
struct SplitMemo {
    int nm[2] = {-1, -1}, nj[2] = {-1, -1}, G[2] = {1, 1}, kper[2] = {1, 1};
};

// The split of a large move's (moved bead x partner bead) pairs over a
// workgroup, for pair loops (MC and force-biased) that hold two partners per
// lane per pass and run all terms of a pass whether or not the second partner
// exists: lane l of group g takes partners l, l + G, ... against moved beads
// [g kper, (g + 1) kper).  Lane i of each wave prices g = i + 1 groups of
// G = BLOCK / g lanes: ceil(ceil(nj / G) / 2) passes, each over
// kper = ceil(nm / g) moved beads plus a quarter for the pass's partner loads,
// 5% more where group boundaries fall inside waves (a wave then spans several
// moved beads and takes the charged-pair branch more often); the cheapest,
// lowest g on ties, is taken by every wave alike.
template <int BLOCK>
__device__ __forceinline__ void pick_split_pairs(int nm, int nj, int lane, int &G, int &kper)
{
    if (nj <= 0) {
        G = 1;
        kper = nm;
        return;
    }
    const int g = lane + 1;
    const int Gl = div_floor_small(BLOCK, g);
    unsigned int key = 0xFFFFFFFFu;
    if (Gl >= 1) {
        const int passes = (div_ceil_small(nj, Gl) + 1) >> 1;
        const int kp = div_ceil_small(nm, g);
        const long long cost = (long long)passes * (kp * ((Gl & 63) ? 105 : 100) + 25);
        key = (unsigned int)min(cost, (long long)(1 << 25) - 1) << 6 | (unsigned int)lane;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) key = min(key, (unsigned int)__shfl_xor((int)key, off, 64));
    const int gb = (int)(key & 63u) + 1;
    G = div_floor_small(BLOCK, gb);
    kper = div_ceil_small(nm, gb);
}

// pick_split_pairs through the memo m (nm and nj uniform over the workgroup)
template <int BLOCK>
__device__ __forceinline__ void pick_split_memo(SplitMemo &m, int nm, int nj, int lane, int &G, int &kper)
{
    if (nm == m.nm[0] && nj == m.nj[0]) {
        G = m.G[0];
        kper = m.kper[0];
    } else if (nm == m.nm[1] && nj == m.nj[1]) {
        G = m.G[1];
        kper = m.kper[1];
    } else {
        pick_split_pairs<BLOCK>(nm, nj, lane, G, kper);
        G = __builtin_amdgcn_readfirstlane(G);
        kper = __builtin_amdgcn_readfirstlane(kper);
        m.nm[1] = m.nm[0];
        m.nj[1] = m.nj[0];
        m.G[1] = m.G[0];
        m.kper[1] = m.kper[0];
        m.nm[0] = nm;
        m.nj[0] = nj;
        m.G[0] = G;
        m.kper[0] = kper;
    }
}

// Draws of one MC step and the move parameters that depend on them only
// (formed ahead of the step, one step per lane).
struct StepDraw {
    double nlu;      // -ln(u) of the Metropolis uniform
    double coef[9];  // rotation coefficients (remc_qcoeffs) of a rotate move
    double q[4];     // rotation quaternion of a rotate move
    float t[3];      // translation vector of a translate move
    int body;
    int rotate;
    int m0, nm;       // the moved body's first bead and bead count
    int prev1, prev2; // the latest two earlier steps of the batch moving the same body (-1: none)
};
constexpr int DRAW_BATCH = 64;

// 64-bit DPP move of a double (both halves with the same control word).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int l2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, true);
    const int h2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, true);
    return __hiloint2double(h2, l2);
}

// Sum over the 64 lanes of a wave in a fixed order, with DPP row shifts and
// row broadcasts (no LDS traffic); the result is uniform.
__device__ __forceinline__ double wave_sum_dpp(double v)
{
    v += dpp_f64<0x111, 0xf>(v); // row_shr:1
    v += dpp_f64<0x112, 0xf>(v); // row_shr:2
    v += dpp_f64<0x114, 0xf>(v); // row_shr:4
    v += dpp_f64<0x118, 0xf>(v); // row_shr:8: lane 15 of each row holds the row sum
    v += dpp_f64<0x142, 0xa>(v); // row_bcast:15: rows 1 and 3 add the previous row's sum
    v += dpp_f64<0x143, 0xc>(v); // row_bcast:31: rows 2 and 3 add lane 31
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Squared distance in units of the box between box-scaled positions (x/box):
// the minimum image of a scaled difference d is d - rint(d).
__device__ __forceinline__ float dist2_unit(float xj, float yj, float zj, float4 a)
{
    float dx = xj - a.x, dy = yj - a.y, dz = zj - a.z;
    dx -= __builtin_rintf(dx);
    dy -= __builtin_rintf(dy);
    dz -= __builtin_rintf(dz);
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}

__device__ __forceinline__ float dist2(float xj, float yj, float zj, float4 a, const DevSys &S)
{
    const float dx = min_image(xj - a.x, S.box, S.inv_box);
    const float dy = min_image(yj - a.y, S.box, S.inv_box);
    const float dz = min_image(zj - a.z, S.box, S.inv_box);
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}

// Dynamic LDS of the MC kernel, carved in this order (mc_dynamic_lds sizes it):
//   draw2:                  StepDraw[DRAW_BATCH]              the draw batch formed ahead
//   cache_quat:             double q[4*nbody]                 body rotations
//   always:                 float4 old[mb], new[mb], olds[mb], news[mb]; int tc[mb]
//                           moved-body staging (olds/news: box-scaled copies)
//   CACHE:                  float x,y,z,q,sig[cl]; int tc[cl]  partner beads [clo, clo+cl)
//   CACHE && cache_rel:     float rx,ry,rz[cl]                 relative positions
//   CACHE && cache_bodies:  int start,end[nbody]; float cx,cy,cz[nbody]
// mb, cl and nbody are the launch-wide maxima (KParams::lds_*).
size_t mc_dynamic_lds(const KParams &p)
{
    const size_t mb = p.lds_mb, cl = p.lds_cl, nb = p.lds_nbody;
    size_t s = 68 * mb;
    if (p.draw2) s += mc_draw_batch_bytes();
    if (p.cache_quat) s += 32 * nb;
    if (cl) s += 24 * cl;
    if (cl && p.cache_rel) s += 12 * cl;
    if (cl && p.cache_bodies) s += 20 * nb;
    return s;
}

size_t mc_draw_batch_bytes() { return sizeof(StepDraw) * DRAW_BATCH; }

// One launch runs nsteps MC steps of every replica.  A step whose moved body
// is small (at most SPEC_MAX beads) may be taken in a speculative window when
// SPEC is set: the next w <= WK such steps are evaluated at once, two slots per
// wave (with cooperating members, each over its partner slice, the partials
// exchanged once per window), each against the conformation at the window's
// start (dE0_k).  The exact dE of step k after the accepted steps j < k of the
// window differs from dE0_k only by the pair terms between the two moved
// bodies,
//   corr(j,k) = sum_{i in k, l in j} [e(n_i,n_l) - e(o_i,n_l)] - [e(n_i,o_l) - e(o_i,o_l)],
// which are formed for every pair of slots in parallel.  Every slot is decided
// on its dE0; the window ends before the first slot k that an accepted earlier
// slot j interacts with (corr(j,k) != 0, or the same body, whose proposal
// depends on j's move), so every kept decision used its exact dE; slot k starts
// the next window.  A slot whose body exactly one earlier slot j moved is
// staged from j's proposal instead and ends the window only if j is rejected.
// Decisions, energies and states are those of taking the steps one at a time,
// up to the FP32 rounding of the reassociated pair sums.
template <int BLOCK, bool CACHE, bool SPEC>
__global__ __launch_bounds__(BLOCK) void k_mc(KParams p, int nsteps, int coop)
{
    constexpr int NW = BLOCK / 64;
    constexpr int WK = SPEC ? 2 * NW : 1; // window slots: at most 32 (the slot masks are 32-bit)
    static_assert(WK <= 32 && WK <= REMC_SLOT_K, "window slots");
    constexpr int WS = SPEC ? WK * SPEC_MAX : 1;
    __shared__ float s_etab[REMC_NTYPES * REMC_NTYPES];
    // (4 E_ij, and with p.sigtab the replica's system's sigma_ij^2 / box^2) per type pair
    __shared__ float2 s_et2[REMC_NTYPES * REMC_NTYPES];
    __shared__ double s_red[NW];
    __shared__ double s_dE;
    __shared__ int s_abort;
    __shared__ StepDraw s_draw0[DRAW_BATCH];
    __shared__ unsigned long long s_elig[2]; // window-eligible steps of each draw batch
    __shared__ float4 s_wold[WS], s_wnew[WS], s_wrel[WS]; // old (w: q c_dh) / new (w: sigma) position, new rel (w: q)
    __shared__ int s_wtc[WS];                             // type | crowder << 8 of the staged beads
    __shared__ float4 s_wcen[WK];                         // new centre of a translated slot
    __shared__ double s_wq[4 * WK];                       // new rotation of a rotated slot
    __shared__ double s_wbase[WK], s_wdE[WK];
    __shared__ unsigned int s_wint[WK]; // bit j of slot k: slot j's move changes slot k's dE
    // bit j of slot k: slot k moves the body slot j moved (j the only earlier
    // such slot), and is staged from slot j's proposal
    __shared__ unsigned int s_wchain[WK];
    __shared__ int s_wnext[WK]; // the accepted slot chained to slot j, or -1
    __shared__ int s_wdec[WK];
    __shared__ int s_wstat[WK]; // 1: the slot is static (dE = 0, nothing to evaluate)
    __shared__ int s_weff;
    __shared__ double s_E;
    __shared__ unsigned long long s_cnt[3];
    extern __shared__ float4 smem[];

    char *lp = (char *)smem;
    StepDraw *s_draw1 = (StepDraw *)lp; // the second batch (p.draw2)
    lp += p.draw2 ? sizeof(StepDraw) * DRAW_BATCH : 0;
    double *s_q = (double *)lp;
    lp += p.cache_quat ? 32 * (size_t)p.lds_nbody : 0;
    float4 *s_old = (float4 *)lp;
    float4 *s_new = s_old + p.lds_mb;
    float4 *s_olds = s_new + p.lds_mb, *s_news = s_olds + p.lds_mb;
    int *s_tc = (int *)(s_news + p.lds_mb);
    float *s_jx = (float *)(s_tc + p.lds_mb);
    float *s_jy = s_jx + p.lds_cl, *s_jz = s_jy + p.lds_cl, *s_jq = s_jz + p.lds_cl, *s_js = s_jq + p.lds_cl;
    int *s_jtc = (int *)(s_js + p.lds_cl);
    const int rel_n = p.cache_rel ? p.lds_cl : 0;
    float *s_rx = (float *)(s_jtc + p.lds_cl), *s_ry = s_rx + rel_n, *s_rz = s_ry + rel_n;
    int *s_bs = (int *)(s_rz + rel_n), *s_be = s_bs + p.lds_nbody;
    float *s_cx = (float *)(s_be + p.lds_nbody), *s_cy = s_cx + p.lds_nbody, *s_cz = s_cy + p.lds_nbody;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int r = blockIdx.x % p.n_rep;
    const int c = blockIdx.x / p.n_rep;
    const DevRep R = p.rep[r];
    const DevSys S = p.sys[R.sys];
    const DevFF ff = p.ff;

    for (int k = tid; k < REMC_NTYPES * REMC_NTYPES; k += BLOCK) {
        s_etab[k] = p.etab[k];
        s_et2[k] = make_float2(p.etab[k], p.sigtab ? p.sstab[(size_t)R.sys * (REMC_NTYPES * REMC_NTYPES) + k] : 0.0f);
    }
    if (tid == 0) s_abort = 0;

    // the tracked energy and the counters live in LDS, updated by thread 0
    // (fewer loop-carried registers per thread)
    if (tid == 0) {
        s_E = p.E[r];
        s_cnt[0] = s_cnt[1] = s_cnt[2] = 0ull;
    }
    const double beta = remc_beta(p.T[r]);
    unsigned long long step = p.step[r];

    // private copy c of this replica's dynamic state (always current in HBM)
    const size_t cb = (size_t)c * p.copy_beads + R.bead_off;
    const size_t cmo = (size_t)c * p.copy_bodies + R.body_off;
    float *gx = p.px + cb, *gy = p.py + cb, *gz = p.pz + cb;
    float *grx = p.rx + cb, *gry = p.ry + cb, *grz = p.rz + cb;
    float *gcx = p.cx + cmo, *gcy = p.cy + cmo, *gcz = p.cz + cmo;
    double *gq = p.quat + 4 * cmo;
    const int *tcs = p.tc + S.bead_off;
    const float *qs = p.q + S.bead_off, *sgs = p.sig + S.bead_off;
    const int N = S.n_beads;
    const int cjs = (int)(((long long)c * N) / coop), cje = (int)(((long long)(c + 1) * N) / coop);
    unsigned long long *slots = p.slots + (size_t)r * REMC_SLOTS_PER_REP;
    unsigned int n_x = 0; // partial-sum exchanges so far in this launch: selects the slot bank

    // cached partner beads: all of them, or only this member's slice when
    // every move of the launch is split (cache_slice)
    const int clo = (CACHE && p.cache_slice) ? cjs : 0;
    const int chi = CACHE ? ((p.cache_slice) ? cje : N) : 0;
    const bool cache_bodies = CACHE && p.cache_bodies;
    const bool cache_rel = CACHE && p.cache_rel;
    const bool cache_quat = p.cache_quat != 0;
    if (CACHE) {
        for (int b = clo + tid; b < chi; b += BLOCK) {
            const int k = b - clo;
            s_jx[k] = gx[b];
            s_jy[k] = gy[b];
            s_jz[k] = gz[b];
            s_jq[k] = qs[b];
            s_js[k] = sgs[b];
            s_jtc[k] = tcs[b];
            if (cache_rel) {
                s_rx[k] = grx[b];
                s_ry[k] = gry[b];
                s_rz[k] = grz[b];
            }
        }
        if (cache_bodies)
            for (int m = tid; m < S.n_bodies; m += BLOCK) {
                s_bs[m] = p.bstart[S.body_off + m];
                s_be[m] = p.bend[S.body_off + m];
                s_cx[m] = gcx[m];
                s_cy[m] = gcy[m];
                s_cz[m] = gcz[m];
            }
    }
    if (cache_quat)
        for (int m = tid; m < 4 * S.n_bodies; m += BLOCK) s_q[m] = gq[m];
    __syncthreads();

    int b0 = 0, have = 0; // s_draw holds the draws of steps [b0, b0 + have) of this launch
    unsigned long long elig = 0ull; // s_elig of the batch, read once per batch
    // With p.draw2 the draws of steps [nb0, nb0 + DRAW_BATCH), nb0 = b0 +
    // DRAW_BATCH/2, are formed ahead into the other buffer by four waves that
    // a window's staging leaves idle; the refill then only switches buffers.
    int dbuf = 0;        // the buffer of the current batch
    bool nb_ok = false;  // the other buffer holds the batch from nb0
    int nb0 = 0;
    int pre_w = 0;       // slots of the next window, staged ahead (0: none)
    int pre_it = -1;     // the large move staged ahead (launch step index; -1: none)
    SplitMemo smemo;     // the split of the latest large-move shapes
    int it = 0;
    while (it < nsteps) {
        // The thread index is re-read opaquely in every step, so that the
        // per-thread LDS addresses derived from it are formed where they are
        // used (one or two instructions each) instead of being hoisted out of
        // the loop, where they outnumber the registers and go to scratch.
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & 63, wave = tid >> 6;
        const unsigned long long t_step = p.rec ? wall_clock64() : 0ull;
        int off = it - b0;
        bool refill = off >= have;
        if (SPEC && !refill && off + WK > have && b0 + have < nsteps) {
            // fewer than WK draws left: form the next batch now if the run of
            // window-eligible steps from here reaches the batch's end (else a
            // large move ends the window before it anyway)
            const unsigned long long rest = (1ull << (have - off)) - 1ull; // have - off < WK <= 32
            refill = ((~(elig >> off)) & rest) == 0ull;
        }
        // The draws of DRAW_BATCH steps from launch step base (RNG counter
        // cbase), one step per lane, into dst, formed by four waves side by
        // side (role 0..3): each forms the counter words itself and then its
        // part (role 0 the body, its bead range, the window eligibility and the
        // same-body links; role 1 a rotation's quaternion and coefficients;
        // role 2 a translation; role 3 -ln u).
        auto form_draws = [&](StepDraw *dst, int eb, int base, unsigned long long cbase, int role) {
            if (role < 0 || role >= 4) return;
            const bool valid = base + lane < nsteps;
            remc_draws dd{};
            if (valid) dd = remc_step_draws(cbase + lane, R.rng_id, p.key0, p.key1, S.n_bodies);
            StepDraw &sd = dst[lane];
            if (role == 0) {
                bool el = false;
                const int body_l = valid ? dd.body : -1;
                if (valid) {
                    sd.body = dd.body;
                    sd.rotate = dd.rotate;
                    const int gb = S.body_off + dd.body;
                    sd.m0 = cache_bodies ? lds_ld(s_bs + dd.body) : p.bstart[gb];
                    sd.nm = (cache_bodies ? lds_ld(s_be + dd.body) : p.bend[gb]) - sd.m0;
                    if (SPEC) el = sd.nm <= SPEC_MAX;
                }
                if (SPEC) {
                    const unsigned long long m = __ballot(el);
                    if (lane == 0) s_elig[eb] = m;
                    // the latest two earlier steps of the batch that move the same
                    // body: the lanes whose body index agrees with this lane's in
                    // every bit (one ballot per bit); lanes past the launch's last
                    // step lie above every valid lane, so never count as earlier
                    const int nbits = S.n_bodies > 1 ? 32 - __builtin_clz((unsigned int)(S.n_bodies - 1)) : 0;
                    unsigned long long same = ~0ull;
                    for (int bt = 0; bt < nbits; ++bt) {
                        const bool on = ((body_l >> bt) & 1) != 0;
                        const unsigned long long mb = __ballot(on);
                        same &= on ? mb : ~mb;
                    }
                    const unsigned long long below = same & ((1ull << lane) - 1ull);
                    int p1 = -1, p2 = -1;
                    if (below != 0ull) {
                        p1 = 63 - __builtin_clzll(below);
                        const unsigned long long rest = below & ~(1ull << p1);
                        if (rest != 0ull) p2 = 63 - __builtin_clzll(rest);
                    }
                    if (valid) {
                        sd.prev1 = p1;
                        sd.prev2 = p2;
                    }
                }
            } else if (role == 1) {
                if (valid && dd.rotate) {
                    double qq[4], cf[9];
                    remc_move_quat(dd.a, dd.b, dd.c, R.rot_ch, R.rot_sh, qq);
                    remc_qcoeffs(qq, cf);
                    for (int i = 0; i < 4; ++i) sd.q[i] = qq[i];
                    for (int i = 0; i < 9; ++i) sd.coef[i] = cf[i];
                }
            } else if (role == 2) {
                if (valid && !dd.rotate) remc_move_translation(dd.a, dd.b, dd.c, R.tstep, &sd.t[0], &sd.t[1], &sd.t[2]);
            } else {
                if (valid) sd.nlu = remc_neglog_u(dd.u);
            }
        };
        if (refill) {
            if (nb_ok && it >= nb0) {
                // the batch formed ahead covers this step and the rest of the old one
                dbuf ^= 1;
                b0 = nb0;
                have = min(DRAW_BATCH, nsteps - nb0);
                off = it - b0;
                nb_ok = false;
                if (SPEC) elig = s_elig[dbuf];
            } else {
                pre_w = 0; // a window staged ahead came from the old batch
                form_draws(dbuf ? s_draw1 : s_draw0, dbuf, it, step, wave);
                b0 = it;
                have = min(DRAW_BATCH, nsteps - it);
                off = 0;
                nb_ok = false;
                __syncthreads();
                if (SPEC) elig = s_elig[dbuf];
            }
        }
        StepDraw *const s_draw = dbuf ? s_draw1 : s_draw0;
        // Staging of a large move (draws Dm) into s_old / s_new, their
        // box-scaled copies and s_tc, by threads t0, t0 + stride, ... over its
        // beads.  It reads only the moved body.
        auto stage_move = [&](const int t0, const int stride, const StepDraw &Dm) {
            // the bead range comes with the draw: the centre and the beads'
            // reads wait on the draw only
            const int bd = Dm.body, mb0 = Dm.m0, mb1 = Dm.m0 + Dm.nm;
            float bcx, bcy, bcz;
            if (cache_bodies) {
                bcx = lds_ld(s_cx + bd);
                bcy = lds_ld(s_cy + bd);
                bcz = lds_ld(s_cz + bd);
            } else {
                bcx = gcx[bd];
                bcy = gcy[bd];
                bcz = gcz[bd];
            }
            float nbx = bcx, nby = bcy, nbz = bcz;
            if (!Dm.rotate) remc_wrap_centre(bcx, bcy, bcz, Dm.t[0], Dm.t[1], Dm.t[2], S.box, &nbx, &nby, &nbz);
            for (int k = t0; k < mb1 - mb0; k += stride) {
                const int b = mb0 + k;
                const bool inc = CACHE && b >= clo && b < chi;
                float rx, ry, rz;
                if (cache_rel) {
                    rx = lds_ld(s_rx + b);
                    ry = lds_ld(s_ry + b);
                    rz = lds_ld(s_rz + b);
                } else {
                    rx = grx[b];
                    ry = gry[b];
                    rz = grz[b];
                }
                float nrx = rx, nry = ry, nrz = rz;
                if (Dm.rotate) remc_qrotate(Dm.coef, rx, ry, rz, &nrx, &nry, &nrz);
                float4 ov, nv;
                if (inc) {
                    const int kb = b - clo;
                    ov = make_float4(lds_ld(s_jx + kb), lds_ld(s_jy + kb), lds_ld(s_jz + kb), lds_ld(s_jq + kb) * ff.c_dh);
                    nv = make_float4(nrx + nbx, nry + nby, nrz + nbz, lds_ld(s_js + kb));
                } else {
                    ov = make_float4(gx[b], gy[b], gz[b], qs[b] * ff.c_dh);
                    nv = make_float4(nrx + nbx, nry + nby, nrz + nbz, sgs[b]);
                }
                s_old[k] = ov;
                s_new[k] = nv;
                // box-scaled about the body's old centre: the beads near it, which
                // give the steep contact terms, keep every bit of their offsets
                const float ib = S.inv_box;
                s_olds[k] = make_float4((ov.x - bcx) * ib, (ov.y - bcy) * ib, (ov.z - bcz) * ib, ov.w * ib);
                s_news[k] = make_float4((nv.x - bcx) * ib, (nv.y - bcy) * ib, (nv.z - bcz) * ib, 0.5f * nv.w);
                s_tc[k] = inc ? lds_ld(s_jtc + (b - clo)) : tcs[b];
            }
        };
        // Staging of a speculative window's slots [0, wn) from step off_ of
        // the draw batch sd, one thread st per slot bead (slot st / SPEC_MAX,
        // bead st % SPEC_MAX): the moved beads' old and new positions, the
        // slot's static and chained flags, and its new rotation or centre.  It
        // reads only the slots' own bodies.
        auto stage_window = [&](const int st, const int off_, const int wn, const StepDraw *sd) {
            const int k = st / SPEC_MAX, i = st % SPEC_MAX;
            if (k < wn) {
                const StepDraw &Dk = sd[off_ + k];
                const int bk = Dk.body;
                const int m0k = Dk.m0, nmk = Dk.nm;
                // One earlier slot j of the window moves the same body: slot k
                // is chained, staged from the state j's proposal leaves, and
                // stands if j is accepted.  Two or more: staged from the
                // window's start, it stands only if all of them are rejected.
                const int pj = Dk.prev1 - off_;
                const bool chained = pj >= 0 && Dk.prev2 < off_;
                if (i == 0) {
                    s_wchain[k] = chained ? 1u << pj : 0u;
                    s_wnext[k] = -1;
                }
                if (i < nmk) {
                    // the moved beads come from the partner cache when it holds
                    // them (a member's slice may not), else from HBM
                    const int b = m0k + i;
                    const bool inc = b >= clo && b < chi;
                    // typed loads: LDS reads where the data is cached (a
                    // select between the two pointers would be a flat load)
                    float rx, ry, rz, cx, cy, cz, ox, oy, oz;
                    if (cache_rel) {
                        rx = lds_ld(s_rx + b);
                        ry = lds_ld(s_ry + b);
                        rz = lds_ld(s_rz + b);
                    } else {
                        rx = grx[b];
                        ry = gry[b];
                        rz = grz[b];
                    }
                    if (cache_bodies) {
                        cx = lds_ld(s_cx + bk);
                        cy = lds_ld(s_cy + bk);
                        cz = lds_ld(s_cz + bk);
                    } else {
                        cx = gcx[bk];
                        cy = gcy[bk];
                        cz = gcz[bk];
                    }
                    if (inc) {
                        ox = lds_ld(s_jx + (b - clo));
                        oy = lds_ld(s_jy + (b - clo));
                        oz = lds_ld(s_jz + (b - clo));
                    } else {
                        ox = gx[b];
                        oy = gy[b];
                        oz = gz[b];
                    }
                    if (chained) {
                        // slot j's proposal, formed as slot j's own staging forms it
                        const StepDraw &Dj = sd[off_ + pj];
                        if (Dj.rotate) remc_qrotate(Dj.coef, rx, ry, rz, &rx, &ry, &rz);
                        else remc_wrap_centre(cx, cy, cz, Dj.t[0], Dj.t[1], Dj.t[2], S.box, &cx, &cy, &cz);
                        ox = rx + cx;
                        oy = ry + cy;
                        oz = rz + cz;
                    }
                    float nrx = rx, nry = ry, nrz = rz, ncx = cx, ncy = cy, ncz = cz;
                    if (Dk.rotate) remc_qrotate(Dk.coef, rx, ry, rz, &nrx, &nry, &nrz);
                    else remc_wrap_centre(cx, cy, cz, Dk.t[0], Dk.t[1], Dk.t[2], S.box, &ncx, &ncy, &ncz);
                    const float qb = inc ? lds_ld(s_jq + (b - clo)) : qs[b];
                    const float sb = inc ? lds_ld(s_js + (b - clo)) : sgs[b];
                    s_wold[st] = make_float4(ox, oy, oz, qb * ff.c_dh);
                    s_wnew[st] = make_float4(nrx + ncx, nry + ncy, nrz + ncz, sb);
                    // a static slot: one bead whose move leaves it where it is
                    if (i == 0)
                        s_wstat[k] = nmk == 1 && ox == nrx + ncx && oy == nry + ncy && oz == nrz + ncz ? 1 : 0;
                    s_wrel[st] = make_float4(nrx, nry, nrz, qb);
                    s_wtc[st] = inc ? lds_ld(s_jtc + (b - clo)) : tcs[b];
                    if (i == 0 && !Dk.rotate) s_wcen[k] = make_float4(ncx, ncy, ncz, 0.0f);
                }
                // a rotated slot's new rotation, formed by the slot's thread 1
                // beside bead 0's positions (SPEC_MAX > 1)
                if (i == 1 && Dk.rotate) {
                    double qo[4], qn[4];
                    if (cache_quat)
                        for (int t = 0; t < 4; ++t) qo[t] = lds_ld(s_q + 4 * bk + t);
                    else
                        for (int t = 0; t < 4; ++t) qo[t] = gq[4 * bk + t];
                    if (chained && sd[off_ + pj].rotate) {
                        remc_qmul(sd[off_ + pj].q, qo, qn);
                        for (int t = 0; t < 4; ++t) qo[t] = qn[t];
                    }
                    remc_qmul(Dk.q, qo, qn);
                    for (int t = 0; t < 4; ++t) s_wq[4 * k + t] = qn[t];
                }
            }
        };

        int w = 0;
        if (SPEC) {
            const unsigned long long run = ~(elig >> off);
            w = run == 0ull ? 64 : (int)__builtin_ctzll(run);
            w = __builtin_amdgcn_readfirstlane(min(w, WK)); // the same in every lane
        }
        // a window staged ahead, during the exchange of the large move before it
        const bool pre_ok = SPEC && pre_w > 0 && pre_w == w;
        pre_w = 0;
        if (SPEC && w > 0) {
            // ---------------- speculative window of w small moves ----------------
            // staging: thread t -> slot k = t / SPEC_MAX, bead i = t % SPEC_MAX
            // (waves 0 .. NW/2 - 1); waves NW/2 .. NW/2 + 3 meanwhile form the
            // next draw batch when it is due
            if (tid < WK) s_wint[tid] = 0u;
            if (NW >= 8 && p.draw2 && !nb_ok && off >= DRAW_BATCH / 2 && b0 + DRAW_BATCH / 2 < nsteps) {
                nb0 = b0 + DRAW_BATCH / 2;
                form_draws(dbuf ? s_draw0 : s_draw1, dbuf ^ 1, nb0, step - (unsigned long long)(it - nb0), wave - NW / 2);
                nb_ok = true;
            }
            if (!pre_ok && tid < WK * SPEC_MAX) stage_window(tid, off, w, s_draw);
            __syncthreads();

            // dE0 of slots k = wave and wave + NW: the moved beads against every
            // other bead; with cooperating members, against this member's
            // partner slice, the members' partials then exchanged once for the
            // whole window
            const unsigned long long wtag = step & 0x7FFFFull;
            unsigned long long *bank = slots + (n_x & 1u) * (REMC_SLOT_K * REMC_MAX_COOP);
            const int wjs = coop > 1 ? cjs : 0, wje = coop > 1 ? cje : N;
            // Static slots: a one-bead body whose move leaves its position as it
            // is (the rotation of a sphere) has dE = 0 exactly, with no partner
            // loop, no partial to exchange and no correction.  The other
            // (working) slots are dealt to the waves in order: ranks wave and
            // wave + NW.
            const bool st_l = lane < w && s_wstat[lane] != 0;
            const unsigned long long statm = __ballot(st_l);
            const bool wk_l = lane < w && !st_l;
            const int rank_l = __popcll(__ballot(wk_l) & ((1ull << lane) - 1ull));
            const int wv = __builtin_amdgcn_readfirstlane(wave);
            const unsigned long long ma = __ballot(wk_l && rank_l == wv), mb = __ballot(wk_l && rank_l == wv + NW);
            int ka = ma ? (int)__builtin_ctzll(ma) : WK, kb = mb ? (int)__builtin_ctzll(mb) : WK;
            if (kb < w && s_draw[off + ka].nm == 1 && s_draw[off + kb].nm == 1 &&
                ((s_wtc[ka * SPEC_MAX] & s_wtc[kb * SPEC_MAX]) >> 8) != 0) {
                // both slots of the wave move one crowder bead: one pass over the
                // partners serves both (the same per-lane order of terms as below)
                const float4 oa = s_wold[ka * SPEC_MAX], na = s_wnew[ka * SPEC_MAX];
                const float4 ob = s_wold[kb * SPEC_MAX], nb = s_wnew[kb * SPEC_MAX];
                const int ba = s_draw[off + ka].m0, bb = s_draw[off + kb].m0;
                const float lim2a = crowder_screen2(oa, na, S, ff), lim2b = crowder_screen2(ob, nb, S, ff);
                float acc_a = 0.0f, comp_a = 0.0f, acc_b = 0.0f, comp_b = 0.0f;
                for (int j = wjs + lane; j < wje; j += 64) {
                    const int jc = j - clo;
                    const float xj = s_jx[jc], yj = s_jy[jc], zj = s_jz[jc];
                    // the terms at the new positions only where a partner of the
                    // wave is near enough to give any (crowders are sparse)
                    const float r2a = dist2(xj, yj, zj, oa, S), r2b = dist2(xj, yj, zj, ob, S);
                    if (!__any((j != ba && r2a < lim2a) || (j != bb && r2b < lim2b))) continue;
                    if (j != ba) {
                        const float eo = pair_crowder(r2a, ff);
                        const float en = pair_crowder(dist2(xj, yj, zj, na, S), ff);
                        kahan_add(acc_a, comp_a, en - eo);
                    }
                    if (j != bb) {
                        const float eo = pair_crowder(r2b, ff);
                        const float en = pair_crowder(dist2(xj, yj, zj, nb, S), ff);
                        kahan_add(acc_b, comp_b, en - eo);
                    }
                }
                const double va = wave_sum_dpp((double)acc_a - (double)comp_a);
                const double vb = wave_sum_dpp((double)acc_b - (double)comp_b);
                if (lane == 0) {
                    if (coop > 1) {
                        __hip_atomic_store(bank + ka * REMC_MAX_COOP + c, slot_pack(va, wtag), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(bank + kb * REMC_MAX_COOP + c, slot_pack(vb, wtag), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    } else {
                        s_wbase[ka] = va;
                        s_wbase[kb] = vb;
                    }
                }
                ka = kb = WK; // both done
            }
            for (int pass = 0; pass < 2; ++pass) {
                const int k = pass == 0 ? ka : kb;
                if (k >= w) continue;
                const int m0k = s_draw[off + k].m0, nmk = s_draw[off + k].nm, m1k = m0k + nmk;
                const bool crow_m = (s_wtc[k * SPEC_MAX] >> 8) != 0;
                float acc = 0.0f, comp = 0.0f;
                if (nmk == 1 && crow_m) {
                    // one crowder bead: partners screened as in the two-slot pass
                    const float4 o = s_wold[k * SPEC_MAX], n = s_wnew[k * SPEC_MAX];
                    const float lim2 = crowder_screen2(o, n, S, ff);
                    for (int j = wjs + lane; j < wje; j += 64) {
                        const int jc = j - clo;
                        const float xj = s_jx[jc], yj = s_jy[jc], zj = s_jz[jc];
                        const float r2o = dist2(xj, yj, zj, o, S);
                        if (!__any(j != m0k && r2o < lim2)) continue;
                        if (j != m0k) {
                            const float eo = pair_crowder(r2o, ff);
                            const float en = pair_crowder(dist2(xj, yj, zj, n, S), ff);
                            kahan_add(acc, comp, en - eo);
                        }
                    }
                }
                for (int i = 0; i < (nmk == 1 && crow_m ? 0 : nmk); ++i) {
                    const float4 o = s_wold[k * SPEC_MAX + i], n = s_wnew[k * SPEC_MAX + i];
                    const float *erow = s_etab + (s_wtc[k * SPEC_MAX + i] & 0xff) * REMC_NTYPES;
                    for (int j = wjs + lane; j < wje; j += 64) {
                        if (j >= m0k && j < m1k) continue;
                        const int jc = j - clo; // the partner range lies in the cache
                        const float xj = s_jx[jc], yj = s_jy[jc], zj = s_jz[jc];
                        const int tcj = s_jtc[jc];
                        float eo, en;
                        if (crow_m || (tcj >> 8) != 0) {
                            eo = pair_crowder(dist2(xj, yj, zj, o, S), ff);
                            en = pair_crowder(dist2(xj, yj, zj, n, S), ff);
                        } else {
                            const float eij = erow[tcj & 0xff];
                            const float sij = 0.5f * (n.w + s_js[jc]);
                            const float sij2 = sij * sij;
                            const float qq = o.w * s_jq[jc];
                            eo = pair_lj_dh(dist2(xj, yj, zj, o, S), eij, sij2, qq, ff);
                            en = pair_lj_dh(dist2(xj, yj, zj, n, S), eij, sij2, qq, ff);
                        }
                        kahan_add(acc, comp, en - eo);
                    }
                }
                const double v = wave_sum_dpp((double)acc - (double)comp);
                if (lane == 0) {
                    if (coop > 1)
                        __hip_atomic_store(bank + k * REMC_MAX_COOP + c, slot_pack(v, wtag), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    else
                        s_wbase[k] = v;
                }
            }
            // pair corrections between the moved bodies of every two slots j < k
            // (zero where either is static)
            if (tid < WK * WK) {
                const int jj = tid / WK, kk = tid % WK;
                if (jj < kk && kk < w) {
                    const int bj = s_draw[off + jj].body, bk = s_draw[off + kk].body;
                    float cv = 0.0f;
                    // the same body: slot kk's proposal depends on slot jj's
                    // move, unless kk is chained to jj (the decision handles it)
                    bool touches = bj == bk && s_wchain[kk] == 0u;
                    if (bj != bk && ((statm >> jj) & 1ull) == 0ull && ((statm >> kk) & 1ull) == 0ull) {
                        const int nmj = s_draw[off + jj].nm, nmk = s_draw[off + kk].nm;
                        const bool crow = ((s_wtc[jj * SPEC_MAX] | s_wtc[kk * SPEC_MAX]) >> 8) != 0;
                        float acc = 0.0f, comp = 0.0f;
                        for (int i = 0; i < nmk; ++i) {
                            const float4 o = s_wold[kk * SPEC_MAX + i], n = s_wnew[kk * SPEC_MAX + i];
                            const int ti = s_wtc[kk * SPEC_MAX + i] & 0xff;
                            for (int l = 0; l < nmj; ++l) {
                                const float4 O = s_wold[jj * SPEC_MAX + l], Nn = s_wnew[jj * SPEC_MAX + l];
                                float e_nn, e_on, e_no, e_oo;
                                if (crow) {
                                    e_nn = pair_crowder(dist2(Nn.x, Nn.y, Nn.z, n, S), ff);
                                    e_on = pair_crowder(dist2(Nn.x, Nn.y, Nn.z, o, S), ff);
                                    e_no = pair_crowder(dist2(O.x, O.y, O.z, n, S), ff);
                                    e_oo = pair_crowder(dist2(O.x, O.y, O.z, o, S), ff);
                                } else {
                                    const int sl = jj * SPEC_MAX + l;
                                    const float eij = s_etab[ti * REMC_NTYPES + (s_wtc[sl] & 0xff)];
                                    const float sij = 0.5f * (n.w + Nn.w);
                                    const float sij2 = sij * sij;
                                    const float qq = o.w * s_wrel[sl].w;
                                    e_nn = pair_lj_dh(dist2(Nn.x, Nn.y, Nn.z, n, S), eij, sij2, qq, ff);
                                    e_on = pair_lj_dh(dist2(Nn.x, Nn.y, Nn.z, o, S), eij, sij2, qq, ff);
                                    e_no = pair_lj_dh(dist2(O.x, O.y, O.z, n, S), eij, sij2, qq, ff);
                                    e_oo = pair_lj_dh(dist2(O.x, O.y, O.z, o, S), eij, sij2, qq, ff);
                                }
                                kahan_add(acc, comp, (e_nn - e_on) - (e_no - e_oo));
                            }
                        }
                        cv = acc - comp;
                        touches = cv != 0.0f;
                    }
                    if (touches) atomicOr(&s_wint[kk], 1u << jj);
                }
            }
            __syncthreads();

            // The window's run of small moves ends at a large move of this
            // batch: waves 1.. stage it while wave 0 gathers the partials and
            // decides.  It reads only the large move's body, which no slot of
            // the window moves, so the staging holds whatever the decisions.
            const int nxo = off + w;
            const bool nx_big = w < WK && nxo < have && it + w < nsteps && ((elig >> nxo) & 1ull) == 0ull;
            if (wave != 0 && nx_big) stage_move(tid - 64, BLOCK - 64, s_draw[nxo]);
            if (nx_big) pre_it = it + w;

            // The commit's data (everything but the decisions) is read into
            // registers by the upper half of the workgroup meanwhile: thread
            // BLOCK/2 + t commits slot t / SPEC_MAX, bead t % SPEC_MAX.
            const int ct = tid - BLOCK / 2;
            const bool cthr = ct >= 0 && ct < WK * SPEC_MAX;
            const int ck = cthr ? ct / SPEC_MAX : 0, ci = cthr ? ct % SPEC_MAX : 0;
            bool cin = false, crot = false;
            int cb = 0, cbk = 0;
            float4 cnew = make_float4(0.f, 0.f, 0.f, 0.f), crel = cnew, ccen = cnew;
            double cq[4] = {0.0, 0.0, 0.0, 0.0};
            if (cthr && ck < w) {
                const StepDraw &Dk = s_draw[off + ck];
                cbk = Dk.body;
                crot = Dk.rotate != 0;
                cin = ci < Dk.nm;
                cb = Dk.m0 + ci;
                if (cin) {
                    cnew = s_wnew[ct];
                    if (crot) crel = s_wrel[ct];
                    if (ci == 0) {
                        if (crot)
                            for (int t = 0; t < 4; ++t) cq[t] = s_wq[4 * ck + t];
                        else
                            ccen = s_wcen[ck];
                    }
                }
            }

            // decisions in step order (wave 0); lane l carries slot l's dE
            if (wave == 0) {
                // the slot's uniform and masks, read before the wait for the partials
                const double nll = lane < w ? s_draw[off + lane].nlu : 0.0;
                const unsigned int il = lane < w ? s_wint[lane] : 0u;
                const unsigned int chl = lane < w ? s_wchain[lane] : 0u;
                double dEl = 0.0;
                if (coop > 1) {
                    // lane k gathers slot k's coop partials (bank[k][m]) and adds
                    // them in member order: the same total in every member
                    // Every lane reads all REMC_MAX_COOP entries of its slot row
                    // (the row exists for every lane: REMC_SLOT_K >= 64) with no
                    // branch between the loads, so a poll costs one memory round
                    // trip, not one per member.
                    unsigned long long v[REMC_MAX_COOP];
                    int spins = 0;
                    bool failed = false;
                    for (;;) {
#pragma unroll
                        for (int m = 0; m < REMC_MAX_COOP; ++m)
                            v[m] = __hip_atomic_load(bank + lane * REMC_MAX_COOP + m, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                        bool mine = true;
#pragma unroll
                        for (int m = 0; m < REMC_MAX_COOP; ++m)
                            mine = mine && (m >= coop || (v[m] & SLOT_TAG_MASK) == wtag);
                        if (__all(mine || lane >= w || st_l)) break;
                        if (++spins >= SLOT_SPIN_LIMIT) {
                            failed = true;
                            break;
                        }
                        if ((spins & 255) == 0 && __hip_atomic_load(p.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                            failed = true;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    if (failed) {
                        if (lane == 0) {
                            __hip_atomic_fetch_or(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            s_abort = 1;
                        }
                    }
#pragma unroll
                    for (int m = 0; m < REMC_MAX_COOP; ++m)
                        if (m < coop) dEl += slot_value(v[m]);
                    if (lane >= w || st_l) dEl = 0.0;
                } else {
                    dEl = lane < w && !st_l ? s_wbase[lane] : 0.0;
                }
                // Every slot is decided at once on its dE0.  That is step k's
                // exact dE unless an accepted earlier slot j of the window moved a
                // body that interacts with slot k's (corr(j,k) != 0) or moved the
                // same body, or slot k is chained to a slot j (staged from j's
                // proposal) that is rejected.  The window ends before the first
                // such slot; the decisions before it are exact, and it starts the
                // next window, staged from the updated conformation.
                const int dl = lane < w ? remc_metropolis_l(dEl, beta, nll) : 0;
                const unsigned long long acc0 = __ballot(dl != 0);
                const bool inter = (il & (unsigned int)acc0) != 0u || (chl & ~(unsigned int)acc0) != 0u;
                const unsigned long long stop = __ballot(inter) | (w < 64 ? (1ull << w) : 0ull);
                const int weff = (int)__builtin_ctzll(stop);
                if (lane < weff) {
                    s_wdec[lane] = dl;
                    s_wdE[lane] = dEl;
                }
                const unsigned long long accm = weff < 64 ? acc0 & ((1ull << weff) - 1ull) : acc0;
                const unsigned long long boltzm = __ballot(dl == 2) & accm;
                if (lane == 0) {
                    // the tracked energy gains the accepted dE in step order, as one step at a time does
                    double Ew = s_E;
                    for (unsigned long long m = accm; m; m &= m - 1ull) {
                        const int k = (int)__builtin_ctzll(m);
                        Ew += __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(dEl), k),
                                               __builtin_amdgcn_readlane(__double2loint(dEl), k));
                    }
                    s_E = Ew;
                    const int nacc = __popcll(accm), nbz = __popcll(boltzm);
                    s_cnt[0] += (unsigned long long)(nacc - nbz);
                    s_cnt[1] += (unsigned long long)nbz;
                    s_cnt[2] += (unsigned long long)(weff - nacc);
                }
                // an accepted chained slot supersedes its (accepted) predecessor's
                // proposal in what the two moves both change
                if (lane < weff && dl != 0 && chl != 0u) s_wnext[__builtin_ctz(chl)] = lane;
                if (lane == 0) s_weff = weff;
            }
            __syncthreads();
            if (s_abort) break;
            if (coop > 1) ++n_x;
            const int weff = s_weff;

            // commit the accepted moves: copies of what staging formed.  Bodies
            // are distinct except for a chained pair, where the later slot writes
            // the positions and what its own move type changes (rotation: the
            // relative positions and the quaternion; translation: the centre),
            // and the earlier slot only what the later one leaves, so every
            // location has one writer.
            if (cthr && cin && ck < weff && s_wdec[ck]) {
                const int nx = s_wnext[ck];
                const bool later = nx >= 0;
                const bool later_rot = later && s_draw[off + nx].rotate;
                const int b = cb, bk = cbk;
                if (!later) {
                    gx[b] = cnew.x;
                    gy[b] = cnew.y;
                    gz[b] = cnew.z;
                    if (b >= clo && b < chi) {
                        s_jx[b - clo] = cnew.x;
                        s_jy[b - clo] = cnew.y;
                        s_jz[b - clo] = cnew.z;
                    }
                }
                if (crot && !later_rot) {
                    grx[b] = crel.x;
                    gry[b] = crel.y;
                    grz[b] = crel.z;
                    if (cache_rel) {
                        s_rx[b] = crel.x;
                        s_ry[b] = crel.y;
                        s_rz[b] = crel.z;
                    }
                }
                if (ci == 0) {
                    if (crot) {
                        if (!later_rot)
                            for (int t = 0; t < 4; ++t) {
                                gq[4 * bk + t] = cq[t];
                                if (cache_quat) s_q[4 * bk + t] = cq[t];
                            }
                    } else if (!later || later_rot) {
                        gcx[bk] = ccen.x;
                        gcy[bk] = ccen.y;
                        gcz[bk] = ccen.z;
                        if (cache_bodies) {
                            s_cx[bk] = ccen.x;
                            s_cy[bk] = ccen.y;
                            s_cz[bk] = ccen.z;
                        }
                    }
                }
            }
            if (c == 0 && p.rec && tid < weff) {
                const unsigned long long rec_start = p.rec_start[r];
                const unsigned long long st = step + tid;
                const unsigned long long idx = st - rec_start;
                if (st >= rec_start && idx < (unsigned long long)p.rec_max) {
                    const StepDraw &Dk = s_draw[off + tid];
                    remc_step_record rr;
                    rr.body = Dk.body;
                    rr.move = Dk.rotate ? 1 : 0;
                    rr.accepted = s_wdec[tid];
                    rr.ticks = (int32_t)((wall_clock64() - t_step) / (unsigned long long)weff);
                    rr.delta = s_wdE[tid];
                    p.rec[(size_t)r * p.rec_max + idx] = rr;
                }
            }
            __syncthreads();
            it += weff;
            step += weff;
            continue;
        }

        // ---------------------------- one step ----------------------------
        const StepDraw &D = s_draw[off];
        const bool pre_mv = SPEC && pre_it == it; // staged during the window before it
        pre_it = -1;
        const int body = D.body;
        const bool rotate = D.rotate != 0;
        const int m0 = D.m0, m1 = D.m0 + D.nm; // the bead range formed with the draw
        float cx, cy, cz;
        if (cache_bodies) {
            cx = lds_ld(s_cx + body);
            cy = lds_ld(s_cy + body);
            cz = lds_ld(s_cz + body);
        } else {
            cx = gcx[body];
            cy = gcy[body];
            cz = gcz[body];
        }
        const int nm = m1 - m0;
        const bool crow_m = (CACHE && m0 >= clo && m0 < chi) ? (s_jtc[m0 - clo] >> 8) != 0
                                                             : p.bcrow[S.body_off + body] != 0;
        float ncx = cx, ncy = cy, ncz = cz;
        if (!rotate) remc_wrap_centre(cx, cy, cz, D.t[0], D.t[1], D.t[2], S.box, &ncx, &ncy, &ncz);
        if (!pre_mv) stage_move(tid, BLOCK, D);
        __syncthreads();

        // Moves with few pair terms are evaluated whole by every member (no
        // exchange); larger ones are split over the members' partner slices.
        // (the body's bead range is uniform: the integer work below is scalar)
        const int m0u = __builtin_amdgcn_readfirstlane(m0), nmu = __builtin_amdgcn_readfirstlane(nm);
        const bool split = coop > 1 && (long long)nmu * (long long)(N - nmu) >= (long long)p.coop_min_pairs;
        // this member's share of dE = sum_{i in body} sum_{j in share} e(new_i,j) - e(old_i,j),
        // split into groups of moved beads x G lanes over the share's partner
        // beads: an even part of the N - nm partners (index v, bead v, or
        // v + nm past the moved body) when the member can read every bead,
        // else its cached slice without the moved body's overlap [ovs, ovs + ov)
        const bool even = split && !(CACHE && p.cache_slice);
        const int np = N - nmu;
        const int js = !split ? 0 : even ? share_bound(c, np, coop) : cjs;
        const int je = !split ? N : even ? share_bound(c + 1, np, coop) : cje;
        const int ovs = even ? m0u : max(js, m0u);
        const int ov = even ? nmu : max(0, min(je, m0u + nmu) - ovs);
        const int nj = even ? je - js : (je - js) - ov;
        int G, kper;
        pick_split_memo<BLOCK>(smemo, nmu, nj, lane, G, kper);
        const int g = div_floor_small(tid, G), l = tid - g * G;
        const int k0 = min(nm, g * kper), k1 = min(nm, k0 + kper);
        // compensated (Kahan) FP32 sum of the differences en - eo per thread
        float acc = 0.0f, comp = 0.0f;
        // two partner beads in registers per pass (jj0 and jj0 + G): each moved
        // bead's staging is read once for both
        for (int jj0 = l; jj0 < nj; jj0 += 2 * G) {
            float xj[2], yj[2], zj[2], qj[2], sj[2];
            int tcj[2];
            bool on[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int jj = jj0 + u * G;
                on[u] = jj < nj;
                const int ju = on[u] ? jj : jj0; // an index that exists, for the partner that is off
                const int j = js + ju + (js + ju >= ovs ? ov : 0);
                if (CACHE) {
                    const int kj = j - clo;
                    xj[u] = s_jx[kj];
                    yj[u] = s_jy[kj];
                    zj[u] = s_jz[kj];
                    qj[u] = s_jq[kj];
                    sj[u] = s_js[kj];
                    tcj[u] = s_jtc[kj];
                } else {
                    xj[u] = gx[j];
                    yj[u] = gy[j];
                    zj[u] = gz[j];
                    qj[u] = qs[j];
                    sj[u] = sgs[j];
                    tcj[u] = tcs[j];
                }
            }
            const bool cr0 = crow_m || (tcj[0] >> 8) != 0, cr1 = crow_m || (tcj[1] >> 8) != 0;
            if (__all(!cr0 && !(on[1] && cr1))) {
                // Every active lane's partners are non-crowder beads (a partner
                // that is off repeats partner 0 and is masked out), the common
                // case: the four terms of a moved bead (old and new position
                // against both partners) are formed side by side with no
                // per-lane branch, four independent chains per wave.  The
                // Debye-Hueckel part is skipped as a whole when no moved bead
                // of the wave is charged.  Distances are taken in units of the
                // box, where the minimum image is d - rint(d): the box-scaled
                // staging holds (x - c)/box about the moved body's old centre c
                // and q_i c_dh / box, the partners are scaled about c alike, and
                // the constants carry the box (sigma^2 / box^2, box k_dh), so
                // every term is the pair energy of the same formula.
                const float ib = S.inv_box, ib2 = ib * ib, kdh = ff.k_dh * S.box;
                float xs[2], ys[2], zs[2], hj[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    xs[u] = (xj[u] - cx) * ib;
                    ys[u] = (yj[u] - cy) * ib;
                    zs[u] = (zj[u] - cz) * ib;
                    hj[u] = 0.5f * sj[u];
                }
                const int t0 = tcj[0] & 0xff, t1 = tcj[1] & 0xff;
                // the pass over this lane's moved beads; with the type table
                // (TAB) sigma_ij^2 / box^2 comes with E_ij in one LDS read
                auto fast_pass = [&](auto tab) {
                    constexpr bool TAB = decltype(tab)::value;
                    for (int k = k0; k < k1; ++k) {
                        const float4 o = s_olds[k], n = s_news[k];
                        const float2 *erow = s_et2 + (s_tc[k] & 0xff) * REMC_NTYPES;
                        const float2 e0 = erow[t0], e1 = erow[t1];
                        float ss0, ss1;
                        if (TAB) {
                            ss0 = e0.y;
                            ss1 = e1.y;
                        } else {
                            const float h0 = n.w + hj[0], h1 = n.w + hj[1];
                            ss0 = (h0 * h0) * ib2;
                            ss1 = (h1 * h1) * ib2;
                        }
                        const float ro0 = dist2_unit(xs[0], ys[0], zs[0], o), rn0 = dist2_unit(xs[0], ys[0], zs[0], n);
                        const float ro1 = dist2_unit(xs[1], ys[1], zs[1], o), rn1 = dist2_unit(xs[1], ys[1], zs[1], n);
                        float eo0, en0, eo1, en1;
                        if (__any(o.w != 0.0f)) {
                            const float qq0 = o.w * qj[0], qq1 = o.w * qj[1];
                            eo0 = pair_lj_dh_k(ro0, e0.x, ss0, qq0, kdh);
                            en0 = pair_lj_dh_k(rn0, e0.x, ss0, qq0, kdh);
                            eo1 = pair_lj_dh_k(ro1, e1.x, ss1, qq1, kdh);
                            en1 = pair_lj_dh_k(rn1, e1.x, ss1, qq1, kdh);
                        } else {
                            eo0 = pair_lj(ro0, e0.x, ss0);
                            en0 = pair_lj(rn0, e0.x, ss0);
                            eo1 = pair_lj(ro1, e1.x, ss1);
                            en1 = pair_lj(rn1, e1.x, ss1);
                        }
                        // the two partners' differences are added (one rounding, as
                        // each difference has), then to the compensated sum
                        kahan_add(acc, comp, (en0 - eo0) + (on[1] ? en1 - eo1 : 0.0f));
                    }
                };
                if (p.sigtab)
                    fast_pass(std::true_type{});
                else
                    fast_pass(std::false_type{});
                continue;
            }
            if (__all(cr0 && (!on[1] || cr1))) {
                // every active lane's pairs are crowder pairs (the moved body or
                // both partners are crowders): the same four-chain form with
                // the crowder term only, in box units (cutoff^2 / box^2,
                // C sigma_c^6 / box^6)
                const float ib = S.inv_box, ib2 = ib * ib;
                const float cut2 = ff.crow_cut2 * ib2, ck = ff.crow_k * (ib2 * ib2 * ib2);
                float xs[2], ys[2], zs[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    xs[u] = (xj[u] - cx) * ib;
                    ys[u] = (yj[u] - cy) * ib;
                    zs[u] = (zj[u] - cz) * ib;
                }
                for (int k = k0; k < k1; ++k) {
                    const float4 o = s_olds[k], n = s_news[k];
                    const float eo0 = pair_crowder_u(dist2_unit(xs[0], ys[0], zs[0], o), cut2, ck);
                    const float en0 = pair_crowder_u(dist2_unit(xs[0], ys[0], zs[0], n), cut2, ck);
                    const float eo1 = pair_crowder_u(dist2_unit(xs[1], ys[1], zs[1], o), cut2, ck);
                    const float en1 = pair_crowder_u(dist2_unit(xs[1], ys[1], zs[1], n), cut2, ck);
                    kahan_add(acc, comp, (en0 - eo0) + (on[1] ? en1 - eo1 : 0.0f));
                }
                continue;
            }
            for (int k = k0; k < k1; ++k) {
                const float4 o = s_old[k], n = s_new[k];
                const int ti = s_tc[k] & 0xff;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (!on[u]) continue;
                    float eo, en;
                    if (u == 0 ? cr0 : cr1) {
                        eo = pair_crowder(dist2(xj[u], yj[u], zj[u], o, S), ff);
                        en = pair_crowder(dist2(xj[u], yj[u], zj[u], n, S), ff);
                    } else {
                        const float eij = s_etab[ti * REMC_NTYPES + (tcj[u] & 0xff)];
                        const float sij = 0.5f * (n.w + sj[u]);
                        const float sij2 = sij * sij;
                        const float qq = o.w * qj[u];
                        eo = pair_lj_dh(dist2(xj[u], yj[u], zj[u], o, S), eij, sij2, qq, ff);
                        en = pair_lj_dh(dist2(xj[u], yj[u], zj[u], n, S), eij, sij2, qq, ff);
                    }
                    kahan_add(acc, comp, en - eo);
                }
            }
        }
        const double v = wave_sum_dpp((double)acc - (double)comp);
        if (lane == 0) s_red[wave] = v;
        __syncthreads();
        double dE = 0.0;
        if (!split) {
#pragma unroll
            for (int w2 = 0; w2 < NW; ++w2) dE += s_red[w2];
        } else {
            // When the next step starts a window in this draw batch (no refill
            // due), waves 1.. stage it while wave 0 waits for the members'
            // partials: the staging reads only the window's bodies, which this
            // step's move does not change.
            // The refill rule of the next step is applied here as the step
            // applies it; a refill that switches to the batch formed ahead
            // leaves the window's draws known.
            int nw = 0, noff = off + 1;
            const StepDraw *nsd = s_draw;
            if (SPEC && it + 1 < nsteps) {
                bool refill1 = noff >= have;
                if (!refill1 && noff + WK > have && b0 + have < nsteps) {
                    const unsigned long long rest = (1ull << (have - noff)) - 1ull; // have - noff < WK <= 32
                    refill1 = ((~(elig >> noff)) & rest) == 0ull;
                }
                unsigned long long el1 = elig;
                if (refill1 && nb_ok && it + 1 >= nb0) {
                    noff = it + 1 - nb0;
                    el1 = s_elig[dbuf ^ 1];
                    nsd = dbuf ? s_draw0 : s_draw1;
                    refill1 = false;
                }
                if (!refill1) {
                    const unsigned long long run = ~(el1 >> noff);
                    nw = run == 0ull ? 64 : (int)__builtin_ctzll(run);
                    nw = __builtin_amdgcn_readfirstlane(min(nw, WK));
                }
            }
            if (wave != 0 && nw > 0 && tid - 64 < WK * SPEC_MAX) stage_window(tid - 64, noff, nw, nsd);
            if (wave == 0) {
                double part = 0.0;
#pragma unroll
                for (int w2 = 0; w2 < NW; ++w2) part += s_red[w2];
                // slots alternate with the split-step count: a member reuses a
                // slot only after every member has read it (they all wrote
                // their partial of the split step in between)
                const unsigned long long tag = step & 0x7FFFFull;
                unsigned long long *slot = slots + (n_x & 1u) * (REMC_SLOT_K * REMC_MAX_COOP);
                if (lane == 0) __hip_atomic_store(slot + c, slot_pack(part, tag), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                unsigned long long wv = 0ull;
                bool ok = lane >= coop;
                int spins = 0;
                bool failed = false;
                while (!__all(ok)) {
                    if (!ok) {
                        wv = __hip_atomic_load(slot + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = (wv & SLOT_TAG_MASK) == tag;
                    }
                    if (++spins >= SLOT_SPIN_LIMIT) {
                        failed = true;
                        break;
                    }
                    if ((spins & 255) == 0 && __hip_atomic_load(p.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        failed = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (failed) {
                    if (lane == 0) {
                        __hip_atomic_fetch_or(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        s_abort = 1;
                    }
                } else {
                    // members' partials summed in member order: identical in every member
                    double tot = 0.0;
                    for (int m = 0; m < coop; ++m) tot += slot_value(__shfl(wv, m, 64));
                    if (lane == 0) s_dE = tot;
                }
            }
            __syncthreads();
            if (s_abort) break;
            dE = s_dE;
            ++n_x;
            pre_w = nw;
        }

        const int decision = remc_metropolis_l(dE, beta, D.nlu);
        if (decision) {
            for (int k = tid; k < nm; k += BLOCK) {
                const int b = m0 + k;
                const float4 n = s_new[k];
                gx[b] = n.x;
                gy[b] = n.y;
                gz[b] = n.z;
                if (CACHE && b >= clo && b < chi) {
                    s_jx[b - clo] = n.x;
                    s_jy[b - clo] = n.y;
                    s_jz[b - clo] = n.z;
                }
                if (rotate) {
                    float rx, ry, rz;
                    if (cache_rel) {
                        rx = lds_ld(s_rx + b);
                        ry = lds_ld(s_ry + b);
                        rz = lds_ld(s_rz + b);
                    } else {
                        rx = grx[b];
                        ry = gry[b];
                        rz = grz[b];
                    }
                    float nrx, nry, nrz;
                    remc_qrotate(D.coef, rx, ry, rz, &nrx, &nry, &nrz);
                    grx[b] = nrx;
                    gry[b] = nry;
                    grz[b] = nrz;
                    if (cache_rel) {
                        s_rx[b] = nrx;
                        s_ry[b] = nry;
                        s_rz[b] = nrz;
                    }
                }
            }
            if (tid == 0) {
                if (rotate) {
                    double qo[4], qn[4];
                    if (cache_quat)
                        for (int t = 0; t < 4; ++t) qo[t] = lds_ld(s_q + 4 * body + t);
                    else
                        for (int t = 0; t < 4; ++t) qo[t] = gq[4 * body + t];
                    remc_qmul(D.q, qo, qn);
                    for (int t = 0; t < 4; ++t) {
                        gq[4 * body + t] = qn[t];
                        if (cache_quat) s_q[4 * body + t] = qn[t];
                    }
                } else {
                    gcx[body] = ncx;
                    gcy[body] = ncy;
                    gcz[body] = ncz;
                    if (cache_bodies) {
                        s_cx[body] = ncx;
                        s_cy[body] = ncy;
                        s_cz[body] = ncz;
                    }
                }
            }
        }
        if (tid == 0) {
            if (decision) s_E += dE;
            ++s_cnt[decision == 1 ? 0 : decision == 2 ? 1 : 2];
        }
        if (c == 0 && p.rec && tid == 0) {
            const unsigned long long rec_start = p.rec_start[r];
            unsigned long long idx = step - rec_start;
            if (step >= rec_start && idx < (unsigned long long)p.rec_max) {
                remc_step_record rr;
                rr.body = body;
                rr.move = rotate ? 1 : 0;
                rr.accepted = decision;
                rr.ticks = (int32_t)(wall_clock64() - t_step); // 100 MHz ticks spent on the step
                rr.delta = dE;
                p.rec[(size_t)r * p.rec_max + idx] = rr;
            }
        }
        __syncthreads();
        ++it;
        ++step;
    }
    if (c == 0 && tid == 0 && !s_abort) {
        p.E[r] = s_E;
        p.step[r] = step;
        p.cnt[3 * r + 0] += s_cnt[0];
        p.cnt[3 * r + 1] += s_cnt[1];
        p.cnt[3 * r + 2] += s_cnt[2];
    }
}

__global__ void k_reset_slots(unsigned long long *slots, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) slots[i] = ~0ull;
}

template <int BLOCK, bool CACHE, bool SPEC>
static hipError_t launch_mc_t(const KParams &p, int grid, int nsteps, int coop, hipStream_t s)
{
    const size_t lds = mc_dynamic_lds(p);
    hipError_t e = hipFuncSetAttribute((const void *)k_mc<BLOCK, CACHE, SPEC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_mc<BLOCK, CACHE, SPEC>), dim3(grid), dim3(BLOCK), lds, s, p, nsteps, coop);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// force-biased (smart) MC, Rossky, Doll & Friedman (1978), for rigid bodies
//
// coop workgroups (members) per replica, each with its own copy of the
// replica's state.  CACHE: the replica's beads, relative positions, centres
// and rotations in LDS (kept current on commit).  Wave 0 forms the draws of
// 64 steps at a time (one step per lane).  Each step: draws; force and torque
// on the chosen body at its old position (pass A; a large move split over the
// members' partner slices, the partials exchanged); the proposal
// d = beta A F + sqrt(2A) n for a translation, the rotation vector
// theta = beta A tau + sqrt(2A) n for a rotation (formed by every thread);
// staging of the new position; dE and the force or torque at the new position
// (pass B, exchanged likewise); the
// acceptance exponent x = beta dE + (|-d - beta A F'|^2 - |d - beta A F|^2)/(4A)
// (accept if x < -ln u); commit.  A = t^2/6 for translations and r^2/6 for
// rotations, from the replica's translate and rotate steps, so the random
// part of a move has the plain move's RMS size.  Pair energies in FP32 with
// Kahan sums of (new - old), forces and torques accumulated in FP64.  Runs of
// one-bead crowder moves are taken in windows (see the window code below).
// --------------------------------------------------------------------------

// Energy of one pair (pair_lj_dh / pair_crowder, the same operations) and
// g = (dE/dr)/r, so that the gradient of the pair energy with respect to the
// moved bead is g (x_i - x_j).
__device__ __forceinline__ void pair_eg(float r2, bool crow, float eij, float sij2, float qq, float inv_xi,
                                        const DevFF &ff, float &e, float &g)
{
    // The energy and its derivative share the reciprocal (root), x6 and the
    // Debye-Hueckel exponential.  For r2 >= 1e-36 the energy is formed by the
    // operations of pair_crowder / pair_lj_dh<false> on the same inputs, so it
    // is the same number; below (coincident beads) both give the 1e30 cap.
    const float r2c = fmaxf(r2, 1e-36f);
    if (crow) {
        const float rc = __builtin_amdgcn_rcpf(r2c);
        const float rc3 = rc * rc * rc;
        e = r2 < ff.crow_cut2 ? fminf(ff.crow_k * rc3, 1e30f) : 0.0f;
        g = fmaxf(-6.0f * e * rc, -1e30f); // d(C r^-6)/dr / r = -6 E / r^2
        return;
    }
    const float rinv = __builtin_amdgcn_rsqf(r2c);
    const float s2 = sij2 * (rinv * rinv);
    const float x6 = s2 * s2 * s2;
    const float base = __builtin_fmaf(x6, x6, -x6);
    const bool flip = (eij > 0.0f) & (x6 > 0.5f);
    const float r = r2c * rinv;
    const float dh = qq * __builtin_amdgcn_exp2f(r * ff.k_dh) * rinv;
    e = fminf(eij * (flip ? base + 0.5f : -base) + dh, 1e30f);
    const float dbase = (2.0f * x6 - 1.0f) * (-6.0f * x6 * rinv); // d[x6 (x6 - 1)]/dr (eij holds 4 E)
    const float dlj = flip ? eij * dbase : -eij * dbase;
    const float ddh = -dh * (inv_xi + rinv);
    g = fmaxf(fminf((dlj + ddh) * rinv, 1e30f), -1e30f);
}

template <int BLOCK, bool CACHE>
__global__ __launch_bounds__(BLOCK) void k_mc_fb(KParams p, int nsteps, int coop)
{
    constexpr int NW = BLOCK / 64;
    __shared__ float s_etab[REMC_NTYPES * REMC_NTYPES];
    __shared__ double s_red[7][NW];
    __shared__ double s_ft[7];
    __shared__ int s_abort;
    __shared__ double s_qr[4]; // the rotation of the step's proposal
    __shared__ float s_ncen[3]; // the centre of the step's proposal
    __shared__ remc_fb_draws s_fbd[64];
    __shared__ int s_fm0[64], s_fnm[64], s_fcrow[64]; // the drawn body's first bead, bead count, crowder flag
    // windows of one-bead crowder moves (slots k and k + NW on wave k, at most
    // FBW: four exchanged values per slot within REMC_SLOT_K)
    constexpr int FBW = 2 * NW < 16 ? 2 * NW : 16;
    static_assert(4 * FBW <= REMC_SLOT_K, "window exchange rows");
    __shared__ unsigned long long s_felig;               // steps of the batch a window may take
    __shared__ double s_wA[3 * FBW], s_wB[4 * FBW];      // force at the old position; dE and force at the new
    __shared__ float4 s_wold[FBW], s_wnew[FBW], s_wnrel[FBW], s_wncen[FBW];
    __shared__ double s_wdisp[3 * FBW], s_wfwd[FBW], s_wqp[4 * FBW], s_wdE[FBW];
    __shared__ unsigned int s_wint[FBW];
    __shared__ int s_wdec[FBW], s_weff;
    __shared__ int s_wstat[FBW]; // 1: the slot leaves its bead where it is (the rotation of a sphere)
    extern __shared__ float4 smem[];
    float4 *s_old = smem;                // old position, q_i c_dh
    float4 *s_orel = s_old + p.lds_mb;   // old relative position, sigma
    float4 *s_new = s_orel + p.lds_mb;   // new position
    float4 *s_nrel = s_new + p.lds_mb;   // new relative position
    int *s_tc = (int *)(s_nrel + p.lds_mb);
    // CACHE: the replica's partner beads with their relative positions
    // (p.lds_cl of them) and its body centres and rotations (p.lds_nbody) in
    // LDS; else all of them are read from HBM
    constexpr bool cached = CACHE, cache_rel = CACHE, cache_bodies = CACHE;
    const int cl = CACHE ? p.lds_cl : 0, nb_c = CACHE ? p.lds_nbody : 0;
    float *s_jx = (float *)(s_tc + p.lds_mb), *s_jy = s_jx + cl, *s_jz = s_jy + cl;
    float *s_jq = s_jz + cl, *s_js = s_jq + cl;
    int *s_jtc = (int *)(s_js + cl);
    float *s_rx = (float *)(s_jtc + cl), *s_ry = s_rx + cl, *s_rz = s_ry + cl;
    char *const bq_at = (char *)(s_rz + cl);
    double *s_bq = (double *)(bq_at + ((8 - ((bq_at - (char *)smem) & 7)) & 7)); // 8-byte aligned
    float *s_bcx = (float *)(s_bq + 4 * nb_c), *s_bcy = s_bcx + nb_c, *s_bcz = s_bcy + nb_c;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = blockIdx.x % p.n_rep;
    const int c = blockIdx.x / p.n_rep; // member c of the replica's coop workgroups
    const DevRep R = p.rep[r];
    const DevSys S = p.sys[R.sys];
    const DevFF ff = p.ff;
    for (int k = tid; k < REMC_NTYPES * REMC_NTYPES; k += BLOCK) s_etab[k] = p.etab[k];
    if (tid == 0) s_abort = 0;
    const double beta = remc_beta(p.T[r]);
    const double a_t = (double)R.tstep * (double)R.tstep / 6.0, a_r = R.rstep * R.rstep / 6.0;
    const float inv_xi = -ff.k_dh * 0.69314718f; // 1/xi from k_dh = -log2(e)/xi
    unsigned long long step = p.step[r];
    double E = p.E[r];
    unsigned long long n_acc = 0, n_accb = 0, n_rej = 0;
    // member c's private copy of the replica's dynamic state; every member
    // commits every move to its own copy
    const size_t cb = (size_t)c * p.copy_beads + R.bead_off;
    const size_t cmo = (size_t)c * p.copy_bodies + R.body_off;
    float *gx = p.px + cb, *gy = p.py + cb, *gz = p.pz + cb;
    float *grx = p.rx + cb, *gry = p.ry + cb, *grz = p.rz + cb;
    float *gcx = p.cx + cmo, *gcy = p.cy + cmo, *gcz = p.cz + cmo;
    double *gq = p.quat + 4 * cmo;
    const int *tcs = p.tc + S.bead_off;
    const float *qs = p.q + S.bead_off, *sgs = p.sig + S.bead_off;
    const int N = S.n_beads;
    const int cjs = (int)(((long long)c * N) / coop), cje = (int)(((long long)(c + 1) * N) / coop);
    unsigned long long *slots = p.slots + (size_t)r * REMC_SLOTS_PER_REP;
    unsigned int n_x = 0; // exchanges so far in this launch: selects the slot bank
    if (cached)
        for (int b = tid; b < N; b += BLOCK) {
            s_jx[b] = gx[b];
            s_jy[b] = gy[b];
            s_jz[b] = gz[b];
            s_jq[b] = qs[b];
            s_js[b] = sgs[b];
            s_jtc[b] = tcs[b];
            if (cache_rel) {
                s_rx[b] = grx[b];
                s_ry[b] = gry[b];
                s_rz[b] = grz[b];
            }
        }
    if (cache_bodies) {
        for (int m = tid; m < S.n_bodies; m += BLOCK) {
            s_bcx[m] = gcx[m];
            s_bcy[m] = gcy[m];
            s_bcz[m] = gcz[m];
        }
        for (int m = tid; m < 4 * S.n_bodies; m += BLOCK) s_bq[m] = gq[m];
    }
    __syncthreads();

    // block sum of v[0..n) into s_ft (FP64, fixed order)
    auto block_sum = [&](double *v, int n) {
        for (int k = 0; k < n; ++k) {
            const double t = wave_sum_dpp(v[k]);
            if (lane == 0) s_red[k][wave] = t;
        }
        __syncthreads();
        if (tid < n) {
            double t = 0.0;
            for (int w = 0; w < NW; ++w) t += s_red[tid][w];
            s_ft[tid] = t;
        }
        __syncthreads();
    };
    // vals[0..n) (this member's partials, n <= REMC_SLOT_K) become the
    // members' totals, added in member order so that every member holds the
    // same values: one tagged 64-bit slot per value and member, the slot bank
    // alternating with the exchange count (a member reuses a bank only after
    // every member has read it), bounded waits; lane k of wave 0 gathers value k
    auto exchange = [&](double *vals, int n, unsigned long long tag) {
        unsigned long long *bank = slots + (n_x & 1u) * (REMC_SLOT_K * REMC_MAX_COOP);
        if (wave == 0) {
            if (lane < n)
                __hip_atomic_store(bank + lane * REMC_MAX_COOP + c, slot_pack(vals[lane], tag), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            // every lane reads its whole row (REMC_SLOT_K >= 64 rows) with no
            // branch between the loads: one memory round trip per poll
            unsigned long long v[REMC_MAX_COOP];
            int spins = 0;
            bool failed = false;
            for (;;) {
#pragma unroll
                for (int m = 0; m < REMC_MAX_COOP; ++m)
                    v[m] = __hip_atomic_load(bank + lane * REMC_MAX_COOP + m, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
                bool mine = true;
#pragma unroll
                for (int m = 0; m < REMC_MAX_COOP; ++m)
                    mine = mine && (m >= coop || (v[m] & SLOT_TAG_MASK) == tag);
                if (__all(mine || lane >= n)) break;
                if (++spins >= SLOT_SPIN_LIMIT) {
                    failed = true;
                    break;
                }
                if ((spins & 255) == 0 && __hip_atomic_load(p.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    failed = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (failed) {
                if (lane == 0) {
                    __hip_atomic_fetch_or(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_abort = 1;
                }
            }
            if (lane < n) {
                double t = 0.0;
#pragma unroll
                for (int m = 0; m < REMC_MAX_COOP; ++m)
                    if (m < coop) t += slot_value(v[m]);
                vals[lane] = t;
            }
        }
        __syncthreads();
        ++n_x;
    };

    int fb0 = 0, fhave = 0; // s_fbd holds the draws of steps [fb0, fb0 + fhave) of this launch
    SplitMemo smemo;        // the split of the latest large-move shapes
    for (int it = 0; it < nsteps; ++it, ++step) {
        // the thread index re-read opaquely in every step, as in k_mc: the
        // addresses derived from it are formed where used, not held in registers
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & 63, wave = tid >> 6;
        const unsigned long long t_step = p.rec ? wall_clock64() : 0ull;
        int off = it - fb0;
        bool refill = off >= fhave;
        if (CACHE && p.spec && !refill && off + FBW > fhave && fb0 + fhave < nsteps) {
            // fewer than FBW draws left: form the next batch now if the run of
            // window-eligible steps from here reaches the batch's end
            const unsigned long long rest = (1ull << (fhave - off)) - 1ull; // fhave - off < FBW <= 16
            refill = ((~(s_felig >> off)) & rest) == 0ull;
        }
        if (refill) {
            // wave 0 forms the draws of the next 64 steps, one per lane, with the drawn body's bead range
            if (wave == 0 && it + lane < nsteps) {
                const remc_fb_draws dd = remc_fb_step_draws(step + lane, R.rng_id, p.key0, p.key1, S.n_bodies);
                s_fbd[lane] = dd;
                const int gb = S.body_off + dd.body;
                s_fm0[lane] = p.bstart[gb];
                s_fnm[lane] = p.bend[gb] - p.bstart[gb];
                s_fcrow[lane] = p.bcrow[gb];
            }
            if (wave == 0) {
                const bool el = it + lane < nsteps && s_fnm[lane] == 1 && s_fcrow[lane] != 0;
                const unsigned long long m = __ballot(el);
                if (lane == 0) s_felig = m;
            }
            fb0 = it;
            fhave = min(64, nsteps - it);
            off = 0;
            __syncthreads();
        }
        // a window of w >= 2 one-bead crowder moves, within the batch
        int w = 0;
        if (CACHE && p.spec) {
            const unsigned long long run = ~(s_felig >> off);
            w = min(run == 0ull ? 64 : (int)__builtin_ctzll(run), FBW);
        }
        if (CACHE && w >= 2) {
            // ----- window of w one-bead crowder moves, slots k and k + NW on wave k -----
            // Each slot's force, dE and new force are taken against the
            // conformation at the window's start; a crowder's pair terms vanish
            // beyond the cutoff, so they are exact unless an accepted earlier
            // slot of the window moved within the cutoff of the slot's bead (old
            // or new position, either end) or moved the same body: the window
            // ends before the first such slot, which starts the next one.
            const int js = coop > 1 ? cjs : 0, je = coop > 1 ? cje : N;
            const int bw = off;
            if (wave < w) {
                // slots ka and kb (when it exists) in one pass over the partners
                const int ka = wave, kb = wave + NW;
                const bool two = kb < w;
                const int ba = s_fm0[bw + ka], bb = two ? s_fm0[bw + kb] : ba;
                const float ax = s_jx[ba], ay = s_jy[ba], az = s_jz[ba];
                const float bx = s_jx[bb], by = s_jy[bb], bz = s_jz[bb];
                // partners beyond the cutoff (+0.01 A for rounding) give exactly
                // zero force: a wave with none nearer skips the derivative
                const float limA = (ff.crow_cut + 0.01f) * (ff.crow_cut + 0.01f);
                double fa0 = 0.0, fa1 = 0.0, fa2 = 0.0, fb0 = 0.0, fb1 = 0.0, fb2 = 0.0;
                for (int j = js + lane; j < je; j += 64) {
                    const float xj = s_jx[j], yj = s_jy[j], zj = s_jz[j];
                    {
                        const float4 pa = make_float4(ax, ay, az, 0.0f), pb = make_float4(bx, by, bz, 0.0f);
                        const float r2a = dist2(xj, yj, zj, pa, S), r2b = dist2(xj, yj, zj, pb, S);
                        if (!__any((j != ba && r2a < limA) || (two && j != bb && r2b < limA))) continue;
                    }
                    if (j != ba) {
                        const float dx = min_image(ax - xj, S.box, S.inv_box);
                        const float dy = min_image(ay - yj, S.box, S.inv_box);
                        const float dz = min_image(az - zj, S.box, S.inv_box);
                        float e, g2;
                        pair_eg(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), true, 0.0f, 0.0f, 0.0f, inv_xi,
                                ff, e, g2);
                        fa0 += -(double)(g2 * dx);
                        fa1 += -(double)(g2 * dy);
                        fa2 += -(double)(g2 * dz);
                    }
                    if (two && j != bb) {
                        const float dx = min_image(bx - xj, S.box, S.inv_box);
                        const float dy = min_image(by - yj, S.box, S.inv_box);
                        const float dz = min_image(bz - zj, S.box, S.inv_box);
                        float e, g2;
                        pair_eg(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), true, 0.0f, 0.0f, 0.0f, inv_xi,
                                ff, e, g2);
                        fb0 += -(double)(g2 * dx);
                        fb1 += -(double)(g2 * dy);
                        fb2 += -(double)(g2 * dz);
                    }
                }
                fa0 = wave_sum_dpp(fa0);
                fa1 = wave_sum_dpp(fa1);
                fa2 = wave_sum_dpp(fa2);
                if (lane == 0) {
                    s_wA[3 * ka] = fa0;
                    s_wA[3 * ka + 1] = fa1;
                    s_wA[3 * ka + 2] = fa2;
                }
                if (two) {
                    fb0 = wave_sum_dpp(fb0);
                    fb1 = wave_sum_dpp(fb1);
                    fb2 = wave_sum_dpp(fb2);
                    if (lane == 0) {
                        s_wA[3 * kb] = fb0;
                        s_wA[3 * kb + 1] = fb1;
                        s_wA[3 * kb + 2] = fb2;
                    }
                }
            }
            __syncthreads();
            if (coop > 1) {
                exchange(s_wA, 3 * w, (2ull * step) & 0x7FFFFull);
                if (s_abort) break;
            }
            // proposals (every lane of slot k's wave forms the same values)
            for (int k = wave; k < w; k += NW) {
                const int b = s_fm0[bw + k];
                const remc_fb_draws dk = s_fbd[bw + k];
                const float rx = s_rx[b], ry = s_ry[b], rz = s_rz[b];
                const float cx = s_bcx[dk.body], cy = s_bcy[dk.body], cz = s_bcz[dk.body];
                const double F[3] = {s_wA[3 * k], s_wA[3 * k + 1], s_wA[3 * k + 2]};
                const double tq[3] = {(double)ry * F[2] - (double)rz * F[1], (double)rz * F[0] - (double)rx * F[2],
                                      (double)rx * F[1] - (double)ry * F[0]};
                const double A = dk.rotate ? a_r : a_t;
                const double *fo = dk.rotate ? tq : F;
                double disp[3], qr[4] = {1.0, 0.0, 0.0, 0.0}, cf[9];
                for (int t = 0; t < 3; ++t) disp[t] = beta * A * fo[t] + sqrt(2.0 * A) * dk.n[t];
                float nrx = rx, nry = ry, nrz = rz, ncx = cx, ncy = cy, ncz = cz;
                if (dk.rotate) {
                    const double ang = sqrt(disp[0] * disp[0] + disp[1] * disp[1] + disp[2] * disp[2]);
                    if (ang > 0.0) {
                        double sh, ch;
                        remc_sincos_turns(0.5 * ang / 6.283185307179586476925, &sh, &ch);
                        qr[0] = ch; qr[1] = sh * disp[0] / ang; qr[2] = sh * disp[1] / ang; qr[3] = sh * disp[2] / ang;
                        const double qm = sqrt(qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
                        for (int t = 0; t < 4; ++t) qr[t] = qr[t] / qm;
                    }
                    remc_qcoeffs(qr, cf);
                    remc_qrotate(cf, rx, ry, rz, &nrx, &nry, &nrz);
                } else {
                    const float t0 = (float)disp[0], t1 = (float)disp[1], t2 = (float)disp[2];
                    disp[0] = t0; disp[1] = t1; disp[2] = t2; // the displacement actually made
                    remc_wrap_centre(cx, cy, cz, t0, t1, t2, S.box, &ncx, &ncy, &ncz);
                }
                double fwd = 0.0;
                for (int t = 0; t < 3; ++t) {
                    const double a = disp[t] - beta * A * fo[t];
                    fwd += a * a;
                }
                if (lane == 0) {
                    s_wold[k] = make_float4(s_jx[b], s_jy[b], s_jz[b], 0.0f);
                    s_wnew[k] = make_float4(nrx + ncx, nry + ncy, nrz + ncz, 0.0f);
                    s_wstat[k] = s_jx[b] == nrx + ncx && s_jy[b] == nry + ncy && s_jz[b] == nrz + ncz ? 1 : 0;
                    s_wnrel[k] = make_float4(nrx, nry, nrz, 0.0f);
                    s_wncen[k] = make_float4(ncx, ncy, ncz, 0.0f);
                    for (int t = 0; t < 3; ++t) s_wdisp[3 * k + t] = disp[t];
                    for (int t = 0; t < 4; ++t) s_wqp[4 * k + t] = qr[t];
                    s_wfwd[k] = fwd;
                }
            }
            __syncthreads();
            for (int k = wave; k < w; k += NW) {
                // which earlier slots reach slot k (lane j looks at slot j)
                bool touch = false;
                if (lane < k) {
                    if (s_fbd[bw + lane].body == s_fbd[bw + k].body) {
                        touch = true;
                    } else if (s_wstat[lane] == 0 && s_wstat[k] == 0) {
                        // a static slot's bead stays put: its dE is 0 and its
                        // torque 0 whatever moves around it, and it moves nothing
                        const float4 oj = s_wold[lane], nj_ = s_wnew[lane], ok = s_wold[k], nk = s_wnew[k];
                        touch = dist2(oj.x, oj.y, oj.z, ok, S) < ff.crow_cut2 || dist2(oj.x, oj.y, oj.z, nk, S) < ff.crow_cut2 ||
                                dist2(nj_.x, nj_.y, nj_.z, ok, S) < ff.crow_cut2 ||
                                dist2(nj_.x, nj_.y, nj_.z, nk, S) < ff.crow_cut2;
                    }
                }
                const unsigned long long tm = __ballot(touch);
                if (lane == 0) s_wint[k] = (unsigned int)tm;
            }
            // static slots: dE = 0 and the new force is the old one (the bead
            // is where it was); their entries are zero partials, and the
            // decision takes the old force.  The other slots are dealt to the
            // waves in order, two per wave.
            const bool wk_l = lane < w && s_wstat[lane] == 0;
            const int rank_l = __popcll(__ballot(wk_l) & ((1ull << lane) - 1ull));
            const int wv = __builtin_amdgcn_readfirstlane(wave);
            const unsigned long long mka = __ballot(wk_l && rank_l == wv), mkb = __ballot(wk_l && rank_l == wv + NW);
            if (tid < w && s_wstat[tid] != 0)
                for (int t = 0; t < 4; ++t) s_wB[4 * tid + t] = 0.0;
            if (mka != 0ull) {
                // dE and the force at the new position, slots ka and kb in one pass
                const int ka = (int)__builtin_ctzll(mka), kb = mkb != 0ull ? (int)__builtin_ctzll(mkb) : ka;
                const bool two = mkb != 0ull;
                const int ba = s_fm0[bw + ka], bb = two ? s_fm0[bw + kb] : ba;
                const float4 oa = s_wold[ka], na = s_wnew[ka], ob = s_wold[two ? kb : ka], nb = s_wnew[two ? kb : ka];
                const float limBa = crowder_screen2(oa, na, S, ff), limBb = crowder_screen2(ob, nb, S, ff);
                float acc_a = 0.0f, comp_a = 0.0f, acc_b = 0.0f, comp_b = 0.0f;
                double fa0 = 0.0, fa1 = 0.0, fa2 = 0.0, fb0 = 0.0, fb1 = 0.0, fb2 = 0.0;
                for (int j = js + lane; j < je; j += 64) {
                    const float xj = s_jx[j], yj = s_jy[j], zj = s_jz[j];
                    {
                        const float r2a = dist2(xj, yj, zj, oa, S), r2b = dist2(xj, yj, zj, ob, S);
                        if (!__any((j != ba && r2a < limBa) || (two && j != bb && r2b < limBb))) continue;
                    }
                    if (j != ba) {
                        float eo, en, go, g2;
                        {
                            const float dx = min_image(oa.x - xj, S.box, S.inv_box);
                            const float dy = min_image(oa.y - yj, S.box, S.inv_box);
                            const float dz = min_image(oa.z - zj, S.box, S.inv_box);
                            pair_eg(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), true, 0.0f, 0.0f, 0.0f, inv_xi,
                                    ff, eo, go);
                        }
                        const float dx = min_image(na.x - xj, S.box, S.inv_box);
                        const float dy = min_image(na.y - yj, S.box, S.inv_box);
                        const float dz = min_image(na.z - zj, S.box, S.inv_box);
                        pair_eg(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), true, 0.0f, 0.0f, 0.0f, inv_xi,
                                ff, en, g2);
                        kahan_add(acc_a, comp_a, en - eo);
                        fa0 += -(double)(g2 * dx);
                        fa1 += -(double)(g2 * dy);
                        fa2 += -(double)(g2 * dz);
                    }
                    if (two && j != bb) {
                        float eo, en, go, g2;
                        {
                            const float dx = min_image(ob.x - xj, S.box, S.inv_box);
                            const float dy = min_image(ob.y - yj, S.box, S.inv_box);
                            const float dz = min_image(ob.z - zj, S.box, S.inv_box);
                            pair_eg(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), true, 0.0f, 0.0f, 0.0f, inv_xi,
                                    ff, eo, go);
                        }
                        const float dx = min_image(nb.x - xj, S.box, S.inv_box);
                        const float dy = min_image(nb.y - yj, S.box, S.inv_box);
                        const float dz = min_image(nb.z - zj, S.box, S.inv_box);
                        pair_eg(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), true, 0.0f, 0.0f, 0.0f, inv_xi,
                                ff, en, g2);
                        kahan_add(acc_b, comp_b, en - eo);
                        fb0 += -(double)(g2 * dx);
                        fb1 += -(double)(g2 * dy);
                        fb2 += -(double)(g2 * dz);
                    }
                }
                const double dEa = wave_sum_dpp((double)acc_a - (double)comp_a);
                fa0 = wave_sum_dpp(fa0);
                fa1 = wave_sum_dpp(fa1);
                fa2 = wave_sum_dpp(fa2);
                if (lane == 0) {
                    s_wB[4 * ka] = dEa;
                    s_wB[4 * ka + 1] = fa0;
                    s_wB[4 * ka + 2] = fa1;
                    s_wB[4 * ka + 3] = fa2;
                }
                if (two) {
                    const double dEb = wave_sum_dpp((double)acc_b - (double)comp_b);
                    fb0 = wave_sum_dpp(fb0);
                    fb1 = wave_sum_dpp(fb1);
                    fb2 = wave_sum_dpp(fb2);
                    if (lane == 0) {
                        s_wB[4 * kb] = dEb;
                        s_wB[4 * kb + 1] = fb0;
                        s_wB[4 * kb + 2] = fb1;
                        s_wB[4 * kb + 3] = fb2;
                    }
                }
            }
            __syncthreads();
            if (coop > 1) {
                exchange(s_wB, 4 * w, (2ull * step + 1ull) & 0x7FFFFull);
                if (s_abort) break;
            }
            // decisions (wave 0, lane k = slot k), as one step at a time takes them
            if (wave == 0) {
                int dl = 0;
                double dEk = 0.0;
                if (lane < w) {
                    const int k = lane;
                    const remc_fb_draws dk = s_fbd[bw + k];
                    const float4 nr = s_wnrel[k];
                    const bool stk = s_wstat[k] != 0;
                    const double Fn[3] = {stk ? s_wA[3 * k] : s_wB[4 * k + 1], stk ? s_wA[3 * k + 1] : s_wB[4 * k + 2],
                                          stk ? s_wA[3 * k + 2] : s_wB[4 * k + 3]};
                    const double tn[3] = {(double)nr.y * Fn[2] - (double)nr.z * Fn[1],
                                          (double)nr.z * Fn[0] - (double)nr.x * Fn[2],
                                          (double)nr.x * Fn[1] - (double)nr.y * Fn[0]};
                    const double *fn = dk.rotate ? tn : Fn;
                    const double A = dk.rotate ? a_r : a_t;
                    double rev = 0.0;
                    for (int t = 0; t < 3; ++t) {
                        const double bb = -s_wdisp[3 * k + t] - beta * A * fn[t];
                        rev += bb * bb;
                    }
                    dEk = s_wB[4 * k];
                    const double x = beta * dEk + (rev - s_wfwd[k]) / (4.0 * A);
                    dl = remc_metropolis_l(x, 1.0, dk.nlu);
                }
                const unsigned long long acc0 = __ballot(dl != 0);
                const unsigned int il = lane < w ? s_wint[lane] : 0u;
                const unsigned long long stop = __ballot((il & (unsigned int)acc0) != 0u) | (1ull << w);
                const int weff = (int)__builtin_ctzll(stop);
                if (lane < weff) {
                    s_wdec[lane] = dl;
                    s_wdE[lane] = dEk;
                }
                if (lane == 0) s_weff = weff;
            }
            __syncthreads();
            const int weff = s_weff;
            // commit the accepted slots (distinct bodies: a repeat ends the window)
            if (tid < weff && s_wdec[tid]) {
                const int k = tid, b = s_fm0[bw + k];
                const remc_fb_draws dk = s_fbd[bw + k];
                const float4 n = s_wnew[k];
                gx[b] = n.x;
                gy[b] = n.y;
                gz[b] = n.z;
                s_jx[b] = n.x;
                s_jy[b] = n.y;
                s_jz[b] = n.z;
                if (dk.rotate) {
                    const float4 nr = s_wnrel[k];
                    grx[b] = nr.x;
                    gry[b] = nr.y;
                    grz[b] = nr.z;
                    s_rx[b] = nr.x;
                    s_ry[b] = nr.y;
                    s_rz[b] = nr.z;
                    double qp[4], qo[4], qn[4];
                    for (int t = 0; t < 4; ++t) {
                        qp[t] = s_wqp[4 * k + t];
                        qo[t] = s_bq[4 * dk.body + t];
                    }
                    remc_qmul(qp, qo, qn);
                    for (int t = 0; t < 4; ++t) {
                        gq[4 * dk.body + t] = qn[t];
                        s_bq[4 * dk.body + t] = qn[t];
                    }
                } else {
                    const float4 nc = s_wncen[k];
                    gcx[dk.body] = nc.x;
                    gcy[dk.body] = nc.y;
                    gcz[dk.body] = nc.z;
                    s_bcx[dk.body] = nc.x;
                    s_bcy[dk.body] = nc.y;
                    s_bcz[dk.body] = nc.z;
                }
            }
            if (c == 0 && p.rec && tid < weff) {
                const unsigned long long st = step + tid;
                const unsigned long long rec_start = p.rec_start[r];
                const unsigned long long idx = st - rec_start;
                if (st >= rec_start && idx < (unsigned long long)p.rec_max) {
                    remc_step_record rr;
                    rr.body = s_fbd[bw + tid].body;
                    rr.move = s_fbd[bw + tid].rotate ? 1 : 0;
                    rr.accepted = s_wdec[tid];
                    rr.ticks = (int32_t)((wall_clock64() - t_step) / (unsigned long long)weff);
                    rr.delta = s_wdE[tid];
                    p.rec[(size_t)r * p.rec_max + idx] = rr;
                }
            }
            // the tracked energy and the counters, in step order, in every thread
            for (int k = 0; k < weff; ++k) {
                const int dk = s_wdec[k];
                if (dk) {
                    E += s_wdE[k];
                    if (dk == 1) ++n_acc; else ++n_accb;
                } else {
                    ++n_rej;
                }
            }
            __syncthreads();
            it += weff - 1;
            step += weff - 1;
            continue;
        }
        const remc_fb_draws d = s_fbd[off];
        const int body = d.body;
        const int m0 = s_fm0[off], nm = s_fnm[off], m1 = m0 + nm;
        const bool crow_m = s_fcrow[off] != 0;
        // the body's centre, needed by the proposal after pass A
        const float cx0 = cache_bodies ? s_bcx[body] : gcx[body], cy0 = cache_bodies ? s_bcy[body] : gcy[body],
                    cz0 = cache_bodies ? s_bcz[body] : gcz[body];
        for (int k = tid; k < nm; k += BLOCK) {
            const int b = m0 + k;
            s_old[k] = cached ? make_float4(s_jx[b], s_jy[b], s_jz[b], s_jq[b] * ff.c_dh)
                              : make_float4(gx[b], gy[b], gz[b], qs[b] * ff.c_dh);
            s_orel[k] = cache_rel ? make_float4(s_rx[b], s_ry[b], s_rz[b], s_js[b])
                                  : make_float4(grx[b], gry[b], grz[b], sgs[b]);
            s_tc[k] = cached ? s_jtc[b] : tcs[b];
        }
        __syncthreads();

        // A large move is split over the members' shares of the partners, the
        // partial forces, torques and dE then exchanged; a small one is
        // evaluated whole by every member, with the same result in each.
        const int nmu = __builtin_amdgcn_readfirstlane(nm); // uniform: scalar integer work
        const bool split = coop > 1 && (long long)nmu * (long long)(N - nmu) >= (long long)p.coop_min_pairs;
        // a member's share: an even part of the N - nm partners (index v, bead
        // v, or v + nm past the moved body)
        const int np = N - nmu;
        const int js = split ? share_bound(c, np, coop) : 0;
        const int je = split ? share_bound(c + 1, np, coop) : np;
        const int ovs = m0, ov = nm;
        const int nj = je - js;

        // pass A: force and torque at the old position; the (moved bead x
        // partner bead) pairs split over the workgroup as in the MC kernel,
        // two partner beads per lane per pass (jj0 and jj0 + G)
        int G, kper;
        pick_split_memo<BLOCK>(smemo, nmu, nj, lane, G, kper);
        const int g = div_floor_small(tid, G), l = tid - g * G;
        const int k0 = min(nm, g * kper), k1 = min(nm, k0 + kper);
        // partner jj's coordinates, charge, sigma and type (jj < nj)
        auto load_partner = [&](const int jj, float &xj, float &yj, float &zj, float &qj, float &sj, int &tcj) {
            const int j = js + jj + (js + jj >= ovs ? ov : 0);
            xj = cached ? s_jx[j] : gx[j];
            yj = cached ? s_jy[j] : gy[j];
            zj = cached ? s_jz[j] : gz[j];
            qj = cached ? s_jq[j] : qs[j];
            sj = cached ? s_js[j] : sgs[j];
            tcj = cached ? s_jtc[j] : tcs[j];
        };
        // Each lane sums its pairs' force and torque in FP32 (a few to a few
        // hundred terms: about 1e-6 relative; the forces only shape the
        // proposal, whose ratio uses the same computed forces, and dE keeps
        // its compensated sum) and the lanes' sums are added in FP64.
        double v[7] = {0, 0, 0, 0, 0, 0, 0};
        float f6[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int jj0 = l; jj0 < nj; jj0 += 2 * G) {
            const bool on1 = jj0 + G < nj;
            float xj[2], yj[2], zj[2], qj[2], sj[2];
            int tcj[2];
            bool crow[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                load_partner(u == 0 || !on1 ? jj0 : jj0 + G, xj[u], yj[u], zj[u], qj[u], sj[u], tcj[u]);
                crow[u] = crow_m || (tcj[u] >> 8) != 0;
            }
            for (int k = k0; k < k1; ++k) {
                const float4 o = s_old[k], orl = s_orel[k];
                const int trow = (s_tc[k] & 0xff) * REMC_NTYPES;
                float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const float dx = min_image(o.x - xj[u], S.box, S.inv_box);
                    const float dy = min_image(o.y - yj[u], S.box, S.inv_box);
                    const float dz = min_image(o.z - zj[u], S.box, S.inv_box);
                    const float sij = 0.5f * (orl.w + sj[u]);
                    float e, g2;
                    pair_eg(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), crow[u],
                            s_etab[trow + (tcj[u] & 0xff)], sij * sij, o.w * qj[u], inv_xi, ff, e, g2);
                    if (u == 1 && !on1) g2 = 0.0f;
                    fx -= g2 * dx;
                    fy -= g2 * dy;
                    fz -= g2 * dz;
                }
                f6[0] += fx;
                f6[1] += fy;
                f6[2] += fz;
                f6[3] += orl.y * fz - orl.z * fy;
                f6[4] += orl.z * fx - orl.x * fz;
                f6[5] += orl.x * fy - orl.y * fx;
            }
        }
        for (int k = 0; k < 6; ++k) v[k] = (double)f6[k];
        block_sum(v, 6);
        if (split) {
            exchange(s_ft, 6, (2ull * step) & 0x7FFFFull);
            if (s_abort) break;
        }

        // proposal, formed by every thread from the same totals (identical
        // values everywhere, no broadcast): the displacement or rotation, the
        // new centre, and for a rotation its quaternion and coefficients
        const double A = d.rotate ? a_r : a_t;
        double fo[3], disp[3], qr[4] = {1.0, 0.0, 0.0, 0.0}, cf[9];
        float ncx = cx0, ncy = cy0, ncz = cz0;
        for (int k = 0; k < 3; ++k) fo[k] = s_ft[(d.rotate ? 3 : 0) + k];
        for (int k = 0; k < 3; ++k) disp[k] = beta * A * fo[k] + sqrt(2.0 * A) * d.n[k];
        if (d.rotate) {
            const double ang = sqrt(disp[0] * disp[0] + disp[1] * disp[1] + disp[2] * disp[2]);
            if (ang > 0.0) {
                double sh, ch;
                remc_sincos_turns(0.5 * ang / 6.283185307179586476925, &sh, &ch);
                qr[0] = ch; qr[1] = sh * disp[0] / ang; qr[2] = sh * disp[1] / ang; qr[3] = sh * disp[2] / ang;
                const double qm = sqrt(qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
                for (int k = 0; k < 4; ++k) qr[k] = qr[k] / qm;
            }
            remc_qcoeffs(qr, cf);
        } else {
            const float t0 = (float)disp[0], t1 = (float)disp[1], t2 = (float)disp[2];
            disp[0] = t0; disp[1] = t1; disp[2] = t2; // the displacement actually made
            remc_wrap_centre(cx0, cy0, cz0, t0, t1, t2, S.box, &ncx, &ncy, &ncz);
        }
        for (int k = tid; k < nm; k += BLOCK) {
            const float4 orl = s_orel[k];
            float nrx = orl.x, nry = orl.y, nrz = orl.z;
            if (d.rotate) remc_qrotate(cf, orl.x, orl.y, orl.z, &nrx, &nry, &nrz);
            s_nrel[k] = make_float4(nrx, nry, nrz, 0.0f);
            s_new[k] = make_float4(nrx + ncx, nry + ncy, nrz + ncz, 0.0f);
        }
        // what thread 0 commits (the other threads need neither after this)
        if (tid == 0) {
            for (int k = 0; k < 4; ++k) s_qr[k] = qr[k];
            s_ncen[0] = ncx;
            s_ncen[1] = ncy;
            s_ncen[2] = ncz;
        }
        // the forward half of the proposal ratio
        double fwd = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double a = disp[k] - beta * A * fo[k];
            fwd += a * a;
        }
        __syncthreads();

        // pass B: dE and the force and torque at the new position (same split)
        float acc = 0.0f, comp = 0.0f;
        for (int k = 0; k < 6; ++k) f6[k] = 0.f;
        for (int jj0 = l; jj0 < nj; jj0 += 2 * G) {
            const bool on1 = jj0 + G < nj;
            float xj[2], yj[2], zj[2], qj[2], sj[2];
            int tcj[2];
            bool crow[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                load_partner(u == 0 || !on1 ? jj0 : jj0 + G, xj[u], yj[u], zj[u], qj[u], sj[u], tcj[u]);
                crow[u] = crow_m || (tcj[u] >> 8) != 0;
            }
            for (int k = k0; k < k1; ++k) {
                const float4 o = s_old[k], orl = s_orel[k], n = s_new[k], nrl = s_nrel[k];
                const int trow = (s_tc[k] & 0xff) * REMC_NTYPES;
                float d = 0.f, fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const float eij = s_etab[trow + (tcj[u] & 0xff)];
                    const float sij = 0.5f * (orl.w + sj[u]);
                    const float qq = o.w * qj[u];
                    float eo, en, go, g2;
                    {
                        const float dx = min_image(o.x - xj[u], S.box, S.inv_box);
                        const float dy = min_image(o.y - yj[u], S.box, S.inv_box);
                        const float dz = min_image(o.z - zj[u], S.box, S.inv_box);
                        pair_eg(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), crow[u], eij, sij * sij, qq,
                                inv_xi, ff, eo, go);
                    }
                    const float dx = min_image(n.x - xj[u], S.box, S.inv_box);
                    const float dy = min_image(n.y - yj[u], S.box, S.inv_box);
                    const float dz = min_image(n.z - zj[u], S.box, S.inv_box);
                    pair_eg(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), crow[u], eij, sij * sij, qq, inv_xi,
                            ff, en, g2);
                    const bool use = u == 0 || on1;
                    // the two partners' differences are added (one rounding, as
                    // each difference has), then to the compensated sum
                    d += use ? en - eo : 0.0f;
                    if (!use) g2 = 0.0f;
                    fx -= g2 * dx;
                    fy -= g2 * dy;
                    fz -= g2 * dz;
                }
                kahan_add(acc, comp, d);
                f6[0] += fx;
                f6[1] += fy;
                f6[2] += fz;
                f6[3] += nrl.y * fz - nrl.z * fy;
                f6[4] += nrl.z * fx - nrl.x * fz;
                f6[5] += nrl.x * fy - nrl.y * fx;
            }
        }
        for (int k = 0; k < 6; ++k) v[k] = (double)f6[k];
        v[6] = (double)acc - (double)comp;
        block_sum(v, 7);
        if (split) {
            exchange(s_ft, 7, (2ull * step + 1ull) & 0x7FFFFull);
            if (s_abort) break;
        }
        // the decision, taken by every thread from the same totals
        double rev = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double b = -disp[k] - beta * A * s_ft[(d.rotate ? 3 : 0) + k];
            rev += b * b;
        }
        const double dE = s_ft[6];
        const double x = beta * dE + (rev - fwd) / (4.0 * A);
        const int dec = remc_metropolis_l(x, 1.0, d.nlu);
        if (dec) {
            for (int k = tid; k < nm; k += BLOCK) {
                const int b = m0 + k;
                const float4 n = s_new[k];
                gx[b] = n.x;
                gy[b] = n.y;
                gz[b] = n.z;
                if (cached) {
                    s_jx[b] = n.x;
                    s_jy[b] = n.y;
                    s_jz[b] = n.z;
                }
                if (d.rotate) {
                    const float4 nrl = s_nrel[k];
                    grx[b] = nrl.x;
                    gry[b] = nrl.y;
                    grz[b] = nrl.z;
                    if (cache_rel) {
                        s_rx[b] = nrl.x;
                        s_ry[b] = nrl.y;
                        s_rz[b] = nrl.z;
                    }
                }
            }
            if (tid == 0) {
                if (d.rotate) {
                    double qo[4], qn[4];
                    for (int k = 0; k < 4; ++k) qo[k] = cache_bodies ? s_bq[4 * body + k] : gq[4 * body + k];
                    remc_qmul(s_qr, qo, qn);
                    for (int k = 0; k < 4; ++k) {
                        gq[4 * body + k] = qn[k];
                        if (cache_bodies) s_bq[4 * body + k] = qn[k];
                    }
                } else {
                    gcx[body] = s_ncen[0];
                    gcy[body] = s_ncen[1];
                    gcz[body] = s_ncen[2];
                    if (cache_bodies) {
                        s_bcx[body] = s_ncen[0];
                        s_bcy[body] = s_ncen[1];
                        s_bcz[body] = s_ncen[2];
                    }
                }
            }
            E += dE;
            if (dec == 1) ++n_acc; else ++n_accb;
        } else {
            ++n_rej;
        }
        if (c == 0 && p.rec && tid == 0) {
            const unsigned long long rec_start = p.rec_start[r];
            const unsigned long long idx = step - rec_start;
            if (step >= rec_start && idx < (unsigned long long)p.rec_max) {
                remc_step_record rr;
                rr.body = body;
                rr.move = d.rotate ? 1 : 0;
                rr.accepted = dec;
                rr.ticks = (int32_t)(wall_clock64() - t_step);
                rr.delta = dE;
                p.rec[(size_t)r * p.rec_max + idx] = rr;
            }
        }
        __syncthreads();
    }
    if (c == 0 && tid == 0 && !s_abort) {
        p.E[r] = E;
        p.step[r] = step;
        p.cnt[3 * r + 0] += n_acc;
        p.cnt[3 * r + 1] += n_accb;
        p.cnt[3 * r + 2] += n_rej;
    }
}

size_t mc_fb_dynamic_lds(const KParams &p)
{
    return (4 * sizeof(float4) + sizeof(int)) * (size_t)p.lds_mb +
           (p.lds_cl > 0 ? 36 * (size_t)p.lds_cl + 8 + 44 * (size_t)p.lds_nbody : 0);
}

template <int BLOCK, bool CACHE>
static hipError_t launch_mc_fb_t(const KParams &p, int n_rep, int nsteps, int coop, hipStream_t s)
{
    const size_t lds = mc_fb_dynamic_lds(p);
    hipError_t e = hipFuncSetAttribute((const void *)k_mc_fb<BLOCK, CACHE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_mc_fb<BLOCK, CACHE>), dim3(n_rep * coop), dim3(BLOCK), lds, s, p, nsteps, coop);
    return hipGetLastError();
}

template <int BLOCK, bool CACHE>
static hipError_t fb_occupancy_t(const KParams &p, int *out)
{
    const size_t lds = mc_fb_dynamic_lds(p);
    const void *f = (const void *)k_mc_fb<BLOCK, CACHE>;
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_mc_fb<BLOCK, CACHE>, BLOCK, lds);
}

template <int BLOCK>
static hipError_t fb_occupancy_b(const KParams &p, int *out)
{
    return p.lds_cl > 0 ? fb_occupancy_t<BLOCK, true>(p, out) : fb_occupancy_t<BLOCK, false>(p, out);
}

hipError_t mc_fb_max_active_per_cu(const KParams &p, int block, int *out)
{
    switch (block) {
    case 256: return fb_occupancy_b<256>(p, out);
    case 512: return fb_occupancy_b<512>(p, out);
    case 1024: return fb_occupancy_b<1024>(p, out);
    default: return hipErrorInvalidValue;
    }
}

template <int BLOCK>
static hipError_t launch_mc_fb_b(const KParams &p, int n_rep, int nsteps, int coop, hipStream_t s)
{
    return p.lds_cl > 0 ? launch_mc_fb_t<BLOCK, true>(p, n_rep, nsteps, coop, s)
                        : launch_mc_fb_t<BLOCK, false>(p, n_rep, nsteps, coop, s);
}

hipError_t launch_mc_fb(const KParams &p, int n_rep, int nsteps, int block, int coop, hipStream_t s)
{
    if (n_rep <= 0 || nsteps <= 0) return hipSuccess;
    if (coop < 1 || coop > REMC_MAX_COOP) return hipErrorInvalidValue;
    if (coop > 1) {
        const int n = n_rep * REMC_SLOTS_PER_REP;
        hipLaunchKernelGGL(k_reset_slots, dim3((n + 255) / 256), dim3(256), 0, s, p.slots, n);
    }
    switch (block) {
    case 256: return launch_mc_fb_b<256>(p, n_rep, nsteps, coop, s);
    case 512: return launch_mc_fb_b<512>(p, n_rep, nsteps, coop, s);
    case 1024: return launch_mc_fb_b<1024>(p, n_rep, nsteps, coop, s);
    default: return hipErrorInvalidValue;
    }
}

// kernel variant: 0 = no partner cache, 1 = partner cache, 2 = speculative windows
static int mc_variant(const KParams &p) { return p.lds_cl <= 0 ? 0 : p.spec ? 2 : 1; }

template <int BLOCK>
static hipError_t launch_mc_b(const KParams &p, int grid, int nsteps, int coop, hipStream_t s)
{
    switch (mc_variant(p)) {
    case 0: return launch_mc_t<BLOCK, false, false>(p, grid, nsteps, coop, s);
    case 1: return launch_mc_t<BLOCK, true, false>(p, grid, nsteps, coop, s);
    default: return launch_mc_t<BLOCK, true, true>(p, grid, nsteps, coop, s);
    }
}

hipError_t launch_mc(const KParams &p, int n_rep, int nsteps, int block, int coop, hipStream_t s)
{
    if (n_rep <= 0 || nsteps <= 0) return hipSuccess;
    if (coop < 1 || coop > REMC_MAX_COOP) return hipErrorInvalidValue;
    // the speculative kernel reads its partner range from the cache (the whole replica, or with
    // every member on its own slice the member's slice)
    if (p.spec && p.lds_cl <= 0) return hipErrorInvalidValue;
    if (coop > 1) {
        const int n = n_rep * REMC_SLOTS_PER_REP;
        hipLaunchKernelGGL(k_reset_slots, dim3((n + 255) / 256), dim3(256), 0, s, p.slots, n);
    }
    const int grid = n_rep * coop;
    switch (block) {
    case 256: return launch_mc_b<256>(p, grid, nsteps, coop, s);
    case 512: return launch_mc_b<512>(p, grid, nsteps, coop, s);
    case 1024: return launch_mc_b<1024>(p, grid, nsteps, coop, s);
    default: return hipErrorInvalidValue;
    }
}

template <int BLOCK, bool CACHE, bool SPEC>
static hipError_t occupancy_t(const KParams &p, int *out)
{
    const size_t lds = mc_dynamic_lds(p);
    const void *f = (const void *)k_mc<BLOCK, CACHE, SPEC>;
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_mc<BLOCK, CACHE, SPEC>, BLOCK, lds);
}

template <int BLOCK>
static hipError_t occupancy_b(const KParams &p, int *out)
{
    switch (mc_variant(p)) {
    case 0: return occupancy_t<BLOCK, false, false>(p, out);
    case 1: return occupancy_t<BLOCK, true, false>(p, out);
    default: return occupancy_t<BLOCK, true, true>(p, out);
    }
}

hipError_t mc_max_active_per_cu(const KParams &p, int block, int *out)
{
    switch (block) {
    case 256: return occupancy_b<256>(p, out);
    case 512: return occupancy_b<512>(p, out);
    case 1024: return occupancy_b<1024>(p, out);
    default: return hipErrorInvalidValue;
    }
}

template <int BLOCK>
static int static_lds_b()
{
    // the window variant holds the most: its slot staging on top of the rest
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, (const void *)k_mc<BLOCK, true, true>) != hipSuccess) return -1;
    return (int)a.sharedSizeBytes;
}

int mc_static_lds(int block)
{
    switch (block) {
    case 256: return static_lds_b<256>();
    case 512: return static_lds_b<512>();
    case 1024: return static_lds_b<1024>();
    default: return -1;
    }
}

// --------------------------------------------------------------------------
// batched full energy
// --------------------------------------------------------------------------
// One workgroup per (row tile a, offset bo) unit of the half-window tile
// schedule: tile a against tile (a + bo) mod G for bo = 0..G/2 (the last
// offset only for a < G/2 when G is even), so every unordered tile pair is
// visited once.  Unit x = a (G/2 + 1) + bo of each replica writes one FP64
// partial (0 for the skipped units); k_sum_partials adds them in a fixed
// order.  When the grid is small (a few replicas of a thousand beads) each
// unit's partner tile is further cut into `split` equal sub-ranges, one
// workgroup each, so that every SIMD holds several waves.
__global__ __launch_bounds__(FULL_TILE) void k_energy_full(KParams p, int stride, int split, double *partials,
                                                           int noncrowder_only)
{
    __shared__ float s_etab[REMC_NTYPES * REMC_NTYPES];
    __shared__ float s_x[FULL_TILE], s_y[FULL_TILE], s_z[FULL_TILE], s_q[FULL_TILE], s_s[FULL_TILE];
    __shared__ int s_tc[FULL_TILE], s_body[FULL_TILE];
    __shared__ double s_red[FULL_TILE / 64];

    const int tid = threadIdx.x;
    const int x = blockIdx.x / split, sub = blockIdx.x % split, r = blockIdx.y;
    const DevRep R = p.rep[r];
    const DevSys S = p.sys[R.sys];
    const int N = S.n_beads;
    const int G = (N + FULL_TILE - 1) / FULL_TILE;
    const int h = G / 2;
    if (x >= full_units(G)) return; // whole workgroup leaves before any barrier
    const int a = x / (h + 1), bo = x % (h + 1);
    if ((G % 2 == 0) && bo == h && bo > 0 && a >= h) {
        if (tid == 0) partials[(size_t)r * stride + blockIdx.x] = 0.0;
        return;
    }
    const DevFF ff = p.ff;
    for (int k = tid; k < REMC_NTYPES * REMC_NTYPES; k += FULL_TILE) s_etab[k] = p.etab[k];

    const float *gx = p.px + R.bead_off, *gy = p.py + R.bead_off, *gz = p.pz + R.bead_off;
    const int *tcs = p.tc + S.bead_off, *bodies = p.body + S.bead_off;
    const float *qs = p.q + S.bead_off, *sgs = p.sig + S.bead_off;

    const int i = a * FULL_TILE + tid;
    const bool vi = i < N;
    const float xi = vi ? gx[i] : 0.f, yi = vi ? gy[i] : 0.f, zi = vi ? gz[i] : 0.f;
    const int tci = vi ? tcs[i] : 0, bi = vi ? bodies[i] : -1;
    const float qi = vi ? qs[i] * ff.c_dh : 0.f, si = vi ? sgs[i] : 0.f;
    const bool ci = (tci >> 8) != 0;
    const float *erow = s_etab + (tci & 0xff) * REMC_NTYPES;

    float acc = 0.0f, comp = 0.0f;
    {
        const int b = (a + bo) % G;
        const int sw = FULL_TILE / split;
        const int j0 = b * FULL_TILE + sub * sw;
        const int jt = j0 + tid;
        if (tid < sw && jt < N) {
            s_x[tid] = gx[jt]; s_y[tid] = gy[jt]; s_z[tid] = gz[jt];
            s_q[tid] = qs[jt]; s_s[tid] = sgs[jt]; s_tc[tid] = tcs[jt]; s_body[tid] = bodies[jt];
        }
        __syncthreads();
        // partner sub-range [j0, j0 + nj): only loaded entries are read
        const int nj = vi ? min(sw, N - j0) : 0;
        const int kstart = (bo == 0) ? max(tid + 1 - sub * sw, 0) : 0;
        for (int k = kstart; k < nj; ++k) {
            if (s_body[k] == bi) continue;
            const int tcj = s_tc[k];
            const bool cj = (tcj >> 8) != 0;
            if (noncrowder_only && (ci || cj)) continue;
            float dx = min_image(s_x[k] - xi, S.box, S.inv_box);
            float dy = min_image(s_y[k] - yi, S.box, S.inv_box);
            float dz = min_image(s_z[k] - zi, S.box, S.inv_box);
            float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
            float e;
            if (ci || cj) {
                e = pair_crowder(r2, ff);
            } else {
                const float sij = 0.5f * (si + s_s[k]);
                e = pair_lj_dh(r2, erow[tcj & 0xff], sij * sij, qi * s_q[k], ff);
            }
            kahan_add(acc, comp, e);
        }
    }
    double v = wave_sum((double)acc - (double)comp);
    if ((tid & 63) == 0) s_red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < FULL_TILE / 64; ++w) t += s_red[w];
        partials[(size_t)r * stride + blockIdx.x] = t;
    }
}

// one wave per replica: lane l adds units l, l+64, ... in order, then a
// fixed butterfly across the wave (the same total on every call)
__global__ __launch_bounds__(64) void k_sum_partials(KParams p, int stride, int split, const double *partials,
                                                     double *out, int write_E)
{
    const int r = blockIdx.x, lane = threadIdx.x;
    const DevSys S = p.sys[p.rep[r].sys];
    const int U = full_units((S.n_beads + FULL_TILE - 1) / FULL_TILE) * split;
    double t = 0.0;
    for (int x = lane; x < U; x += 64) t += partials[(size_t)r * stride + x];
    t = wave_sum(t);
    if (lane == 0) {
        if (out) out[r] = t;
        if (write_E) p.E[r] = t;
    }
}

int full_split_default(int n_rep, int max_tiles)
{
    const long long units = (long long)full_units(max_tiles) * n_rep;
    return units >= 2048 ? 1 : units >= 1024 ? 2 : FULL_SPLIT_MAX;
}

hipError_t launch_energy_full(const KParams &p, int n_rep, int max_tiles, double *partials, double *out,
                              int noncrowder_only, int write_E, int split, hipStream_t s)
{
    if (n_rep <= 0) return hipSuccess;
    const int max_units = full_units(max_tiles);
    if (split != 1 && split != 2 && split != FULL_SPLIT_MAX) split = full_split_default(n_rep, max_tiles);
    const int stride = max_units * FULL_SPLIT_MAX;
    hipLaunchKernelGGL(k_energy_full, dim3(max_units * split, n_rep), dim3(FULL_TILE), 0, s, p, stride, split,
                       partials, noncrowder_only);
    hipLaunchKernelGGL(k_sum_partials, dim3(n_rep), dim3(64), 0, s, p, stride, split, partials, out, write_E);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// reference-shaped plain kernel (one replica, FP32 tiles, host FP64 sum)
// --------------------------------------------------------------------------
template <int B>
__global__ __launch_bounds__(B) void k_energy_plain(KParams p, int replica, float *partials)
{
    __shared__ float s_etab[REMC_NTYPES * REMC_NTYPES];
    __shared__ float s_x[B], s_y[B], s_z[B], s_q[B], s_s[B];
    __shared__ int s_tc[B], s_body[B];
    __shared__ float s_res[B];
    const int tx = threadIdx.x;
    const int bx = blockIdx.x, by = blockIdx.y;
    const int G = gridDim.x;
    if (bx < by) return; // lower triangle culled, whole block (CudaFunctions.cu:483-486)
    const DevRep R = p.rep[replica];
    const DevSys S = p.sys[R.sys];
    const DevFF ff = p.ff;
    const int N = S.n_beads;
    for (int k = tx; k < REMC_NTYPES * REMC_NTYPES; k += B) s_etab[k] = p.etab[k];
    const float *gx = p.px + R.bead_off, *gy = p.py + R.bead_off, *gz = p.pz + R.bead_off;
    const int *tcs = p.tc + S.bead_off, *bodies = p.body + S.bead_off;
    const float *qs = p.q + S.bead_off, *sgs = p.sig + S.bead_off;

    const int jt = bx * B + tx;
    if (jt < N) {
        s_x[tx] = gx[jt]; s_y[tx] = gy[jt]; s_z[tx] = gz[jt];
        s_q[tx] = qs[jt]; s_s[tx] = sgs[jt]; s_tc[tx] = tcs[jt]; s_body[tx] = bodies[jt];
    } else {
        s_body[tx] = -2; // padding
    }
    __syncthreads();
    const int i = by * B + tx;
    float sum = 0.0f;
    if (i < N) {
        const float xi = gx[i], yi = gy[i], zi = gz[i];
        const int tci = tcs[i], bi = bodies[i];
        const float qi = qs[i] * ff.c_dh, si = sgs[i];
        const bool ci = (tci >> 8) != 0;
        for (int k = 0; k < B; ++k) {
            if (s_body[k] == bi || s_body[k] < 0) continue; // same body or padding
            const int tcj = s_tc[k];
            const bool cj = (tcj >> 8) != 0;
            float dx = min_image(s_x[k] - xi, S.box, S.inv_box);
            float dy = min_image(s_y[k] - yi, S.box, S.inv_box);
            float dz = min_image(s_z[k] - zi, S.box, S.inv_box);
            float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
            if (ci || cj) {
                sum += pair_crowder(r2, ff);
            } else {
                const float sij = 0.5f * (si + s_s[k]);
                sum += pair_lj_dh<false>(r2, s_etab[(tci & 0xff) * REMC_NTYPES + (tcj & 0xff)], sij * sij, qi * s_q[k], ff);
            }
        }
        if (bx == by) sum *= 0.5f; // diagonal tile counts both orders (CudaFunctions.cu:651-653)
    }
    s_res[tx] = sum;
    __syncthreads();
    for (int off = B / 2; off > 0; off >>= 1) {
        if (tx < off) s_res[tx] += s_res[tx + off];
        __syncthreads();
    }
    if (tx == 0) partials[bx * G + by] = s_res[0];
}

hipError_t launch_energy_plain(const KParams &p, int replica, int n_beads, int block, float *partials, hipStream_t s)
{
    const int G = (n_beads + block - 1) / block;
    switch (block) {
    case 32: hipLaunchKernelGGL(k_energy_plain<32>, dim3(G, G), dim3(32), 0, s, p, replica, partials); break;
    case 64: hipLaunchKernelGGL(k_energy_plain<64>, dim3(G, G), dim3(64), 0, s, p, replica, partials); break;
    case 128: hipLaunchKernelGGL(k_energy_plain<128>, dim3(G, G), dim3(128), 0, s, p, replica, partials); break;
    case 256: hipLaunchKernelGGL(k_energy_plain<256>, dim3(G, G), dim3(256), 0, s, p, replica, partials); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// replica exchange
// --------------------------------------------------------------------------
__global__ void k_exchange(KParams p)
{
    const int sidx = blockIdx.x;
    const DevSys S = p.sys[sidx];
    const int offset = p.ex_offset[sidx];
    const unsigned long long round = p.ex_round[sidx];
    int *slots = p.slot_rep + S.slot_off;
    const double *lad = p.ladder + S.slot_off;
    int nsw = 0;
    for (int k = offset + 2 * (int)threadIdx.x; k < S.n_rep - 1; k += 2 * (int)blockDim.x) {
        const int ri = slots[k], rj = slots[k + 1];
        const double u = remc_exchange_u(round, (uint32_t)k, (uint32_t)sidx, p.key0, p.key1);
        if (remc_exchange_accept(lad[k], lad[k + 1], p.E[ri], p.E[rj], u)) {
            slots[k] = rj;
            slots[k + 1] = ri;
            p.T[ri] = lad[k + 1];
            p.T[rj] = lad[k];
            ++nsw;
        }
    }
    for (int off = 32; off > 0; off >>= 1) nsw += __shfl_xor(nsw, off, 64);
    if (threadIdx.x == 0) {
        p.ex_swaps[sidx] += (unsigned long long)nsw;
        p.ex_round[sidx] = round + 1;
        p.ex_offset[sidx] = 1 - offset;
    }
}

hipError_t launch_exchange(const KParams &p, int n_sys, hipStream_t s)
{
    if (n_sys <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_exchange, dim3(n_sys), dim3(64), 0, s, p);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// adjoint: gradient of the full energy, every replica of a range in one launch
//
// k_adjoint_rows: one workgroup of ADJ_WAVES waves per (row tile of 64 beads,
// partner sub-range, replica).  The sub-range is staged through LDS in chunks
// of ADJ_BLOCK beads; wave w takes the staged beads w, w + ADJ_WAVES, ...
// against its lane's row bead, so every staged value is read as a broadcast.
// A lane accumulates in FP64, from FP32 pair derivatives (the same branches as
// the pair term), dE/dx_i of its row bead and the parameter derivatives
// dE/dlambda = sum dE/dE_ij (T_ij - e0), dE/de0 = -lambda sum dE/dE_ij and
// dE/dxi = sum DH r / xi^2 over the unordered pairs it owns.  The waves' row
// gradients are added through LDS in wave order and written once per
// (sub-range, bead), their parameter sums once per workgroup.
// k_adjoint_bodies: one wave per (replica, body) adds each bead's sub-range
// partials in sub-range order, forms F = -sum dE/dx_i and the torque
// sum rel_i x F_i about the body centre with fixed-order wave sums; the first
// wave of each replica adds the parameter partials.  The result does not
// depend on scheduling.  Sub-ranges are only cut when the launch holds fewer
// than ADJ_TARGET_WG (row tile, replica) workgroups.
// --------------------------------------------------------------------------
struct AdjFF {
    float table[REMC_NTYPES * REMC_NTYPES]; // raw table T_ij
    float lambda, e0, xi, c_lj, c_dh;
};

constexpr int ADJ_ROWS = 64;                 // row beads per workgroup (one per lane)
constexpr int ADJ_WAVES = 4;                 // waves per workgroup
constexpr int ADJ_BLOCK = 64 * ADJ_WAVES;    // staged partner beads per chunk
constexpr int ADJ_TARGET_WG = 2048;          // workgroups a launch aims at

__global__ __launch_bounds__(ADJ_BLOCK) void k_adjoint_rows(KParams p, AdjFF af, int rep_first, int tiles, int split,
                                                            int sub_len, int bead_base, size_t bead_stride,
                                                            double *part_grad, double *param_part)
{
    __shared__ float s_tab[REMC_NTYPES * REMC_NTYPES];
    __shared__ float s_x[ADJ_BLOCK], s_y[ADJ_BLOCK], s_z[ADJ_BLOCK], s_q[ADJ_BLOCK], s_s[ADJ_BLOCK];
    __shared__ int s_tc[ADJ_BLOCK], s_body[ADJ_BLOCK];
    __shared__ double s_g[ADJ_WAVES - 1][3][ADJ_ROWS];
    __shared__ double s_pp[ADJ_WAVES][3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tile = blockIdx.x / split, sub = blockIdx.x % split;
    const int rr = blockIdx.y;
    const DevRep R = p.rep[rep_first + rr];
    const DevSys S = p.sys[R.sys];
    const DevFF ff = p.ff;
    const int N = S.n_beads;
    const int j0 = sub * sub_len;
    if (tile * ADJ_ROWS >= N || j0 >= N) return; // the whole workgroup, before any barrier
    const int j1 = min(N, j0 + sub_len);
    for (int k = tid; k < REMC_NTYPES * REMC_NTYPES; k += ADJ_BLOCK) s_tab[k] = af.table[k];
    const float *gx = p.px + R.bead_off, *gy = p.py + R.bead_off, *gz = p.pz + R.bead_off;
    const int *tcs = p.tc + S.bead_off, *bodies = p.body + S.bead_off;
    const float *qs = p.q + S.bead_off, *sgs = p.sig + S.bead_off;
    const int i = tile * ADJ_ROWS + lane;
    const bool vi = i < N;
    const float xi = vi ? gx[i] : 0.f, yi = vi ? gy[i] : 0.f, zi = vi ? gz[i] : 0.f;
    const int tci = vi ? tcs[i] : 0, bi = vi ? bodies[i] : -1;
    const float qi = vi ? qs[i] : 0.f, si = vi ? sgs[i] : 0.f;
    const bool ci = (tci >> 8) != 0;
    double gxa = 0.0, gya = 0.0, gza = 0.0, dl = 0.0, de0 = 0.0, dxi = 0.0;
    const float inv_xi2 = 1.0f / (af.xi * af.xi);
    for (int c0 = j0; c0 < j1; c0 += ADJ_BLOCK) {
        const int nc = min(ADJ_BLOCK, j1 - c0);
        __syncthreads(); // the previous chunk is consumed (first pass: the table is written)
        if (tid < nc) {
            const int jt = c0 + tid;
            s_x[tid] = gx[jt]; s_y[tid] = gy[jt]; s_z[tid] = gz[jt];
            s_q[tid] = qs[jt]; s_s[tid] = sgs[jt]; s_tc[tid] = tcs[jt]; s_body[tid] = bodies[jt];
        }
        __syncthreads();
        if (vi) {
            for (int k = wave; k < nc; k += ADJ_WAVES) {
                if (s_body[k] == bi) continue;
                const int tcj = s_tc[k];
                const bool cj = (tcj >> 8) != 0;
                const float dx = min_image(xi - s_x[k], S.box, S.inv_box);
                const float dy = min_image(yi - s_y[k], S.box, S.inv_box);
                const float dz = min_image(zi - s_z[k], S.box, S.inv_box);
                const float r2 = fmaxf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), 1e-36f);
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                float dedr;
                if (ci || cj) {
                    const float t = ff.crow_s2 * (rinv * rinv);
                    const float t6 = t * t * t;
                    dedr = r2 < ff.crow_cut2 ? -6.0f * ff.c_crow * t6 * rinv : 0.0f;
                } else {
                    const int ti = tci & 0xff, tj = tcj & 0xff;
                    const float traw = s_tab[ti * REMC_NTYPES + tj];
                    const float eij = af.lambda * (traw - af.e0);
                    const float sij = 0.5f * (si + s_s[k]);
                    const float s2 = sij * sij * (rinv * rinv);
                    const float x6 = s2 * s2 * s2;
                    const float base = 4.0f * x6 * (x6 - 1.0f);
                    const float dbase = 4.0f * (2.0f * x6 - 1.0f) * (-6.0f * x6 * rinv);
                    const bool flip = (eij > 0.0f) & (x6 > 0.5f);
                    const float dlj = flip ? eij * dbase : -eij * dbase;
                    const float dlj_de = flip ? base + 2.0f : -base;
                    const float dh = af.c_dh * qi * s_q[k] * __builtin_amdgcn_exp2f(r * ff.k_dh) * rinv;
                    const float ddh = -dh * (1.0f / af.xi + rinv);
                    dedr = af.c_lj * dlj + ddh;
                    // each unordered pair once for the parameter sums
                    if (s_body[k] > bi) {
                        const float de_deij = af.c_lj * dlj_de;
                        dl += (double)(de_deij * (traw - af.e0));
                        de0 += (double)(-de_deij * af.lambda);
                        dxi += (double)(dh * r * inv_xi2);
                    }
                }
                const float kf = dedr * rinv;
                gxa += (double)(kf * dx);
                gya += (double)(kf * dy);
                gza += (double)(kf * dz);
            }
        }
    }
    if (wave > 0) {
        s_g[wave - 1][0][lane] = gxa;
        s_g[wave - 1][1][lane] = gya;
        s_g[wave - 1][2][lane] = gza;
    }
    const double v0 = wave_sum_dpp(dl), v1 = wave_sum_dpp(de0), v2 = wave_sum_dpp(dxi);
    if (lane == 0) {
        s_pp[wave][0] = v0;
        s_pp[wave][1] = v1;
        s_pp[wave][2] = v2;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int w = 0; w < ADJ_WAVES - 1; ++w) {
            gxa += s_g[w][0][lane];
            gya += s_g[w][1][lane];
            gza += s_g[w][2][lane];
        }
        if (vi) {
            double *pg = part_grad + 3 * ((size_t)sub * bead_stride + (size_t)(R.bead_off - bead_base) + i);
            pg[0] = gxa;
            pg[1] = gya;
            pg[2] = gza;
        }
        if (lane < 3) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < ADJ_WAVES; ++w) t += s_pp[w][lane];
            param_part[3 * (((size_t)rr * tiles + tile) * split + sub) + lane] = t;
        }
    }
}

// One wave per (replica, body): the lanes stride over the body's beads, then
// wave sums (fixed order, so every call gives the same bits); the first wave of
// workgroup (0, replica) also adds the replica's parameter partials.
__global__ __launch_bounds__(256) void k_adjoint_bodies(KParams p, int rep_first, int tiles, int split, int sub_len,
                                                        int bead_base, size_t bead_stride, int body_base,
                                                        const double *part_grad, const double *param_part,
                                                        double *force_torque, double *dparams)
{
    const int rr = blockIdx.y;
    const DevRep R = p.rep[rep_first + rr];
    const DevSys S = p.sys[R.sys];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int N = S.n_beads;
    const int nsub = (N + sub_len - 1) / sub_len, ntile = (N + ADJ_ROWS - 1) / ADJ_ROWS;
    if (blockIdx.x == 0 && wave == 0) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        for (int a = lane; a < ntile * nsub; a += 64) {
            const int tile = a / nsub, sub = a - tile * nsub;
            const double *pp = param_part + 3 * (((size_t)rr * tiles + tile) * split + sub);
            t0 += pp[0];
            t1 += pp[1];
            t2 += pp[2];
        }
        t0 = wave_sum_dpp(t0);
        t1 = wave_sum_dpp(t1);
        t2 = wave_sum_dpp(t2);
        if (lane == 0) {
            dparams[3 * rr + 0] = t0;
            dparams[3 * rr + 1] = t1;
            dparams[3 * rr + 2] = t2;
        }
    }
    const int m = blockIdx.x * (blockDim.x >> 6) + wave;
    if (m >= S.n_bodies) return;
    const int b0 = p.bstart[S.body_off + m], b1 = p.bend[S.body_off + m];
    const float *rx = p.rx + R.bead_off, *ry = p.ry + R.bead_off, *rz = p.rz + R.bead_off;
    const double *pg = part_grad + 3 * (size_t)(R.bead_off - bead_base);
    double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
    for (int i = b0 + lane; i < b1; i += 64) {
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        for (int s = 0; s < nsub; ++s) {
            const double *q = pg + 3 * ((size_t)s * bead_stride + i);
            g0 += q[0];
            g1 += q[1];
            g2 += q[2];
        }
        const double f0 = -g0, f1 = -g1, f2 = -g2;
        F[0] += f0; F[1] += f1; F[2] += f2;
        const double ax = rx[i], ay = ry[i], az = rz[i];
        T[0] += ay * f2 - az * f1;
        T[1] += az * f0 - ax * f2;
        T[2] += ax * f1 - ay * f0;
    }
    for (int k = 0; k < 3; ++k) {
        F[k] = wave_sum_dpp(F[k]);
        T[k] = wave_sum_dpp(T[k]);
    }
    if (lane == 0) {
        double *o = force_torque + 6 * ((size_t)(R.body_off - body_base) + m);
        for (int k = 0; k < 3; ++k) {
            o[k] = F[k];
            o[3 + k] = T[k];
        }
    }
}

AdjPlan adjoint_plan(int n_rep, int max_beads, size_t bead_stride, int target_wg)
{
    AdjPlan a{};
    a.tiles = (max_beads + ADJ_ROWS - 1) / ADJ_ROWS;
    const long long target = target_wg > 0 ? target_wg : ADJ_TARGET_WG;
    const long long units = (long long)n_rep * a.tiles;
    const long long want = units >= target ? 1 : (target + units - 1) / units;
    const long long per = (max_beads + want - 1) / want;
    a.sub_len = (int)std::max<long long>(ADJ_ROWS, (per + ADJ_ROWS - 1) / ADJ_ROWS * ADJ_ROWS);
    a.split = (max_beads + a.sub_len - 1) / a.sub_len;
    a.part_doubles = 3 * (size_t)a.split * bead_stride;
    a.param_doubles = 3 * (size_t)n_rep * a.tiles * a.split;
    return a;
}

hipError_t launch_adjoint(const KParams &p, const void *adj_ff, const AdjPlan &a, int rep_first, int n_rep,
                          int max_bodies, int bead_base, size_t bead_stride, int body_base, double *scratch,
                          double *force_torque, double *dparams, hipStream_t s)
{
    if (n_rep <= 0) return hipSuccess;
    double *part_grad = scratch, *param_part = scratch + a.part_doubles;
    const AdjFF af = *(const AdjFF *)adj_ff;
    hipLaunchKernelGGL(k_adjoint_rows, dim3(a.tiles * a.split, n_rep), dim3(ADJ_BLOCK), 0, s, p, af, rep_first,
                       a.tiles, a.split, a.sub_len, bead_base, bead_stride, part_grad, param_part);
    hipLaunchKernelGGL(k_adjoint_bodies, dim3((max_bodies + 3) / 4, n_rep), dim3(256), 0, s, p, rep_first, a.tiles,
                       a.split, a.sub_len, bead_base, bead_stride, body_base, part_grad, param_part, force_torque,
                       dparams);
    return hipGetLastError();
}

size_t adjoint_ff_size() { return sizeof(AdjFF); }

void adjoint_ff_fill(void *dst, const float *table, double lambda, double e0, double xi, double c_lj, double c_dh)
{
    AdjFF *a = (AdjFF *)dst;
    for (int k = 0; k < REMC_NTYPES * REMC_NTYPES; ++k) a->table[k] = table[k];
    a->lambda = (float)lambda;
    a->e0 = (float)e0;
    a->xi = (float)xi;
    a->c_lj = (float)c_lj;
    a->c_dh = (float)c_dh;
}

} // namespace remc
