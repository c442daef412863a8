// remc_analysis.hip -- device reductions over the live replica conformations.
//
//   k_drms             distance RMS of every replica of a system against a
//                      reference conformation, over bead pairs of different
//                      non-crowder bodies ("post processor/pdbgensrc/
//                      drms.cpp:7-58"): one workgroup per (replica, 256-bead
//                      row tile); FP32 distances, FP64 per-thread and
//                      per-tile sums, tiles added in fixed order.
//   k_centre_hist      histogram of body-centre distances (minimum image,
//                      FP64) by body-type pair over every replica of a system
//                      ("post processor/gr.py:93-140"): one thread per
//                      (replica, body a), workgroup-local LDS histogram.
#include "remc_internal.h"

namespace remc {

constexpr int DRMS_TILE = 256;

// FP32 Euclidean distance with every operation rounded (no contraction), so
// that a host restatement in numpy float32 gives the same value.
__device__ __forceinline__ float plain_dist(float dx, float dy, float dz)
{
    const float s = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
    return __fsqrt_rn(s);
}

__global__ __launch_bounds__(DRMS_TILE) void k_drms(KParams p, int sys_id, const float *rx, const float *ry,
                                                    const float *rz, double *partials, int n_tiles)
{
    __shared__ float s_x[DRMS_TILE], s_y[DRMS_TILE], s_z[DRMS_TILE], s_rx[DRMS_TILE], s_ry[DRMS_TILE],
        s_rz[DRMS_TILE];
    __shared__ int s_b[DRMS_TILE];
    __shared__ double s_red[DRMS_TILE / 64];
    const DevSys S = p.sys[sys_id];
    const int k = blockIdx.y; // the system's k-th replica
    const DevRep R = p.rep[S.rep_first + k];
    const int N = S.n_beads;
    const int it = blockIdx.x;
    const int i = it * DRMS_TILE + threadIdx.x;
    const float *px = p.px + R.bead_off, *py = p.py + R.bead_off, *pz = p.pz + R.bead_off;
    const int *body = p.body + S.bead_off;
    const int *crow = p.bcrow + S.body_off;
    float xi = 0.f, yi = 0.f, zi = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
    int bi = -1;
    if (i < N) {
        bi = body[i];
        if (crow[bi]) bi = -1;
        xi = px[i];
        yi = py[i];
        zi = pz[i];
        ax = rx[i];
        ay = ry[i];
        az = rz[i];
    }
    double acc = 0.0;
    // row tile it against column tiles jt >= it; within the diagonal tile only j > i
    for (int jt = it; jt < n_tiles; ++jt) {
        const int j0 = jt * DRMS_TILE;
        const int jl = j0 + threadIdx.x;
        __syncthreads();
        if (jl < N) {
            int bj = body[jl];
            s_b[threadIdx.x] = crow[bj] ? -1 : bj;
            s_x[threadIdx.x] = px[jl];
            s_y[threadIdx.x] = py[jl];
            s_z[threadIdx.x] = pz[jl];
            s_rx[threadIdx.x] = rx[jl];
            s_ry[threadIdx.x] = ry[jl];
            s_rz[threadIdx.x] = rz[jl];
        } else {
            s_b[threadIdx.x] = -1;
        }
        __syncthreads();
        if (bi < 0) continue;
        const int jn = min(DRMS_TILE, N - j0);
        const int jstart = jt == it ? threadIdx.x + 1 : 0;
        for (int t = jstart; t < jn; ++t) {
            const int bj = s_b[t];
            if (bj < 0 || bj == bi) continue;
            const float d = plain_dist(xi - s_x[t], yi - s_y[t], zi - s_z[t]);
            const float e = plain_dist(ax - s_rx[t], ay - s_ry[t], az - s_rz[t]);
            acc += (double)fabsf(__fsub_rn(d, e)); // |float - float| as the reference's abs of magnitudes
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < DRMS_TILE / 64; ++w) t += s_red[w];
        partials[(size_t)k * n_tiles + it] = t;
    }
}

__global__ void k_drms_sum(const double *partials, int n_tiles, int n_rep, double inv_pairs, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_rep) return;
    double t = 0.0;
    for (int i = 0; i < n_tiles; ++i) t += partials[(size_t)k * n_tiles + i];
    out[k] = t * inv_pairs;
}

hipError_t launch_drms(const KParams &p, int sys_id, int n_beads, int n_rep, const float *rx, const float *ry,
                       const float *rz, double *partials, double inv_pairs, double *out, hipStream_t s)
{
    if (n_rep <= 0 || n_beads <= 0) return hipSuccess;
    const int n_tiles = (n_beads + DRMS_TILE - 1) / DRMS_TILE;
    hipLaunchKernelGGL(k_drms, dim3(n_tiles, n_rep), dim3(DRMS_TILE), 0, s, p, sys_id, rx, ry, rz, partials, n_tiles);
    hipLaunchKernelGGL(k_drms_sum, dim3((n_rep + 63) / 64), dim3(64), 0, s, partials, n_tiles, n_rep, inv_pairs, out);
    return hipGetLastError();
}

constexpr int HIST_BLOCK = 256;

// counts[(ti * n_types + tj) * nbin + bin] for ti <= tj; a pair of bodies of
// the same type counts once from each side, a pair of different types once
// (gr.py's loops over typeI <= typeJ and over both chains).  A body's point is
// its centre, or with `site` the position of its bead site[m] (gr.py takes the
// residue nearest each chain's centroid).
__global__ __launch_bounds__(HIST_BLOCK) void k_centre_hist(KParams p, int sys_id, const int *btype, const int *site,
                                                            int n_types, int nbin, double lo, double dr,
                                                            unsigned long long *counts)
{
    extern __shared__ unsigned int s_hist[];
    const DevSys S = p.sys[sys_id];
    const int k = blockIdx.y;
    const DevRep R = p.rep[S.rep_first + k];
    const int M = S.n_bodies;
    const int nh = n_types * n_types * nbin;
    for (int t = threadIdx.x; t < nh; t += HIST_BLOCK) s_hist[t] = 0u;
    __syncthreads();
    const int a = blockIdx.x * HIST_BLOCK + threadIdx.x;
    const float *cx = p.cx + R.body_off, *cy = p.cy + R.body_off, *cz = p.cz + R.body_off;
    const float *bx = p.px + R.bead_off, *by = p.py + R.bead_off, *bz = p.pz + R.bead_off;
    const double L = (double)S.box;
    if (a < M && btype[a] >= 0) {
        const int ta = btype[a];
        const double xa = site ? bx[site[a]] : cx[a], ya = site ? by[site[a]] : cy[a], za = site ? bz[site[a]] : cz[a];
        for (int b = 0; b < M; ++b) {
            const int tb = btype[b];
            if (b == a || tb < 0 || tb < ta) continue;
            const double xb = site ? bx[site[b]] : cx[b], yb = site ? by[site[b]] : cy[b], zb = site ? bz[site[b]] : cz[b];
            double dx = __dsub_rn(xa, xb), dy = __dsub_rn(ya, yb), dz = __dsub_rn(za, zb);
            if (L > 0.0) {
                dx = __dsub_rn(dx, __dmul_rn(L, round(__ddiv_rn(dx, L))));
                dy = __dsub_rn(dy, __dmul_rn(L, round(__ddiv_rn(dy, L))));
                dz = __dsub_rn(dz, __dmul_rn(L, round(__ddiv_rn(dz, L))));
            }
            const double r = __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)));
            const double f = floor(__ddiv_rn(__dsub_rn(r, lo), dr));
            if (!(f >= 0.0 && f < (double)nbin)) continue;
            atomicAdd(&s_hist[(ta * n_types + tb) * nbin + (int)f], 1u);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nh; t += HIST_BLOCK)
        if (s_hist[t]) atomicAdd(&counts[t], (unsigned long long)s_hist[t]);
}

hipError_t launch_centre_hist(const KParams &p, int sys_id, int n_bodies, int n_rep, const int *btype, const int *site,
                              int n_types, int nbin, double lo, double dr, unsigned long long *counts, hipStream_t s)
{
    if (n_rep <= 0 || n_bodies <= 0) return hipSuccess;
    const size_t lds = sizeof(unsigned int) * (size_t)n_types * n_types * nbin;
    hipLaunchKernelGGL(k_centre_hist, dim3((n_bodies + HIST_BLOCK - 1) / HIST_BLOCK, n_rep), dim3(HIST_BLOCK), lds, s,
                       p, sys_id, btype, site, n_types, nbin, lo, dr, counts);
    return hipGetLastError();
}

// g = counts / (frames 4 pi r^2 dr n_ti n_tj / V), r the bin centre (gr.py:140-155)
__global__ void k_gr_normalise(const unsigned long long *counts, const int *type_count, int n_types, int nbin,
                               double lo, double dr, double frames_over_volume, double *gr)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_types * n_types * nbin) return;
    const int b = t % nbin, pair = t / nbin, ti = pair / n_types, tj = pair % n_types;
    if (ti > tj) {
        gr[t] = 0.0;
        return;
    }
    const double r = lo + (0.5 + (double)b) * dr;
    const double four_pi = 12.566370614359172953850573533118;
    const double norm = frames_over_volume * four_pi * r * r * dr * (double)type_count[ti] * (double)type_count[tj];
    gr[t] = norm > 0.0 ? (double)counts[t] / norm : 0.0;
}

hipError_t launch_gr_normalise(const unsigned long long *counts, const int *type_count, int n_types, int nbin, double lo,
                               double dr, double frames_over_volume, double *gr, hipStream_t s)
{
    const int n = n_types * n_types * nbin;
    hipLaunchKernelGGL(k_gr_normalise, dim3((n + 255) / 256), dim3(256), 0, s, counts, type_count, n_types, nbin, lo, dr,
                       frames_over_volume, gr);
    return hipGetLastError();
}

} // namespace remc
