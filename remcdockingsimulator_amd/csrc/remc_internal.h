// remc_internal.h -- device data layout shared by the kernels and the C-ABI
// implementation.  Everything lives in HBM as structure-of-arrays:
//
//   static, per bead (all systems concatenated, index sys.bead_off + b):
//     tc   int    type | crowder<<8
//     q    float  charge
//     sig  float  vdW radius
//     body int    body index local to the system
//   static, per body (index sys.body_off + m): bstart, bend (local bead
//     indices), bcrow (crowder flag)
//   dynamic, per replica bead (index rep.bead_off + b): px,py,pz positions,
//     rx,ry,rz positions relative to the body centre
//   dynamic, per replica body (index rep.body_off + m): cx,cy,cz centre,
//     quat[4*(...)] rotation (w,x,y,z) in FP64
//   dynamic, per replica: E (FP64, kcal/mol), T (FP64), step (RNG counter),
//     cnt[3] accept/accept_boltz/reject
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "remc.h"

namespace remc {

struct DevSys {
    int bead_off, n_beads, body_off, n_bodies;
    int max_body, rep_first, n_rep, slot_off;
    float box, inv_box;
    int pad0, pad1;
};

struct DevRep {
    int sys, bead_off, body_off;
    uint32_t rng_id;
    float tstep, pad;
    double rot_ch, rot_sh;
    double rstep; // rotate step (rad)
};

// prescaled force field: etab[i*20+j] = lambda*(table-e0)*lj_cf*kbt_cf
struct DevFF {
    float c_crow;    // lj_cf*kbt_cf (crowder term coefficient)
    float crow_s2;   // crowder sigma^2
    float crow_cut2; // crowder cutoff^2
    float c_dh;      // dh_cf*kbt_cf
    float k_dh;      // -log2(e)/xi, exp(-r/xi) = exp2(r*k_dh)
    float eps_r;     // 1e-38 added to r
    float crow_k;    // c_crow * crow_s2^3: crowder term = crow_k / r^6
    float crow_cut;  // crowder cutoff
};

struct KParams {
    // static
    const int *tc;
    const float *q;
    const float *sig;
    const int *body;
    const int *bstart;
    const int *bend;
    const int *bcrow;
    const DevSys *sys;
    const DevRep *rep;
    const float *etab;
    const float *sstab; // per system: (0.5 (sigma_i + sigma_j))^2 / box^2 of type pairs (when sigtab)
    int sigtab;         // 1: every non-crowder bead's sigma is its type's
    DevFF ff;
    // dynamic
    float *px, *py, *pz, *rx, *ry, *rz, *cx, *cy, *cz;
    double *quat;
    double *E, *T;
    unsigned long long *step;
    unsigned long long *cnt;
    // exchange ladder
    int *slot_rep;          // per system slots: slot_rep[sys.slot_off + k]
    double *ladder;         // ladder[sys.slot_off + k]
    unsigned long long *ex_round;  // per system
    int *ex_offset;         // per system
    unsigned long long *ex_swaps;  // per system cumulative
    // cooperative MC: private state copies and partial-dE slots
    int n_rep;
    int coop_min_pairs;     // moves with fewer pair terms are not split across members
    size_t copy_beads;      // stride between state copies (beads)
    size_t copy_bodies;     // stride between state copies (bodies)
    unsigned long long *slots; // n_rep * REMC_SLOTS_PER_REP
    int *err;               // set when a member waited too long
    // MC dynamic LDS: moved-body staging size and partner-bead/body caches
    int lds_mb;             // largest body (beads)
    int lds_cl;             // cached partner beads per workgroup (0 = no cache)
    int lds_nbody;          // cached bodies per workgroup (0 = none)
    int cache_slice;        // 1: each member caches only its partner slice
    int cache_bodies;       // 1: body ranges and centres cached
    int cache_rel;          // 1: relative positions cached (full partner cache only)
    int cache_quat;         // 1: body rotations cached (with the body table)
    int spec;               // 1: speculative windows of small moves (needs the full caches)
    int draw2;              // 1: a second draw batch in dynamic LDS, formed ahead
    // recording
    remc_step_record *rec;
    unsigned long long *rec_start;
    int rec_max;
    uint32_t key0, key1;
};

// launchers (remc_kernels.hip)
hipError_t launch_mc(const KParams &p, int n_rep, int nsteps, int block, int coop, hipStream_t s);
size_t mc_dynamic_lds(const KParams &p);
// static LDS of the MC kernel at this block size (its largest variant), or -1
int mc_static_lds(int block);
size_t mc_draw_batch_bytes(); // LDS of one batch of MC step draws
hipError_t launch_mc_fb(const KParams &p, int n_rep, int nsteps, int block, int coop, hipStream_t s);
hipError_t mc_fb_max_active_per_cu(const KParams &p, int block, int *out);
size_t mc_fb_dynamic_lds(const KParams &p);
hipError_t mc_max_active_per_cu(const KParams &p, int block, int *out);
// split: partner sub-ranges per full-energy unit (1, 2 or FULL_SPLIT_MAX;
// anything else takes full_split_default, a rule on the grid's size)
hipError_t launch_energy_full(const KParams &p, int n_rep, int max_tiles, double *partials, double *out,
                              int noncrowder_only, int write_E, int split, hipStream_t s);
int full_split_default(int n_rep, int max_tiles);
hipError_t launch_energy_plain(const KParams &p, int replica, int n_beads, int block, float *partials, hipStream_t s);
hipError_t launch_exchange(const KParams &p, int n_sys, hipStream_t s);
// adjoint launch plan for n_rep replicas of at most max_beads beads whose
// beads span bead_stride entries of the replica arrays: row tiles, partner
// sub-ranges and the scratch they need (per-sub-range bead gradients, then
// per-workgroup parameter partials)
struct AdjPlan {
    int tiles, split, sub_len;
    size_t part_doubles, param_doubles;
};
// target_wg: the workgroups a launch aims at (0: the size rule's 2048);
// sub-ranges are cut only while the launch holds fewer
AdjPlan adjoint_plan(int n_rep, int max_beads, size_t bead_stride, int target_wg = 0);
hipError_t launch_adjoint(const KParams &p, const void *adj_ff, const AdjPlan &a, int rep_first, int n_rep,
                          int max_bodies, int bead_base, size_t bead_stride, int body_base, double *scratch,
                          double *force_torque, double *dparams, hipStream_t s);
hipError_t launch_drms(const KParams &p, int sys_id, int n_beads, int n_rep, const float *rx, const float *ry,
                       const float *rz, double *partials, double inv_pairs, double *out, hipStream_t s);
hipError_t launch_centre_hist(const KParams &p, int sys_id, int n_bodies, int n_rep, const int *btype, const int *site,
                              int n_types, int nbin, double lo, double dr, unsigned long long *counts, hipStream_t s);
hipError_t launch_gr_normalise(const unsigned long long *counts, const int *type_count, int n_types, int nbin, double lo,
                               double dr, double frames_over_volume, double *gr, hipStream_t s);
size_t adjoint_ff_size();
void adjoint_ff_fill(void *dst, const float *table, double lambda, double e0, double xi, double c_lj, double c_dh);

constexpr int FULL_TILE = 256;
// full-energy work units per replica of a system with `tiles` bead tiles:
// row tile x offset 0..tiles/2 of the half-window tile schedule
__host__ __device__ inline int full_units(int tiles) { return tiles * (tiles / 2 + 1); }
// largest number of partner sub-ranges a full-energy unit is cut into
constexpr int FULL_SPLIT_MAX = 4;
constexpr int SPEC_MAX = 16;     // largest body (beads) a speculative window slot takes
static_assert(SPEC_MAX > 1, "a window slot's thread 1 forms its rotation");
constexpr int REMC_MAX_COOP = 8; // workgroups per replica in the cooperative MC kernel
constexpr int REMC_SLOT_K = 64;  // partial sums per member and exchange (MC window slots, force-biased window values)
// two banks of REMC_SLOT_K x REMC_MAX_COOP tagged partial sums per replica
constexpr int REMC_SLOTS_PER_REP = 2 * REMC_SLOT_K * REMC_MAX_COOP;
static_assert(REMC_SLOT_K >= 64, "every lane of a polling wave reads a slot row of its own");

} // namespace remc
