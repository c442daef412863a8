/*
 * remc.h -- C ABI of the MI355X replica-exchange Monte Carlo hot path.
 *
 * The reference drives its GPU through the extern "C" block of
 * "simulation engine/src/cudaExterns.h" (lines 11-78) plus the device-mirror
 * methods of Replica (Replica.cpp:1108-1121, 1308-1417, 1530-1576, 1601-1655).
 * This header is the replacement boundary: plain pointers and sizes, no GPU
 * runtime types, every call returns an int status (0 = OK, < 0 = error; the
 * text is in remc_last_error()).  The reference aborts through cutilCheckMsg
 * instead; nothing here calls exit().
 *
 * Ownership: the library owns all device memory.  Host arrays passed in are
 * copied before the call returns.  A context is bound to one device, uses one
 * stream and is not thread-safe (callers serialise), as the reference's
 * one-host-thread-per-GPU model assumes (main.cpp:899-906).  Unlike the
 * reference's per-context globals (const_boxdim, LJTexture), box size and force
 * field constants are per context/system.
 *
 * Units: positions in Angstrom, energies in kcal/mol, temperatures in K.
 */
#ifndef REMC_H
#define REMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REMC_NTYPES 20
#define REMC_MAX_BODY_BEADS 1024 /* largest rigid body the MC kernel stages in LDS */

/* status codes */
#define REMC_OK 0
#define REMC_ERR_ARG (-1)
#define REMC_ERR_STATE (-2)
#define REMC_ERR_HIP (-3)
#define REMC_ERR_ALLOC (-4)
#define REMC_ERR_LIMIT (-5)
#define REMC_ERR_DEVICE (-6)

#if defined(__GNUC__)
#define REMC_API __attribute__((visibility("default")))
#else
#define REMC_API
#endif

typedef struct remc_ctx remc_ctx;

/* Force field.  E_ij = lambda*(table_ij - e0) in kT units; contact term
 * LJ = -4 E_ij x6 (x6 - 1), x6 = ((sigma_i+sigma_j)/2r)^6, flipped to
 * -LJ + 2E_ij when E_ij > 0 and r < 2^(1/6) sigma_ij; Debye-Hueckel
 * DH = q_i q_j exp(-r/xi)/r; crowder pairs (6/r)^6 below 15 A;
 * E = (sum LJ * lj_cf + sum DH * dh_cf) * kbt_cf.
 * Replaces copyLJPotentialDataToDevice/bindLJTexture (cudaExterns.h:14,44)
 * and the compile-time constants of definitions.h:11-35,61,89. */
typedef struct remc_forcefield {
    float table[REMC_NTYPES * REMC_NTYPES]; /* symmetric, row-major */
    double lambda, e0, xi;
    double lj_cf, dh_cf, kbt_cf;
    double crowder_sigma, crowder_cutoff;
} remc_forcefield;

/* One system of rigid bodies.  Beads are ordered by body; body b owns beads
 * [body_start[b], body_start[b+1]).  The conformation given here is the
 * initial state of every replica added for the system.  Replaces the packing
 * of ReplicaDataToDevice (Replica.cpp:1308-1417): where the reference builds
 * AoS float4 pos=(x,y,z,molecule) and meta=(aa,q,vdW,role) padded to a block
 * multiple, this layout is SoA and unpadded. */
typedef struct remc_system_desc {
    int n_beads, n_bodies;
    const int *body_start;        /* n_bodies + 1 */
    const int *type;              /* n_beads, 0..19 (ignored for crowders) */
    const float *charge;          /* n_beads */
    const float *sigma;           /* n_beads, vdW radius */
    const unsigned char *crowder; /* n_bodies, 1 = repulsive crowder */
    float box;                    /* periodic edge (CUDA_setBoxDimension); <= 0: none */
    const float *x, *y, *z;       /* bead positions */
    const float *rx, *ry, *rz;    /* positions relative to the body centre */
    const float *cx, *cy, *cz;    /* body centres (n_bodies) */
    const double *quat;           /* 4*n_bodies, (w,x,y,z) per body */
} remc_system_desc;

typedef struct remc_counters {
    uint64_t accept;       /* dE < 0            (Replica::accept)  */
    uint64_t accept_boltz; /* Boltzmann accepts (Replica::acceptA) */
    uint64_t reject;       /* rejections        (Replica::reject)  */
    uint64_t steps;        /* MC steps taken = RNG step counter    */
} remc_counters;

typedef struct remc_step_record {
    int32_t body, move, accepted; /* move 1 = rotate; accepted 0/1/2 */
    int32_t ticks;                /* device wall-clock ticks (100 MHz) the step took */
    double delta;                      /* dE, kcal/mol */
} remc_step_record;

/* ------------------------------------------------------------- context --- */
/* Replaces cudaSetDevice + per-thread init in MCthreadableFunction (main.cpp:899-940). */
REMC_API int remc_ctx_create(int device, remc_ctx **out);
/* Replaces Replica::FreeDevice (Replica.cpp:1618-1655) and unbindLJTexture. */
REMC_API int remc_ctx_destroy(remc_ctx *ctx);
REMC_API const char *remc_last_error(void);
REMC_API int remc_set_forcefield(remc_ctx *ctx, const remc_forcefield *ff);
/* Seed of the counter-based RNG (Philox4x32-10).  Replaces Replica::initRNGs
 * (Replica.cpp:265-290), whose GSL streams are time-seeded. */
REMC_API int remc_set_seed(remc_ctx *ctx, uint64_t seed);

/* Adds a system; *sys_id receives its index.  Must precede the first compute call. */
REMC_API int remc_system_add(remc_ctx *ctx, const remc_system_desc *desc, int *sys_id);
/* Adds n_rep replicas of a system (Replica::copy, REMCSimulation main.cpp:1161-1188).
 * temperature[k] is both replica k's start temperature and the k-th rung of the
 * system's exchange ladder (ascending).  Step sizes stay with the replica, as in
 * the reference (exchangeReplicas swaps temperatures only).  *first_replica
 * receives the global index of the first new replica. */
REMC_API int remc_replicas_add(remc_ctx *ctx, int sys_id, int n_rep, const double *temperature,
                      const float *translate_step, const double *rotate_step, int *first_replica);
/* Number of replicas/beads/bodies known to the context. */
REMC_API int remc_get_sizes(remc_ctx *ctx, int *n_systems, int *n_replicas, int *max_beads);

/* -------------------------------------------------------------- energy --- */
/* Plain per-replica energy: the reference-shaped tiled kernel (block x block
 * tiles, lower triangle culled, FP32 partials, host FP64 sum) on the current
 * device conformation of one replica.  Replaces CUDA_EonDevice
 * (cudaExterns.h:23; CudaFunctions.cu:304-385) / Replica::EonDevice
 * (Replica.cpp:1601-1616).  block = 0 picks the reference's rule (32 if
 * N < 1024 else 64, main.cpp:1746-1747). */
REMC_API int remc_energy_plain(remc_ctx *ctx, int replica, int block, double *out);
/* Batched full energy of every replica in one launch: FP32 pair terms with
 * Kahan sums per thread, FP64 across threads and tiles.  out[n_replicas].
 * Same quantity as Replica::E() (Replica.cpp:603-745). */
REMC_API int remc_energy_all(remc_ctx *ctx, double *out);
/* remc_energy_all without any allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Its kernels' partials live
 * in a static 4 MiB device array per device instead of the context's buffers:
 * at most 524,288 doubles, of which it needs R (4 U + 1) for R replicas,
 * U = G (G/2 + 1) and G = ceil(N/256) for the largest system's bead count N
 * (config 5, 64 replicas of 11,000 beads: 242,240); larger contexts get
 * REMC_ERR_LIMIT.
 * The context must already be set up by an earlier call (any energy, state or
 * MC call), else REMC_ERR_STATE.  Every context on one device shares the
 * array: calls that overlap on one device race. */
REMC_API int remc_energy_all_noalloc(remc_ctx *ctx, double *out);
/* Sets every replica's tracked energy to remc_energy_all's value. */
REMC_API int remc_refresh_energies(remc_ctx *ctx);
/* Energy of pairs between non-crowder bodies only, all replicas; the
 * extra term of Replica::sample (Replica.cpp:1723-1743). */
REMC_API int remc_energy_noncrowder_all(remc_ctx *ctx, double *out);

/* ------------------------------------------------------------------ MC --- */
/* nsteps MC steps on every replica in one persistent launch (one or more
 * cooperating workgroups per replica, remc_set_mc_coop): device-side move
 * proposal, dE of the moved body, Metropolis in FP64.  Replaces the MCSearch loop (Replica.cpp:368-478) with its per-step
 * MoleculeDataToDevice + EonDevice round trips. */
REMC_API int remc_mc_steps(remc_ctx *ctx, int nsteps);
/* The reference protocol for one replica: host proposal, host->device copy
 * of the moved body, remc_energy_plain, device->host partials, host
 * Metropolis, copy-back on reject (Replica.cpp:386-474).  Uses the same RNG
 * stream and move arithmetic as remc_mc_steps. */
REMC_API int remc_mc_steps_plain(remc_ctx *ctx, int replica, int nsteps, int block);
/* nsteps force-biased (smart) MC steps on every replica (Rossky, Doll &
 * Friedman, J. Chem. Phys. 69, 4628 (1978)), one persistent launch: the
 * chosen body's translation is drawn as d = beta A F + sqrt(2A) n, or its
 * rotation vector as beta A tau + sqrt(2A) n, from the force F on the body or
 * the torque tau about its centre (n standard normal; A = t^2/6 for
 * translations, r^2/6 for rotations, t and r the replica's translate and
 * rotate steps), and accepted with the proposal ratio included, so the chain
 * keeps exp(-beta E).  No counterpart in the reference: it replaces the
 * uniform proposal of Replica::MCSearch (Replica.cpp:386-420) with one driven
 * by the gradient that remc_adjoint returns.  Shares the replicas, energies,
 * step and decision counters with remc_mc_steps.  A replica runs on the
 * workgroups per replica set by remc_set_mc_coop (or the MC autotuner), else
 * on as many as give each its own CU, at most what runs co-resident; a large
 * move is split over them.  Runs of one-bead crowder moves are taken in
 * windows (one per wave) when the replica fits in LDS, switched with
 * remc_set_mc_window; decisions and states equal those of one step at a time
 * up to rounding.  block: workgroup size (256, 512 or 1024; 0 = the size an
 * autotuner picks on first use: 128 steps at each size on the live state,
 * which is then restored bit for bit). */
REMC_API int remc_mc_steps_fb(remc_ctx *ctx, int nsteps, int block);
/* One replica-exchange round over every system's temperature ladder, pairs
 * (k,k+1) for k = offset, offset+2, ...; the offset alternates per call
 * starting at 0 (main.cpp:1427-1453).  *n_swaps (may be NULL) receives the
 * number of swaps. */
REMC_API int remc_exchange(remc_ctx *ctx, int *n_swaps);
/* rounds x (steps_per_round MC steps + one exchange round), captured once in a
 * hipGraph and replayed.  Replaces the REMCSimulation/MCthreadableFunction
 * lock-step loop (main.cpp:971-1012, 1378-1469). */
REMC_API int remc_run(remc_ctx *ctx, int rounds, int steps_per_round);
/* enable = 1: every round remc_run captures re-evaluates every replica's
 * energy (as remc_refresh_energies) between the MC launch and the exchange,
 * so the exchanges read full energies as the reference's do
 * (main.cpp:1427-1453: its potential is re-evaluated every step).  Default 0:
 * the exchanges read the tracked energies, the FP64 sums of the accepted
 * steps' FP32 dE, which differ from the full evaluation by the pair terms'
 * rounding (about 1e-6 kT after a 20,000-step run of config 4).  No
 * counterpart in the reference. */
REMC_API int remc_set_run_refresh(remc_ctx *ctx, int enable);
/* Records per-step decisions for the first max_steps steps after this call
 * (0 disables). */
REMC_API int remc_set_recording(remc_ctx *ctx, int max_steps);
REMC_API int remc_get_records(remc_ctx *ctx, int replica, remc_step_record *out, int n);

/* --------------------------------------------------------------- state --- */
REMC_API int remc_get_energies(remc_ctx *ctx, double *out);
REMC_API int remc_get_temperatures(remc_ctx *ctx, double *out);
REMC_API int remc_get_counters(remc_ctx *ctx, remc_counters *out);
/* slot_replica[k] for system sys: replica holding the k-th ladder rung. */
REMC_API int remc_get_ladder(remc_ctx *ctx, int sys_id, int *slot_replica);
/* Conformation of one replica: bead arrays of the system's size, body arrays of
 * its body count, quat 4 per body.  Any pointer may be NULL on get. */
REMC_API int remc_get_state(remc_ctx *ctx, int replica, float *x, float *y, float *z,
                   float *rx, float *ry, float *rz, float *cx, float *cy, float *cz, double *quat);
REMC_API int remc_set_state(remc_ctx *ctx, int replica, const float *x, const float *y, const float *z,
                   const float *rx, const float *ry, const float *rz,
                   const float *cx, const float *cy, const float *cz, const double *quat);

/* ---------------------------------------------------------- checkpoint --- */
/* What a checkpoint holds beyond remc_get_state/remc_set_state: the tracked
 * energies, the accept counters with each replica's RNG step counter, every
 * system's ladder and exchange state (round counter, offset, swap total).
 * A fresh context given the saved conformations and these continues the saved
 * run bit for bit.  Together they replace saveSimulation and the resume path
 * (main.cpp:89-122, 336-372), whose load never restores a run. */
REMC_API int remc_set_energies(remc_ctx *ctx, const double *energies);
REMC_API int remc_set_counters(remc_ctx *ctx, const remc_counters *in);
/* slot_replica[k] = replica holding rung k; temperatures follow the rungs. */
REMC_API int remc_set_ladder(remc_ctx *ctx, int sys_id, const int *slot_replica);
REMC_API int remc_get_exchange_state(remc_ctx *ctx, int sys_id, uint64_t *round, int *offset, uint64_t *swaps);
REMC_API int remc_set_exchange_state(remc_ctx *ctx, int sys_id, uint64_t round, int offset, uint64_t swaps);

/* ------------------------------------------------------------ analysis --- */
/* Distance RMS of every replica of system sys_id against a reference
 * conformation of its beads (ref_x/y/z, n_beads each): the mean over pairs of
 * beads in different non-crowder bodies of | |r_i - r_j| - |r_i - r_j|_ref |
 * (plain Euclidean distances in FP32, sums in FP64), as DRMS over the
 * non-crowder molecules ("post processor/pdbgensrc/drms.cpp:7-58").
 * out[k] for the system's k-th replica. */
REMC_API int remc_drms(remc_ctx *ctx, int sys_id, const float *ref_x, const float *ref_y, const float *ref_z,
                       double *out);
/* Histogram of body-centre distances (minimum image, FP64) over every
 * replica of system sys_id, by body type ("post processor/gr.py:93-140"):
 * body_type[m] in [0, n_types), or -1 to leave body m out; bins of width
 * (hi - lo) / nbin from lo.  counts[(ti * n_types + tj) * nbin + bin] for
 * ti <= tj: a pair of one type counts from both bodies, a pair of two types
 * once.  At most 12288 bins in all (n_types^2 x nbin). */
REMC_API int remc_centre_histogram(remc_ctx *ctx, int sys_id, const int *body_type, int n_types, int nbin, double lo,
                                   double hi, uint64_t *counts);

/* Radial distribution g(r) of body pairs by type over every replica of
 * system sys_id, each replica one frame, as "post processor/gr.py:93-155"
 * computes it from PDB frames: the histogram of remc_centre_histogram (same
 * type and bin conventions; a body's point is the position of its bead
 * site_bead[m], which gr.py takes as the residue nearest the chain's centroid,
 * or its centre when site_bead is NULL), divided by
 * n_frames 4 pi r^2 dr n_ti n_tj / V with r the bin centre, n_t the number of
 * bodies of type t and V = box^3 (the system must be periodic).
 * gr[(ti * n_types + tj) * nbin + bin] for ti <= tj (0 below the diagonal). */
REMC_API int remc_radial_distribution(remc_ctx *ctx, int sys_id, const int *body_type, const int *site_bead,
                                      int n_types, int nbin, double lo, double hi, double *gr);

/* ------------------------------------------------------------- adjoint --- */
/* Per-body force and torque about the body centre (6 doubles per body:
 * Fx Fy Fz Tx Ty Tz, kcal/mol/A and kcal/mol/rad, F = -dE/dx) and the
 * gradient of E with respect to (lambda, e0, xi) for one replica.  No
 * counterpart in the reference (its gradient-free MC has none). */
REMC_API int remc_adjoint(remc_ctx *ctx, int replica, double *force_torque, double *dparams);
/* The same for every replica in one launch: force_torque holds 6 doubles per
 * body of each replica's system, replicas in order (6 x the total body count
 * over replicas); dparams 3 per replica.  Either output may be NULL, not both.
 * No counterpart in the reference. */
REMC_API int remc_adjoint_all(remc_ctx *ctx, double *force_torque, double *dparams);

/* -------------------------------------------------------------- timing --- */
/* Kernel timing with HIP events recorded on the context's stream around every
 * launch of a kernel class: 0 = MC, 1 = full energy, 2 = plain energy,
 * 3 = exchange, 4 = adjoint.  enable resets the accumulators. */
REMC_API int remc_timing_enable(remc_ctx *ctx, int enable);
REMC_API int remc_timing_get(remc_ctx *ctx, int kernel_class, double *total_ms, int64_t *launches);
/* Blocks until the context's stream is idle. */
REMC_API int remc_synchronize(remc_ctx *ctx);
/* Selects the MC kernel's workgroup size (0 = autotune on first use; else
 * 256, 512 or 1024). */
REMC_API int remc_set_mc_block(remc_ctx *ctx, int block);
REMC_API int remc_get_mc_block(remc_ctx *ctx, int *block);
/* The force-biased kernel's autotuned workgroup size (0 until remc_mc_steps_fb
 * has been called with block 0). */
REMC_API int remc_get_fb_block(remc_ctx *ctx, int *block);
/* Workgroups per replica of the MC kernel (0 = autotune; else 1, 2, 4 or 8).
 * With more than one, each workgroup keeps a private copy of its replica's
 * state, evaluates dE over a slice of the partner beads, and the members add
 * their partials through device memory (once per split step, once per
 * speculative window); the grid must fit co-resident on the device
 * (checked). */
REMC_API int remc_set_mc_coop(remc_ctx *ctx, int coop);
REMC_API int remc_get_mc_coop(remc_ctx *ctx, int *coop);
/* MC kernel tuning: lds_cache = 0 keeps partner beads in HBM instead of the
 * LDS cache; split_min_pairs (> 0) is the pair-term count from which a move is
 * split across cooperating workgroups (default 8192; 0 keeps the current). */
REMC_API int remc_set_mc_tuning(remc_ctx *ctx, int lds_cache, int split_min_pairs);
/* Speculative windows of the MC kernels (default on): up to two steps per wave
 * (32 with 1024-thread workgroups) of consecutive moves of small bodies (at
 * most 16 beads) are evaluated at once against the window's starting
 * conformation (with cooperating workgroups, each over its partner slice); the
 * window ends before the first step that an accepted earlier step of the
 * window interacts with (exact pair corrections between the moved bodies) or
 * whose body it moved.  A step whose body exactly one earlier step of the
 * window moved is proposed from that step's proposal instead, and ends the
 * window only if that step is rejected.  Used when a partner cache fits in
 * LDS (the replica, or each member's slice).  Decisions and states equal those
 * of one step at a time up to FP32 rounding.  The force-biased kernel's windows
 * of one-bead crowder moves follow the same switch.  0 turns the windows off.
 * No counterpart in the reference. */
REMC_API int remc_set_mc_window(remc_ctx *ctx, int enable);

#ifdef __cplusplus
}
#endif
#endif
