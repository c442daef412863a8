// remc_api.hip -- implementation of the C ABI declared in include/remc.h.
//
// Host-side responsibilities: validate inputs, lay systems and replicas out in
// HBM (see remc_internal.h), launch kernels on the context's stream, run the
// reference's per-step protocol for the plain path, capture MC+exchange rounds
// in a hipGraph, and time launches with HIP events.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "remc.h"
#include "remc_internal.h"
#include "remc_math.h"

using namespace remc;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return fail(REMC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

struct HostSys {
    int n_beads, n_bodies;
    std::vector<int> body_start, type, crowder;
    std::vector<float> charge, sigma;
    float box;
    std::vector<float> x, y, z, rx, ry, rz, cx, cy, cz;
    std::vector<double> quat;
    std::vector<int> replicas;    // global replica ids, ladder order
    std::vector<double> ladder;   // ascending temperatures
    int max_body;
};

struct HostRep {
    int sys;
    double T;
    float tstep;
    double rstep;
};

template <typename T>
hipError_t dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **)p, n * sizeof(T));
}

template <typename T>
hipError_t upload(T *d, const std::vector<T> &h, hipStream_t s)
{
    if (h.empty()) return hipSuccess;
    return hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

} // namespace

// Scratch of remc_energy_all_noalloc: the full-energy kernels' FP64 unit
// partials and per-replica totals, in device memory the library never
// allocates (4 MiB; each device the library runs on holds its copy until the
// process ends).  In host code `scratch` is a stand-in for the symbol, not a
// device address: its address on a device comes from hipGetSymbolAddress.
constexpr size_t REMC_NOALLOC_DOUBLES = 524288;
static __device__ __align__(256) double scratch[REMC_NOALLOC_DOUBLES];

namespace {

constexpr int REMC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[REMC_MAX_DEVICES] = {};

// The device address of `scratch` on `device`, looked up once per device with
// that device current (the caller has made it current).
hipError_t scratch_address(int device, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= REMC_MAX_DEVICES) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = (double *)ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

} // namespace

struct remc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    remc_forcefield ff{};
    bool ff_set = false;
    uint64_t seed = 0x52454D43ull;
    std::vector<HostSys> systems;
    std::vector<HostRep> reps;
    bool finalized = false;

    // sizes
    int total_static_beads = 0, total_static_bodies = 0;
    int total_rep_beads = 0, total_rep_bodies = 0, total_slots = 0;
    int max_beads = 0, max_tiles = 0;
    std::vector<int> rep_bead_off, rep_body_off;

    // device buffers
    int *d_tc = nullptr, *d_body = nullptr, *d_bstart = nullptr, *d_bend = nullptr, *d_bcrow = nullptr;
    float *d_q = nullptr, *d_sig = nullptr, *d_etab = nullptr, *d_sstab = nullptr;
    int sigtab = 0;
    DevSys *d_sys = nullptr;
    DevRep *d_rep = nullptr;
    float *d_px = nullptr, *d_py = nullptr, *d_pz = nullptr, *d_rx = nullptr, *d_ry = nullptr, *d_rz = nullptr;
    float *d_cx = nullptr, *d_cy = nullptr, *d_cz = nullptr;
    double *d_quat = nullptr, *d_E = nullptr, *d_T = nullptr;
    unsigned long long *d_step = nullptr, *d_cnt = nullptr;
    int *d_slot = nullptr, *d_exoff = nullptr;
    double *d_ladder = nullptr;
    unsigned long long *d_exround = nullptr, *d_exswaps = nullptr;
    double *d_full_partials = nullptr, *d_out = nullptr;
    float *d_plain_partials = nullptr;
    size_t plain_partials_cap = 0;
    remc_step_record *d_rec = nullptr;
    unsigned long long *d_rec_start = nullptr;
    int rec_max = 0;
    unsigned long long *d_slots = nullptr;
    int *d_err = nullptr;
    double *d_adj = nullptr; // adjoint scratch: bead gradients, tile partials, body forces, parameter gradients
    size_t adj_cap = 0;      // doubles held by d_adj
    int adj_target = 0;      // all-replica adjoint's workgroup target (0: tune on first use)
    void *d_ana = nullptr;   // analysis scratch (reference conformation, partials, histogram)
    size_t ana_bytes = 0;
    DevFF devff{};
    double *noalloc_scratch = nullptr; // this device's `scratch` (remc_energy_all_noalloc)
    int full_split = 0;      // full-energy partner sub-ranges per unit (0: the size rule; set at set-up)
    int fb_block = 0;        // force-biased kernel's workgroup size (0 = autotune on first use)
    bool run_refresh = false; // remc_run: energy refresh between the MC launch and the exchange

    // MC configuration: workgroup size and workgroups per replica (0 = autotune)
    int mc_block = 0;
    int mc_coop = 0;
    int coop_min_pairs = 8192; // moves below this many pair terms are evaluated whole by every member
    int synced_copies = 1;   // state copies [0, synced_copies) hold the current conformation
    int n_cu = 0;
    int lds_per_block = 0, lds_per_cu = 0;
    bool lds_cache = true;   // partner-bead LDS cache of the MC kernel allowed
    bool mc_window = true;   // speculative windows of small moves allowed
    long long min_move_pairs = -1;
    int max_bodies = 0, max_body_beads = 0;

    // graph
    hipGraphExec_t graph = nullptr;
    int graph_steps = -1;

    // timing
    bool timing = false;
    struct EvPair {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<EvPair> pending;
    std::vector<hipEvent_t> free_events;
    double t_total[5] = {0, 0, 0, 0, 0};
    int64_t t_count[5] = {0, 0, 0, 0, 0};

    KParams params() const
    {
        KParams p{};
        p.tc = d_tc; p.q = d_q; p.sig = d_sig; p.body = d_body;
        p.bstart = d_bstart; p.bend = d_bend; p.bcrow = d_bcrow;
        p.sys = d_sys; p.rep = d_rep; p.etab = d_etab; p.ff = devff;
        p.sstab = d_sstab; p.sigtab = sigtab;
        p.px = d_px; p.py = d_py; p.pz = d_pz; p.rx = d_rx; p.ry = d_ry; p.rz = d_rz;
        p.cx = d_cx; p.cy = d_cy; p.cz = d_cz; p.quat = d_quat;
        p.E = d_E; p.T = d_T; p.step = d_step; p.cnt = d_cnt;
        p.slot_rep = d_slot; p.ladder = d_ladder; p.ex_round = d_exround; p.ex_offset = d_exoff;
        p.ex_swaps = d_exswaps;
        p.n_rep = (int)reps.size();
        p.coop_min_pairs = coop_min_pairs;
        p.copy_beads = (size_t)total_rep_beads;
        p.copy_bodies = (size_t)total_rep_bodies;
        p.slots = d_slots;
        p.err = d_err;
        p.rec = rec_max > 0 ? d_rec : nullptr;
        p.rec_start = d_rec_start;
        p.rec_max = rec_max;
        p.key0 = (uint32_t)seed;
        p.key1 = (uint32_t)(seed >> 32);
        return p;
    }
};

namespace {

hipEvent_t get_event(remc_ctx *c)
{
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct TimedLaunch {
    remc_ctx *c;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(remc_ctx *c_, int cls_) : c(c_), cls(cls_)
    {
        if (!c->timing) return;
        a = get_event(c);
        b = get_event(c);
        // a launch that cannot be timed is still made; its events go back to the pool
        if (!a || !b || hipEventRecord(a, c->stream) != hipSuccess) release();
    }
    ~TimedLaunch()
    {
        if (!a) return;
        if (hipEventRecord(b, c->stream) == hipSuccess)
            c->pending.push_back({cls, a, b});
        else
            release();
    }
    void release()
    {
        if (a) c->free_events.push_back(a);
        if (b) c->free_events.push_back(b);
        a = b = nullptr;
    }
};

int collect_timing(remc_ctx *c)
{
    for (auto &pe : c->pending) {
        HIP_TRY(hipEventSynchronize(pe.b));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, pe.a, pe.b));
        c->t_total[pe.cls] += ms;
        c->t_count[pe.cls] += 1;
        c->free_events.push_back(pe.a);
        c->free_events.push_back(pe.b);
    }
    c->pending.clear();
    return REMC_OK;
}

void destroy_graph(remc_ctx *c)
{
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    c->graph_steps = -1;
}

// Two timing events, destroyed with the holder.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t create()
    {
        hipError_t e = hipEventCreate(&a);
        return e != hipSuccess ? e : hipEventCreate(&b);
    }
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Partner split of the full-energy kernels for each (replicas, bead tiles):
// timed once per process and grid size, so every context of a process with
// that size takes the same split and their energies agree bit for bit.
std::mutex g_split_mu;
std::map<std::pair<int, int>, int> g_split_cache;

int autotune_energy(remc_ctx *c)
{
    const int R = (int)c->reps.size();
    const std::pair<int, int> key(R, c->max_tiles);
    {
        std::lock_guard<std::mutex> lock(g_split_mu);
        auto it = g_split_cache.find(key);
        if (it != g_split_cache.end()) {
            c->full_split = it->second;
            return REMC_OK;
        }
    }
    // each candidate on the live conformations (energies only, nothing
    // written back), best of three; the size rule's choice is kept unless
    // another split is at least 3% faster
    const int def = full_split_default(R, c->max_tiles);
    const int cand[3] = {1, 2, FULL_SPLIT_MAX};
    float best_ms[3] = {0.f, 0.f, 0.f};
    EventPair ev;
    HIP_TRY(ev.create());
    const KParams p = c->params();
    for (int k = 0; k < 3; ++k) {
        HIP_TRY(launch_energy_full(p, R, c->max_tiles, c->d_full_partials, c->d_out, 0, 0, cand[k], c->stream));
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            HIP_TRY(hipEventRecord(ev.a, c->stream));
            HIP_TRY(launch_energy_full(p, R, c->max_tiles, c->d_full_partials, c->d_out, 0, 0, cand[k], c->stream));
            HIP_TRY(hipEventRecord(ev.b, c->stream));
            HIP_TRY(hipEventSynchronize(ev.b));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
            best = std::min(best, ms);
        }
        best_ms[k] = best;
    }
    float t_def = 1e30f;
    for (int k = 0; k < 3; ++k)
        if (cand[k] == def) t_def = best_ms[k];
    int pick = def;
    float t_pick = t_def;
    for (int k = 0; k < 3; ++k)
        if (best_ms[k] < 0.97f * t_def && best_ms[k] < t_pick) {
            pick = cand[k];
            t_pick = best_ms[k];
        }
    std::lock_guard<std::mutex> lock(g_split_mu);
    auto ins = g_split_cache.emplace(key, pick); // a context set up meanwhile keeps its choice
    c->full_split = ins.first->second;
    return REMC_OK;
}

int finalize(remc_ctx *c)
{
    if (c->finalized) return REMC_OK;
    if (!c->ff_set) return fail(REMC_ERR_STATE, "force field not set");
    if (c->systems.empty()) return fail(REMC_ERR_STATE, "no systems added");
    if (c->reps.empty()) return fail(REMC_ERR_STATE, "no replicas added");
    HIP_TRY(hipSetDevice(c->device));

    // static layout
    std::vector<int> tc, body, bstart, bend, bcrow;
    std::vector<float> q, sig;
    std::vector<DevSys> dsys;
    int bead_off = 0, body_off = 0, slot_off = 0;
    for (auto &s : c->systems) {
        DevSys d{};
        d.bead_off = bead_off;
        d.n_beads = s.n_beads;
        d.body_off = body_off;
        d.n_bodies = s.n_bodies;
        d.max_body = s.max_body;
        d.rep_first = s.replicas.empty() ? 0 : s.replicas.front();
        d.n_rep = (int)s.replicas.size();
        d.slot_off = slot_off;
        d.box = s.box > 0.f ? s.box : 0.f;
        d.inv_box = s.box > 0.f ? 1.0f / s.box : 0.f;
        dsys.push_back(d);
        for (int m = 0; m < s.n_bodies; ++m) {
            bstart.push_back(s.body_start[m]);
            bend.push_back(s.body_start[m + 1]);
            bcrow.push_back(s.crowder[m]);
            for (int b = s.body_start[m]; b < s.body_start[m + 1]; ++b) {
                tc.push_back((s.type[b] & 0xff) | (s.crowder[m] ? 0x100 : 0));
                body.push_back(m);
                q.push_back(s.charge[b]);
                sig.push_back(s.sigma[b]);
            }
        }
        bead_off += s.n_beads;
        body_off += s.n_bodies;
        slot_off += (int)s.replicas.size();
        c->max_beads = std::max(c->max_beads, s.n_beads);
    }
    c->total_static_beads = bead_off;
    c->total_static_bodies = body_off;
    c->total_slots = slot_off;

    // replica layout
    const int R = (int)c->reps.size();
    std::vector<DevRep> drep(R);
    c->rep_bead_off.assign(R, 0);
    c->rep_body_off.assign(R, 0);
    std::vector<float> px, py, pz, rx, ry, rz, cx, cy, cz;
    std::vector<double> quat, E(R, 0.0), T(R);
    int rbo = 0, rmo = 0;
    for (int r = 0; r < R; ++r) {
        const HostRep &h = c->reps[r];
        const HostSys &s = c->systems[h.sys];
        DevRep d{};
        d.sys = h.sys;
        d.bead_off = rbo;
        d.body_off = rmo;
        d.rng_id = (uint32_t)r;
        d.tstep = h.tstep;
        d.rot_ch = std::cos(h.rstep / 2.0);
        d.rot_sh = std::sin(h.rstep / 2.0);
        d.rstep = h.rstep;
        drep[r] = d;
        c->rep_bead_off[r] = rbo;
        c->rep_body_off[r] = rmo;
        px.insert(px.end(), s.x.begin(), s.x.end());
        py.insert(py.end(), s.y.begin(), s.y.end());
        pz.insert(pz.end(), s.z.begin(), s.z.end());
        rx.insert(rx.end(), s.rx.begin(), s.rx.end());
        ry.insert(ry.end(), s.ry.begin(), s.ry.end());
        rz.insert(rz.end(), s.rz.begin(), s.rz.end());
        cx.insert(cx.end(), s.cx.begin(), s.cx.end());
        cy.insert(cy.end(), s.cy.begin(), s.cy.end());
        cz.insert(cz.end(), s.cz.begin(), s.cz.end());
        quat.insert(quat.end(), s.quat.begin(), s.quat.end());
        T[r] = h.T;
        rbo += s.n_beads;
        rmo += s.n_bodies;
    }
    c->total_rep_beads = rbo;
    c->total_rep_bodies = rmo;
    std::vector<int> slots;
    std::vector<double> ladder;
    for (auto &s : c->systems) {
        slots.insert(slots.end(), s.replicas.begin(), s.replicas.end());
        ladder.insert(ladder.end(), s.ladder.begin(), s.ladder.end());
    }

    // force field
    std::vector<float> etab(REMC_NTYPES * REMC_NTYPES);
    const remc_forcefield &f = c->ff;
    const double lj_scale = f.lj_cf * f.kbt_cf;
    for (int i = 0; i < REMC_NTYPES * REMC_NTYPES; ++i)
        etab[i] = 4.0f * (float)(f.lambda * ((double)f.table[i] - f.e0) * lj_scale); // 4 E_ij (exact scaling)
    c->devff.c_crow = (float)lj_scale;
    c->devff.crow_s2 = (float)(f.crowder_sigma * f.crowder_sigma);
    c->devff.crow_cut2 = (float)(f.crowder_cutoff * f.crowder_cutoff);
    c->devff.crow_cut = (float)f.crowder_cutoff;
    c->devff.c_dh = (float)(f.dh_cf * f.kbt_cf);
    c->devff.k_dh = (float)(-1.4426950408889634 / f.xi);
    c->devff.eps_r = 1e-38f;
    {
        const double s2 = f.crowder_sigma * f.crowder_sigma;
        c->devff.crow_k = (float)(lj_scale * s2 * s2 * s2);
    }

    c->max_tiles = (c->max_beads + FULL_TILE - 1) / FULL_TILE;

    // When every non-crowder bead's sigma is its residue type's (as the
    // reference's amino-acid table makes it), the MC kernel reads each type
    // pair's (0.5 (sigma_i + sigma_j))^2 / box^2 from a table per system: the
    // FP32 operations the kernel would form from the two sigmas, in the same
    // order, so the numbers are the same.
    {
        std::vector<float> sstab;
        bool by_type = true;
        for (const auto &hs : c->systems) {
            float st[REMC_NTYPES];
            bool seen[REMC_NTYPES] = {};
            for (int m = 0; m < hs.n_bodies && by_type; ++m) {
                if (hs.crowder[m]) continue;
                for (int b = hs.body_start[m]; b < hs.body_start[m + 1]; ++b) {
                    const int t = hs.type[b] & 0xff;
                    if (!seen[t]) {
                        seen[t] = true;
                        st[t] = hs.sigma[b];
                    } else if (!(st[t] == hs.sigma[b])) {
                        by_type = false;
                        break;
                    }
                }
            }
            if (!by_type) break;
            const float ib = hs.box > 0.f ? 1.0f / hs.box : 0.f;
            const float ib2 = ib * ib;
            for (int i = 0; i < REMC_NTYPES; ++i)
                for (int j = 0; j < REMC_NTYPES; ++j) {
                    float v = 0.f;
                    if (seen[i] && seen[j]) {
                        const float h = 0.5f * st[i] + 0.5f * st[j];
                        v = (h * h) * ib2;
                    }
                    sstab.push_back(v);
                }
        }
        c->sigtab = by_type ? 1 : 0;
        if (by_type) {
            HIP_TRY(dalloc(&c->d_sstab, sstab.size()));
            HIP_TRY(upload(c->d_sstab, sstab, c->stream));
        }
    }

    hipStream_t s = c->stream;
    HIP_TRY(dalloc(&c->d_tc, tc.size()));
    HIP_TRY(dalloc(&c->d_body, body.size()));
    HIP_TRY(dalloc(&c->d_q, q.size()));
    HIP_TRY(dalloc(&c->d_sig, sig.size()));
    HIP_TRY(dalloc(&c->d_bstart, bstart.size()));
    HIP_TRY(dalloc(&c->d_bend, bend.size()));
    HIP_TRY(dalloc(&c->d_bcrow, bcrow.size()));
    HIP_TRY(dalloc(&c->d_sys, dsys.size()));
    HIP_TRY(dalloc(&c->d_rep, drep.size()));
    HIP_TRY(dalloc(&c->d_etab, etab.size()));
    // REMC_MAX_COOP copies of the dynamic state; copy 0 is the canonical one
    HIP_TRY(dalloc(&c->d_px, px.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_py, py.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_pz, pz.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_rx, rx.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_ry, ry.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_rz, rz.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_cx, cx.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_cy, cy.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_cz, cz.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_quat, quat.size() * REMC_MAX_COOP));
    HIP_TRY(dalloc(&c->d_slots, (size_t)R * REMC_SLOTS_PER_REP));
    HIP_TRY(dalloc(&c->d_err, (size_t)1));
    HIP_TRY(dalloc(&c->d_E, (size_t)R));
    HIP_TRY(dalloc(&c->d_T, (size_t)R));
    HIP_TRY(dalloc(&c->d_step, (size_t)R));
    HIP_TRY(dalloc(&c->d_cnt, (size_t)3 * R));
    HIP_TRY(dalloc(&c->d_slot, slots.size()));
    HIP_TRY(dalloc(&c->d_ladder, ladder.size()));
    HIP_TRY(dalloc(&c->d_exround, c->systems.size()));
    HIP_TRY(dalloc(&c->d_exoff, c->systems.size()));
    HIP_TRY(dalloc(&c->d_exswaps, c->systems.size()));
    HIP_TRY(dalloc(&c->d_full_partials, (size_t)R * full_units(c->max_tiles) * FULL_SPLIT_MAX));
    HIP_TRY(dalloc(&c->d_out, (size_t)R));
    HIP_TRY(dalloc(&c->d_rec_start, (size_t)R));

    HIP_TRY(upload(c->d_tc, tc, s));
    HIP_TRY(upload(c->d_body, body, s));
    HIP_TRY(upload(c->d_q, q, s));
    HIP_TRY(upload(c->d_sig, sig, s));
    HIP_TRY(upload(c->d_bstart, bstart, s));
    HIP_TRY(upload(c->d_bend, bend, s));
    HIP_TRY(upload(c->d_bcrow, bcrow, s));
    HIP_TRY(upload(c->d_sys, dsys, s));
    HIP_TRY(upload(c->d_rep, drep, s));
    HIP_TRY(upload(c->d_etab, etab, s));
    HIP_TRY(upload(c->d_px, px, s));
    HIP_TRY(upload(c->d_py, py, s));
    HIP_TRY(upload(c->d_pz, pz, s));
    HIP_TRY(upload(c->d_rx, rx, s));
    HIP_TRY(upload(c->d_ry, ry, s));
    HIP_TRY(upload(c->d_rz, rz, s));
    HIP_TRY(upload(c->d_cx, cx, s));
    HIP_TRY(upload(c->d_cy, cy, s));
    HIP_TRY(upload(c->d_cz, cz, s));
    HIP_TRY(upload(c->d_quat, quat, s));
    HIP_TRY(upload(c->d_E, E, s));
    HIP_TRY(upload(c->d_T, T, s));
    HIP_TRY(upload(c->d_slot, slots, s));
    HIP_TRY(upload(c->d_ladder, ladder, s));
    std::vector<unsigned long long> zeros_r((size_t)3 * R, 0ull), zeros_s(c->systems.size(), 0ull);
    std::vector<int> zeros_si(c->systems.size(), 0);
    HIP_TRY(hipMemcpyAsync(c->d_step, zeros_r.data(), sizeof(unsigned long long) * R, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->d_cnt, zeros_r.data(), sizeof(unsigned long long) * 3 * R, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->d_rec_start, zeros_r.data(), sizeof(unsigned long long) * R, hipMemcpyHostToDevice, s));
    HIP_TRY(upload(c->d_exround, zeros_s, s));
    HIP_TRY(upload(c->d_exswaps, zeros_s, s));
    HIP_TRY(upload(c->d_exoff, zeros_si, s));
    {
        std::vector<unsigned long long> empty((size_t)R * REMC_SLOTS_PER_REP, ~0ull);
        HIP_TRY(upload(c->d_slots, empty, s));
        int zero = 0;
        HIP_TRY(hipMemcpyAsync(c->d_err, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    }
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, c->device));
    c->n_cu = prop.multiProcessorCount;
    c->lds_per_block = (int)prop.sharedMemPerBlock;
    c->lds_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
    if (c->lds_per_cu <= 0) c->lds_per_cu = c->lds_per_block;
    // move-size facts used to pick the LDS caches of the MC kernel
    c->min_move_pairs = -1;
    for (auto &sy : c->systems)
        for (int m = 0; m < sy.n_bodies; ++m) {
            const long long nm = sy.body_start[m + 1] - sy.body_start[m];
            const long long pairs = nm * (sy.n_beads - nm);
            if (c->min_move_pairs < 0 || pairs < c->min_move_pairs) c->min_move_pairs = pairs;
        }
    c->max_bodies = 0;
    c->max_body_beads = 0;
    for (auto &sy : c->systems) {
        c->max_bodies = std::max(c->max_bodies, sy.n_bodies);
        c->max_body_beads = std::max(c->max_body_beads, sy.max_body);
    }
    HIP_TRY(hipStreamSynchronize(s));
    c->synced_copies = 1;
    c->finalized = true;

    // initial energies of every replica
    KParams p = c->params();
    HIP_TRY(launch_energy_full(p, R, c->max_tiles, c->d_full_partials, nullptr, 0, 1, 0, s));
    HIP_TRY(hipStreamSynchronize(s));
    c->adj_target = 0; // the all-replica adjoint is tuned on its first call
    return autotune_energy(c);
}

#define CHECK_CTX(c)                                                   \
    do {                                                               \
        if (!(c)) return fail(REMC_ERR_ARG, "null context");           \
    } while (0)

#define FINALIZE(c)                         \
    do {                                    \
        int _rc = finalize(c);              \
        if (_rc != REMC_OK) return _rc;     \
    } while (0)

#define CHECK_REP(c, r)                                                                          \
    do {                                                                                         \
        if ((r) < 0 || (r) >= (int)(c)->reps.size()) return fail(REMC_ERR_ARG, "replica %d out of range", (r)); \
    } while (0)

// Copies [synced_copies, coop) of the dynamic state are refreshed from copy 0.
int sync_copies(remc_ctx *c, int coop)
{
    if (c->synced_copies >= coop) return REMC_OK;
    const size_t nb = (size_t)c->total_rep_beads, nm = (size_t)c->total_rep_bodies;
    float *fb[6] = {c->d_px, c->d_py, c->d_pz, c->d_rx, c->d_ry, c->d_rz};
    float *fm[3] = {c->d_cx, c->d_cy, c->d_cz};
    for (int k = c->synced_copies; k < coop; ++k) {
        for (float *a : fb) HIP_TRY(hipMemcpyAsync(a + k * nb, a, nb * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        for (float *a : fm) HIP_TRY(hipMemcpyAsync(a + k * nm, a, nm * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->d_quat + 4 * k * nm, c->d_quat, 4 * nm * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
    }
    c->synced_copies = coop;
    return REMC_OK;
}

// Reads and clears the device error flag of the cooperative MC kernel.
int check_err(remc_ctx *c)
{
    int err = 0;
    HIP_TRY(hipMemcpyAsync(&err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (err) {
        int zero = 0;
        HIP_TRY(hipMemcpyAsync(c->d_err, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return fail(REMC_ERR_DEVICE, "cooperative MC workgroups did not meet within the wait limit; replica state is invalid");
    }
    return REMC_OK;
}

// LDS layout of the MC kernel for a given cooperation: the moved-body staging
// always; the partner beads of the replica (or of the member's slice when
// every move is split) and the body table when they fit the LDS share of one
// workgroup at the occupancy the grid needs.
void fill_mc_lds(remc_ctx *c, KParams &p, int coop, int block)
{
    const int R = (int)c->reps.size();
    const long long wgs = (long long)R * coop;
    const int per_cu = (int)std::max<long long>(1, (wgs + c->n_cu - 1) / c->n_cu);
    // etab, draws, window staging and reduction scratch, as the kernel declares them, and slack
    const int kernel_lds = mc_static_lds(block);
    const int static_lds = (kernel_lds > 0 ? kernel_lds : 48 * 1024) + 2048;
    long long budget = std::min<long long>(c->lds_per_block, c->lds_per_cu / per_cu) - static_lds;
    p.lds_mb = std::max(1, c->max_body_beads);
    budget -= 68LL * p.lds_mb;
    p.lds_cl = 0;
    p.lds_nbody = 0;
    p.cache_slice = 0;
    p.cache_bodies = 0;
    p.cache_rel = 0;
    p.cache_quat = 0;
    p.spec = 0;
    p.draw2 = 0;
    if (!c->lds_cache || budget <= 0) return;
    const long long full = 24LL * c->max_beads;
    const bool all_split = coop > 1 && c->min_move_pairs >= c->coop_min_pairs;
    const long long slice = 24LL * ((c->max_beads + coop - 1) / coop + 1);
    [&] {
        if (full <= budget) {
            p.lds_cl = c->max_beads;
            budget -= full;
        } else if (all_split && slice <= budget) {
            p.lds_cl = (c->max_beads + coop - 1) / coop + 1;
            p.cache_slice = 1;
            budget -= slice;
        } else {
            return;
        }
        // speculative windows read the partner range (the replica, or the member's slice) from LDS
        p.spec = c->mc_window ? 1 : 0;
        if (20LL * c->max_bodies <= budget) {
            p.lds_nbody = c->max_bodies;
            p.cache_bodies = 1;
            budget -= 20LL * c->max_bodies;
        } else {
            return;
        }
        if (!p.cache_slice && 12LL * c->max_beads <= budget) {
            p.cache_rel = 1;
            budget -= 12LL * c->max_beads;
        }
        if (32LL * c->max_bodies <= budget) {
            p.cache_quat = 1;
            budget -= 32LL * c->max_bodies;
        }
    }();
    // a second draw batch, formed ahead by waves the window staging leaves
    // idle (workgroups of 8 waves or more), with what LDS the caches leave
    const long long draws = (long long)mc_draw_batch_bytes();
    if (p.spec && block >= 512 && draws <= budget) {
        p.draw2 = 1;
        budget -= draws;
    }
}

// Largest workgroups-per-replica count whose grid is co-resident.
int max_coop(remc_ctx *c, int block, int *out)
{
    const int R = (int)c->reps.size();
    int m = REMC_MAX_COOP;
    for (; m > 1; m >>= 1) {
        KParams p = c->params();
        fill_mc_lds(c, p, m, block);
        int per_cu = 0;
        HIP_TRY(mc_max_active_per_cu(p, block, &per_cu));
        if ((long long)m * R <= (long long)per_cu * c->n_cu) break;
    }
    *out = m;
    return REMC_OK;
}

int mc_launch(remc_ctx *c, const KParams &p0, int nsteps, int block, int coop)
{
    int rc = sync_copies(c, coop);
    if (rc != REMC_OK) return rc;
    KParams p = p0;
    fill_mc_lds(c, p, coop, block);
    HIP_TRY(launch_mc(p, (int)c->reps.size(), nsteps, block, coop, c->stream));
    c->synced_copies = coop;
    return REMC_OK;
}

// The copy-0 dynamic state of every replica (positions, relative positions,
// centres, rotations, tracked energies, step and decision counters), saved
// so that trial launches can run on the live state.  restore() puts it back
// bit for bit; the destructor restores it once more on every exit (error
// paths included) and releases the saves.
struct StateGuard {
    struct Buf {
        void *dev;
        size_t bytes;
        void *save;
    };
    remc_ctx *c;
    std::vector<Buf> bufs;
    explicit StateGuard(remc_ctx *c_) : c(c_)
    {
        const int R = (int)c->reps.size();
        const size_t nb = (size_t)c->total_rep_beads, nm = (size_t)c->total_rep_bodies;
        bufs = {{c->d_px, sizeof(float) * nb, nullptr},          {c->d_py, sizeof(float) * nb, nullptr},
                {c->d_pz, sizeof(float) * nb, nullptr},          {c->d_rx, sizeof(float) * nb, nullptr},
                {c->d_ry, sizeof(float) * nb, nullptr},          {c->d_rz, sizeof(float) * nb, nullptr},
                {c->d_cx, sizeof(float) * nm, nullptr},          {c->d_cy, sizeof(float) * nm, nullptr},
                {c->d_cz, sizeof(float) * nm, nullptr},          {c->d_quat, sizeof(double) * 4 * nm, nullptr},
                {c->d_E, sizeof(double) * R, nullptr},           {c->d_step, sizeof(unsigned long long) * R, nullptr},
                {c->d_cnt, sizeof(unsigned long long) * 3 * R, nullptr}};
    }
    int save()
    {
        int rc = sync_copies(c, 1);
        if (rc != REMC_OK) return rc;
        for (auto &b : bufs) {
            void *ptr = nullptr;
            HIP_TRY(hipMalloc(&ptr, b.bytes));
            b.save = ptr; // set only once allocated: the destructor copies back from it
            HIP_TRY(hipMemcpyAsync(b.save, b.dev, b.bytes, hipMemcpyDeviceToDevice, c->stream));
        }
        return REMC_OK;
    }
    int restore()
    {
        for (auto &b : bufs)
            if (b.save) HIP_TRY(hipMemcpyAsync(b.dev, b.save, b.bytes, hipMemcpyDeviceToDevice, c->stream));
        c->synced_copies = 1;
        return REMC_OK;
    }
    ~StateGuard()
    {
        (void)restore();
        (void)hipStreamSynchronize(c->stream);
        for (auto &b : bufs)
            if (b.save) (void)hipFree(b.save);
    }
};

int autotune_mc(remc_ctx *c)
{
    // time each (workgroup size, workgroups per replica) on the live state, then restore it exactly
    StateGuard st(c);
    int rc = st.save();
    if (rc != REMC_OK) return rc;
    EventPair ev;
    HIP_TRY(ev.create());
    KParams p = c->params();
    p.rec = nullptr;
    const int blocks[3] = {256, 512, 1024};
    double best = 1e30;
    int best_block = 256, best_coop = 1;
    const int trial_steps = 256;
    for (int bi = 0; bi < 3; ++bi) {
        const int blk = blocks[bi];
        if (c->mc_block && blk != c->mc_block) continue;
        int cmax = 1;
        rc = max_coop(c, blk, &cmax);
        if (rc != REMC_OK) return rc;
        for (int coop = 1; coop <= cmax; coop *= 2) {
            if (c->mc_coop && coop != c->mc_coop) continue;
            // every trial starts from the same saved state, so each one makes the same moves;
            // the best of two repeats discounts launch and clock noise
            float ms_best = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                c->synced_copies = 1;
                rc = mc_launch(c, p, 4, blk, coop); // warm
                if (rc != REMC_OK) return rc;
                HIP_TRY(hipEventRecord(ev.a, c->stream));
                rc = mc_launch(c, p, trial_steps, blk, coop);
                if (rc != REMC_OK) return rc;
                HIP_TRY(hipEventRecord(ev.b, c->stream));
                HIP_TRY(hipEventSynchronize(ev.b));
                rc = check_err(c);
                if (rc != REMC_OK) return rc;
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
                ms_best = std::min(ms_best, ms);
                rc = st.restore();
                if (rc != REMC_OK) return rc;
            }
            if (ms_best < best) {
                best = ms_best;
                best_block = blk;
                best_coop = coop;
            }
        }
    }
    c->mc_block = best_block;
    c->mc_coop = best_coop;
    return REMC_OK;
}

// Launch set-up of the force-biased kernel at one workgroup size: the LDS
// layout (the replica's partner beads with their relative positions, and its
// body centres and rotations, cached when they all fit) and the workgroups
// per replica.
int fb_prepare(remc_ctx *c, int block, KParams *out, int *coop_out)
{
    KParams p = c->params();
    p.lds_mb = std::max(1, c->max_body_beads);
    p.lds_cl = 0;
    const long long static_lds = 13 * 1024; // table, reductions, exchange, 64 batched draws
    if ((long long)mc_fb_dynamic_lds(p) + static_lds > c->lds_per_block)
        return fail(REMC_ERR_LIMIT, "a body of %d beads does not fit the force-biased kernel's LDS staging",
                    c->max_body_beads);
    p.lds_cl = c->max_beads;
    p.lds_nbody = c->max_bodies;
    if (!c->lds_cache || (long long)mc_fb_dynamic_lds(p) + static_lds > c->lds_per_block) {
        p.lds_cl = 0;
        p.lds_nbody = 0;
    }
    // windows of one-bead crowder moves (with the cache), switched with the MC kernel's
    p.spec = c->mc_window ? 1 : 0;
    // workgroups per replica: the MC kernel's setting when one is made, else
    // as many as give every workgroup a CU of its own; never more than run
    // co-resident (the members of a replica wait for each other)
    const int R = (int)c->reps.size();
    int coop = c->mc_coop > 0 ? c->mc_coop : REMC_MAX_COOP;
    if (c->mc_coop <= 0)
        while (coop > 1 && (long long)R * coop > (long long)c->n_cu) coop >>= 1;
    if (coop > 1) {
        int per_cu = 0;
        HIP_TRY(mc_fb_max_active_per_cu(p, block, &per_cu));
        while (coop > 1 && (long long)R * coop > (long long)per_cu * c->n_cu) coop >>= 1;
    }
    *out = p;
    *coop_out = coop;
    return REMC_OK;
}

// Workgroup size of the force-biased kernel: each of 256, 512 and 1024
// threads runs 128 steps on the live state (best of two), which is then
// restored bit for bit.
int autotune_fb(remc_ctx *c)
{
    StateGuard st(c);
    int rc = st.save();
    if (rc != REMC_OK) return rc;
    EventPair ev;
    HIP_TRY(ev.create());
    const int R = (int)c->reps.size();
    const int blocks[3] = {256, 512, 1024};
    float best = 1e30f;
    int best_block = 0;
    for (int bi = 0; bi < 3; ++bi) {
        KParams p{};
        int coop = 1;
        rc = fb_prepare(c, blocks[bi], &p, &coop);
        if (rc == REMC_ERR_LIMIT) continue;
        if (rc != REMC_OK) return rc;
        p.rec = nullptr;
        float ms_best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            rc = sync_copies(c, coop);
            if (rc != REMC_OK) return rc;
            HIP_TRY(launch_mc_fb(p, R, 4, blocks[bi], coop, c->stream)); // warm
            HIP_TRY(hipEventRecord(ev.a, c->stream));
            HIP_TRY(launch_mc_fb(p, R, 128, blocks[bi], coop, c->stream));
            HIP_TRY(hipEventRecord(ev.b, c->stream));
            HIP_TRY(hipEventSynchronize(ev.b));
            rc = check_err(c);
            if (rc != REMC_OK) return rc;
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
            ms_best = std::min(ms_best, ms);
            rc = st.restore();
            if (rc != REMC_OK) return rc;
        }
        if (ms_best < best) {
            best = ms_best;
            best_block = blocks[bi];
        }
    }
    if (!best_block) return fail(REMC_ERR_LIMIT, "no workgroup size of the force-biased kernel fits this context");
    c->fb_block = best_block;
    return REMC_OK;
}

int ensure_mc_config(remc_ctx *c)
{
    if (c->mc_block == 0 || c->mc_coop == 0) return autotune_mc(c);
    int cmax = 1;
    int rc = max_coop(c, c->mc_block, &cmax);
    if (rc != REMC_OK) return rc;
    if (c->mc_coop > cmax)
        return fail(REMC_ERR_LIMIT, "%d workgroups per replica do not fit co-resident (at most %d)", c->mc_coop, cmax);
    return REMC_OK;
}

// Adjoint of replicas [first, first + n): their beads occupy bead_stride
// entries of the replica arrays from bead_base, their bodies n_bodies_out
// entries from body_base.  Outputs: 6 doubles per body in replica order, 3
// parameter derivatives per replica.
// adjoint scratch of plan a for n replicas and n_bodies_out output bodies,
// grown as needed
int adjoint_scratch(remc_ctx *c, const AdjPlan &a, int n, int n_bodies_out)
{
    const size_t need = a.part_doubles + a.param_doubles + 6 * (size_t)n_bodies_out + 3 * (size_t)n;
    if (need > c->adj_cap) {
        if (c->d_adj) HIP_TRY(hipFree(c->d_adj));
        c->d_adj = nullptr;
        c->adj_cap = 0;
        HIP_TRY(hipMalloc(&c->d_adj, need * sizeof(double)));
        c->adj_cap = need;
    }
    return REMC_OK;
}

std::vector<unsigned char> adjoint_ff(const remc_ctx *c)
{
    std::vector<unsigned char> af(adjoint_ff_size());
    adjoint_ff_fill(af.data(), c->ff.table, c->ff.lambda, c->ff.e0, c->ff.xi, c->ff.lj_cf * c->ff.kbt_cf,
                    c->ff.dh_cf * c->ff.kbt_cf);
    return af;
}

// Workgroup target of the all-replica adjoint, from which its partner split
// follows: timed once per process and (replicas, beads), so every context of
// that size takes the same split and the same bits.  Each distinct split is
// run on the live conformations (read only), best of three; the size rule's
// target is kept unless another is at least 3% faster.
std::map<std::pair<int, int>, int> g_adj_cache; // under g_split_mu

int autotune_adjoint(remc_ctx *c)
{
    const int R = (int)c->reps.size();
    const std::pair<int, int> key(R, c->max_beads);
    {
        std::lock_guard<std::mutex> lock(g_split_mu);
        auto it = g_adj_cache.find(key);
        if (it != g_adj_cache.end()) {
            c->adj_target = it->second;
            return REMC_OK;
        }
    }
    const size_t stride = (size_t)c->total_rep_beads;
    const int cand[4] = {2048, 1024, 4096, 8192}; // the size rule's first
    float best_ms[4];
    int split[4];
    EventPair ev;
    HIP_TRY(ev.create());
    const KParams p = c->params();
    const std::vector<unsigned char> af = adjoint_ff(c);
    for (int k = 0; k < 4; ++k) {
        const AdjPlan a = adjoint_plan(R, c->max_beads, stride, cand[k]);
        split[k] = a.split;
        best_ms[k] = 1e30f;
        int same = -1;
        for (int q = 0; q < k; ++q)
            if (split[q] == a.split) same = q;
        if (same >= 0) continue; // the same launch as an earlier candidate
        int rc = adjoint_scratch(c, a, R, c->total_rep_bodies);
        if (rc != REMC_OK) return rc;
        double *ft = c->d_adj + a.part_doubles + a.param_doubles, *dp = ft + 6 * (size_t)c->total_rep_bodies;
        HIP_TRY(launch_adjoint(p, af.data(), a, 0, R, c->max_bodies, 0, stride, 0, c->d_adj, ft, dp, c->stream));
        for (int rep = 0; rep < 3; ++rep) {
            HIP_TRY(hipEventRecord(ev.a, c->stream));
            HIP_TRY(launch_adjoint(p, af.data(), a, 0, R, c->max_bodies, 0, stride, 0, c->d_adj, ft, dp, c->stream));
            HIP_TRY(hipEventRecord(ev.b, c->stream));
            HIP_TRY(hipEventSynchronize(ev.b));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
            best_ms[k] = std::min(best_ms[k], ms);
        }
    }
    int pick = cand[0];
    float t_pick = best_ms[0];
    for (int k = 1; k < 4; ++k)
        if (best_ms[k] < 0.97f * best_ms[0] && best_ms[k] < t_pick) {
            pick = cand[k];
            t_pick = best_ms[k];
        }
    std::lock_guard<std::mutex> lock(g_split_mu);
    auto it = g_adj_cache.emplace(key, pick).first; // a concurrent tuner's entry wins
    c->adj_target = it->second;
    return REMC_OK;
}

int adjoint_range(remc_ctx *c, int first, int n, int max_beads, int max_bodies, int bead_base, size_t bead_stride,
                  int body_base, int n_bodies_out, double *force_torque, double *dparams, int target_wg = 0)
{
    const AdjPlan a = adjoint_plan(n, max_beads, bead_stride, target_wg);
    int rc = adjoint_scratch(c, a, n, n_bodies_out);
    if (rc != REMC_OK) return rc;
    double *ft = c->d_adj + a.part_doubles + a.param_doubles, *dp = ft + 6 * (size_t)n_bodies_out;
    const std::vector<unsigned char> af = adjoint_ff(c);
    {
        TimedLaunch t(c, 4);
        HIP_TRY(launch_adjoint(c->params(), af.data(), a, first, n, max_bodies, bead_base, bead_stride, body_base,
                               c->d_adj, ft, dp, c->stream));
    }
    if (force_torque)
        HIP_TRY(hipMemcpyAsync(force_torque, ft, sizeof(double) * 6 * n_bodies_out, hipMemcpyDeviceToHost, c->stream));
    if (dparams) HIP_TRY(hipMemcpyAsync(dparams, dp, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

} // namespace

extern "C" {

const char *remc_last_error(void) { return g_last_error.c_str(); }

int remc_ctx_create(int device, remc_ctx **out)
{
    if (!out) return fail(REMC_ERR_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(REMC_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(REMC_ERR_ARG, "device %d out of range (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device));
    remc_ctx *c = new remc_ctx();
    c->device = device;
    e = scratch_address(device, &c->noalloc_scratch);
    if (e != hipSuccess) {
        delete c;
        return fail(REMC_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(REMC_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = c;
    return REMC_OK;
}

int remc_ctx_destroy(remc_ctx *c)
{
    if (!c) return REMC_OK;
    // teardown: errors here leave nothing to recover, the memory goes either way
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    collect_timing(c);
    destroy_graph(c);
    void *ptrs[] = {c->d_tc, c->d_body, c->d_bstart, c->d_bend, c->d_bcrow, c->d_q, c->d_sig, c->d_etab, c->d_sstab,
                    c->d_sys, c->d_rep, c->d_px, c->d_py, c->d_pz, c->d_rx, c->d_ry, c->d_rz,
                    c->d_cx, c->d_cy, c->d_cz, c->d_quat, c->d_E, c->d_T, c->d_step, c->d_cnt,
                    c->d_slot, c->d_exoff, c->d_ladder, c->d_exround, c->d_exswaps, c->d_full_partials,
                    c->d_out, c->d_plain_partials, c->d_rec, c->d_rec_start, c->d_slots, c->d_err, c->d_adj,
                    c->d_ana};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return REMC_OK;
}

int remc_set_forcefield(remc_ctx *c, const remc_forcefield *ff)
{
    CHECK_CTX(c);
    if (!ff) return fail(REMC_ERR_ARG, "null force field");
    if (c->finalized) return fail(REMC_ERR_STATE, "force field must be set before the first compute call");
    if (!(ff->xi > 0.0)) return fail(REMC_ERR_ARG, "Debye length must be positive");
    for (int i = 0; i < REMC_NTYPES; ++i)
        for (int j = 0; j < REMC_NTYPES; ++j)
            if (ff->table[i * REMC_NTYPES + j] != ff->table[j * REMC_NTYPES + i])
                return fail(REMC_ERR_ARG, "pair table is not symmetric at (%d,%d)", i, j);
    c->ff = *ff;
    c->ff_set = true;
    return REMC_OK;
}

int remc_set_seed(remc_ctx *c, uint64_t seed)
{
    CHECK_CTX(c);
    c->seed = seed;
    destroy_graph(c);
    return REMC_OK;
}

int remc_system_add(remc_ctx *c, const remc_system_desc *d, int *sys_id)
{
    CHECK_CTX(c);
    if (!d) return fail(REMC_ERR_ARG, "null system description");
    if (c->finalized) return fail(REMC_ERR_STATE, "systems must be added before the first compute call");
    if (d->n_beads <= 0 || d->n_bodies <= 0) return fail(REMC_ERR_ARG, "empty system");
    if (!d->body_start || !d->type || !d->charge || !d->sigma || !d->crowder || !d->x || !d->y || !d->z || !d->rx ||
        !d->ry || !d->rz || !d->cx || !d->cy || !d->cz || !d->quat)
        return fail(REMC_ERR_ARG, "null array in system description");
    if (d->body_start[0] != 0 || d->body_start[d->n_bodies] != d->n_beads)
        return fail(REMC_ERR_ARG, "body_start must run from 0 to n_beads");
    HostSys s;
    s.n_beads = d->n_beads;
    s.n_bodies = d->n_bodies;
    s.max_body = 0;
    for (int m = 0; m < d->n_bodies; ++m) {
        int n = d->body_start[m + 1] - d->body_start[m];
        if (n <= 0) return fail(REMC_ERR_ARG, "body %d is empty or body_start is not increasing", m);
        if (n > REMC_MAX_BODY_BEADS)
            return fail(REMC_ERR_LIMIT, "body %d has %d beads, above the limit %d", m, n, REMC_MAX_BODY_BEADS);
        s.max_body = std::max(s.max_body, n);
    }
    for (int b = 0; b < d->n_beads; ++b) {
        if (d->type[b] < 0 || d->type[b] >= REMC_NTYPES) {
            // crowder beads carry no type (the reference sets aa index -1)
            int m = (int)(std::upper_bound(d->body_start, d->body_start + d->n_bodies + 1, b) - d->body_start) - 1;
            if (!d->crowder[m]) return fail(REMC_ERR_ARG, "bead %d has type %d outside 0..19", b, d->type[b]);
        }
        if (!std::isfinite(d->x[b]) || !std::isfinite(d->y[b]) || !std::isfinite(d->z[b]) ||
            !std::isfinite(d->rx[b]) || !std::isfinite(d->ry[b]) || !std::isfinite(d->rz[b]))
            return fail(REMC_ERR_ARG, "bead %d has a non-finite position", b);
    }
    for (int m = 0; m < d->n_bodies; ++m) {
        if (!std::isfinite(d->cx[m]) || !std::isfinite(d->cy[m]) || !std::isfinite(d->cz[m]))
            return fail(REMC_ERR_ARG, "body %d has a non-finite centre", m);
        for (int k = 0; k < 4; ++k)
            if (!std::isfinite(d->quat[4 * m + k])) return fail(REMC_ERR_ARG, "body %d has a non-finite rotation", m);
    }
    if (!std::isfinite(d->box)) return fail(REMC_ERR_ARG, "box edge is not finite");
    s.body_start.assign(d->body_start, d->body_start + d->n_bodies + 1);
    s.type.resize(d->n_beads);
    for (int b = 0; b < d->n_beads; ++b) s.type[b] = (d->type[b] >= 0 && d->type[b] < REMC_NTYPES) ? d->type[b] : 0;
    s.crowder.resize(d->n_bodies);
    for (int m = 0; m < d->n_bodies; ++m) s.crowder[m] = d->crowder[m] ? 1 : 0;
    s.charge.assign(d->charge, d->charge + d->n_beads);
    s.sigma.assign(d->sigma, d->sigma + d->n_beads);
    s.box = d->box;
    s.x.assign(d->x, d->x + d->n_beads);
    s.y.assign(d->y, d->y + d->n_beads);
    s.z.assign(d->z, d->z + d->n_beads);
    s.rx.assign(d->rx, d->rx + d->n_beads);
    s.ry.assign(d->ry, d->ry + d->n_beads);
    s.rz.assign(d->rz, d->rz + d->n_beads);
    s.cx.assign(d->cx, d->cx + d->n_bodies);
    s.cy.assign(d->cy, d->cy + d->n_bodies);
    s.cz.assign(d->cz, d->cz + d->n_bodies);
    s.quat.assign(d->quat, d->quat + 4 * d->n_bodies);
    c->systems.push_back(std::move(s));
    if (sys_id) *sys_id = (int)c->systems.size() - 1;
    return REMC_OK;
}

int remc_replicas_add(remc_ctx *c, int sys_id, int n_rep, const double *temperature, const float *translate_step,
                      const double *rotate_step, int *first_replica)
{
    CHECK_CTX(c);
    if (c->finalized) return fail(REMC_ERR_STATE, "replicas must be added before the first compute call");
    if (sys_id < 0 || sys_id >= (int)c->systems.size()) return fail(REMC_ERR_ARG, "system %d out of range", sys_id);
    if (n_rep <= 0 || !temperature || !translate_step || !rotate_step) return fail(REMC_ERR_ARG, "bad replica arguments");
    HostSys &s = c->systems[sys_id];
    if (!s.replicas.empty()) return fail(REMC_ERR_STATE, "system %d already has replicas", sys_id);
    for (int k = 0; k < n_rep; ++k) {
        if (!(temperature[k] > 0.0)) return fail(REMC_ERR_ARG, "temperature %d must be positive", k);
        if (k > 0 && !(temperature[k] > temperature[k - 1]))
            return fail(REMC_ERR_ARG, "temperature ladder must be strictly ascending");
    }
    int first = (int)c->reps.size();
    for (int k = 0; k < n_rep; ++k) {
        c->reps.push_back({sys_id, temperature[k], translate_step[k], rotate_step[k]});
        s.replicas.push_back(first + k);
        s.ladder.push_back(temperature[k]);
    }
    if (first_replica) *first_replica = first;
    return REMC_OK;
}

int remc_get_sizes(remc_ctx *c, int *n_systems, int *n_replicas, int *max_beads)
{
    CHECK_CTX(c);
    if (n_systems) *n_systems = (int)c->systems.size();
    if (n_replicas) *n_replicas = (int)c->reps.size();
    if (max_beads) {
        int m = 0;
        for (auto &s : c->systems) m = std::max(m, s.n_beads);
        *max_beads = m;
    }
    return REMC_OK;
}

int remc_energy_plain(remc_ctx *c, int replica, int block, double *out)
{
    CHECK_CTX(c);
    FINALIZE(c);
    CHECK_REP(c, replica);
    if (!out) return fail(REMC_ERR_ARG, "null output");
    const HostSys &s = c->systems[c->reps[replica].sys];
    if (block == 0) block = s.n_beads < 1024 ? 32 : 64;
    if (block != 32 && block != 64 && block != 128 && block != 256)
        return fail(REMC_ERR_ARG, "plain block size must be 32, 64, 128 or 256");
    const int G = (s.n_beads + block - 1) / block;
    const size_t need = (size_t)G * G;
    if (need > c->plain_partials_cap) {
        if (c->d_plain_partials) (void)hipFree(c->d_plain_partials);
        c->d_plain_partials = nullptr;
        HIP_TRY(hipMalloc(&c->d_plain_partials, need * sizeof(float)));
        c->plain_partials_cap = need;
    }
    // blocks of the culled lower triangle never write: clear them first
    HIP_TRY(hipMemsetAsync(c->d_plain_partials, 0, need * sizeof(float), c->stream));
    {
        TimedLaunch t(c, 2);
        HIP_TRY(launch_energy_plain(c->params(), replica, s.n_beads, block, c->d_plain_partials, c->stream));
    }
    std::vector<float> h(need);
    HIP_TRY(hipMemcpyAsync(h.data(), c->d_plain_partials, need * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double sum = 0.0;
    for (size_t i = 0; i < need; ++i) sum += (double)h[i];
    *out = sum;
    return REMC_OK;
}

int remc_energy_all(remc_ctx *c, double *out)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (!out) return fail(REMC_ERR_ARG, "null output");
    const int R = (int)c->reps.size();
    {
        TimedLaunch t(c, 1);
        HIP_TRY(launch_energy_full(c->params(), R, c->max_tiles, c->d_full_partials, c->d_out, 0, 0, c->full_split, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(out, c->d_out, sizeof(double) * R, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_energy_all_noalloc(remc_ctx *c, double *out)
{
    CHECK_CTX(c);
    if (!out) return fail(REMC_ERR_ARG, "null output");
    // setting a context up allocates its buffers: that must have happened in
    // an earlier call, since this one allocates nothing
    if (!c->finalized)
        return fail(REMC_ERR_STATE, "context not set up: make a call that sets it up first (remc_get_energies, "
                                    "remc_energy_all, ...); remc_energy_all_noalloc allocates nothing");
    if (!c->noalloc_scratch) return fail(REMC_ERR_STATE, "no scratch address for device %d", c->device);
    const int R = (int)c->reps.size();
    const size_t stride = (size_t)full_units(c->max_tiles) * FULL_SPLIT_MAX;
    const size_t need = stride * (size_t)R + (size_t)R;
    if (need > REMC_NOALLOC_DOUBLES)
        return fail(REMC_ERR_LIMIT, "remc_energy_all_noalloc needs %zu doubles of scratch for %d replicas of up to %d "
                                    "beads; its static scratch holds %zu", need, R, c->max_beads, REMC_NOALLOC_DOUBLES);
    double *partials = c->noalloc_scratch, *dout = partials + stride * (size_t)R;
    HIP_TRY(launch_energy_full(c->params(), R, c->max_tiles, partials, dout, 0, 0, c->full_split, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, sizeof(double) * R, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_refresh_energies(remc_ctx *c)
{
    CHECK_CTX(c);
    FINALIZE(c);
    const int R = (int)c->reps.size();
    {
        TimedLaunch t(c, 1);
        HIP_TRY(launch_energy_full(c->params(), R, c->max_tiles, c->d_full_partials, nullptr, 0, 1, c->full_split, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_energy_noncrowder_all(remc_ctx *c, double *out)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (!out) return fail(REMC_ERR_ARG, "null output");
    const int R = (int)c->reps.size();
    {
        TimedLaunch t(c, 1);
        HIP_TRY(launch_energy_full(c->params(), R, c->max_tiles, c->d_full_partials, c->d_out, 1, 0, c->full_split, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(out, c->d_out, sizeof(double) * R, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_mc_steps(remc_ctx *c, int nsteps)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (nsteps < 0) return fail(REMC_ERR_ARG, "negative step count");
    if (nsteps == 0) return REMC_OK;
    int rc = ensure_mc_config(c);
    if (rc != REMC_OK) return rc;
    rc = sync_copies(c, c->mc_coop);
    if (rc != REMC_OK) return rc;
    {
        TimedLaunch t(c, 0);
        rc = mc_launch(c, c->params(), nsteps, c->mc_block, c->mc_coop);
    }
    return rc;
}

int remc_mc_steps_fb(remc_ctx *c, int nsteps, int block)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (nsteps < 0) return fail(REMC_ERR_ARG, "negative step count");
    if (block != 0 && block != 256 && block != 512 && block != 1024)
        return fail(REMC_ERR_ARG, "workgroup size %d (256, 512 or 1024; 0 = autotuned)", block);
    if (nsteps == 0) return REMC_OK;
    int rc = REMC_OK;
    if (block == 0) {
        if (c->fb_block == 0) {
            rc = autotune_fb(c);
            if (rc != REMC_OK) return rc;
        }
        block = c->fb_block;
    }
    KParams p{};
    int coop = 1;
    rc = fb_prepare(c, block, &p, &coop);
    if (rc != REMC_OK) return rc;
    // every member starts from the same state
    rc = sync_copies(c, coop);
    if (rc != REMC_OK) return rc;
    {
        TimedLaunch t(c, 0);
        HIP_TRY(launch_mc_fb(p, (int)c->reps.size(), nsteps, block, coop, c->stream));
    }
    c->synced_copies = coop;
    return REMC_OK;
}

int remc_mc_steps_plain(remc_ctx *c, int replica, int nsteps, int block)
{
    CHECK_CTX(c);
    FINALIZE(c);
    CHECK_REP(c, replica);
    if (nsteps < 0) return fail(REMC_ERR_ARG, "negative step count");
    const HostRep &h = c->reps[replica];
    const HostSys &s = c->systems[h.sys];
    const int N = s.n_beads, M = s.n_bodies;
    const int bo = c->rep_bead_off[replica], mo = c->rep_body_off[replica];
    hipStream_t st = c->stream;
    // host mirror of the replica (Replica's own Residue/Molecule arrays)
    std::vector<float> x(N), y(N), z(N), rx(N), ry(N), rz(N), cx(M), cy(M), cz(M);
    std::vector<double> quat(4 * M);
    double E;
    unsigned long long step, cnt[3];
    HIP_TRY(hipMemcpyAsync(x.data(), c->d_px + bo, 4 * N, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(y.data(), c->d_py + bo, 4 * N, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(z.data(), c->d_pz + bo, 4 * N, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(rx.data(), c->d_rx + bo, 4 * N, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(ry.data(), c->d_ry + bo, 4 * N, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(rz.data(), c->d_rz + bo, 4 * N, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(cx.data(), c->d_cx + mo, 4 * M, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(cy.data(), c->d_cy + mo, 4 * M, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(cz.data(), c->d_cz + mo, 4 * M, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(quat.data(), c->d_quat + 4 * mo, 32 * M, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&E, c->d_E + replica, 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&step, c->d_step + replica, 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(cnt, c->d_cnt + 3 * replica, 24, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const double ch = std::cos(h.rstep / 2.0), sh = std::sin(h.rstep / 2.0);
    const float box = s.box > 0.f ? s.box : 0.f;
    const uint32_t k0 = (uint32_t)c->seed, k1 = (uint32_t)(c->seed >> 32);
    std::vector<float> nx(s.max_body), ny(s.max_body), nz(s.max_body);
    for (int it = 0; it < nsteps; ++it, ++step) {
        remc_draws d = remc_step_draws(step, (uint32_t)replica, k0, k1, M);
        const int m0 = s.body_start[d.body], m1 = s.body_start[d.body + 1], nm = m1 - m0;
        double qmv[4], coef[9];
        float ncx = cx[d.body], ncy = cy[d.body], ncz = cz[d.body];
        if (d.rotate) {
            remc_move_quat(d.a, d.b, d.c, ch, sh, qmv);
            remc_qcoeffs(qmv, coef);
            for (int k = 0; k < nm; ++k) {
                float a, b, cc;
                remc_qrotate(coef, rx[m0 + k], ry[m0 + k], rz[m0 + k], &a, &b, &cc);
                nx[k] = a + ncx; ny[k] = b + ncy; nz[k] = cc + ncz;
            }
        } else {
            remc_move_centre(d.a, d.b, d.c, h.tstep, box, cx[d.body], cy[d.body], cz[d.body], &ncx, &ncy, &ncz);
            for (int k = 0; k < nm; ++k) {
                nx[k] = rx[m0 + k] + ncx; ny[k] = ry[m0 + k] + ncy; nz[k] = rz[m0 + k] + ncz;
            }
        }
        // MoleculeDataToDevice (Replica.cpp:424)
        HIP_TRY(hipMemcpyAsync(c->d_px + bo + m0, nx.data(), 4 * nm, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(c->d_py + bo + m0, ny.data(), 4 * nm, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(c->d_pz + bo + m0, nz.data(), 4 * nm, hipMemcpyHostToDevice, st));
        double e_new;
        int rc = remc_energy_plain(c, replica, block, &e_new);
        if (rc != REMC_OK) return rc;
        const double delta = e_new - E;
        const int dec = remc_metropolis(delta, h.T, d.u);
        if (dec) {
            for (int k = 0; k < nm; ++k) {
                x[m0 + k] = nx[k]; y[m0 + k] = ny[k]; z[m0 + k] = nz[k];
            }
            if (d.rotate) {
                for (int k = 0; k < nm; ++k) {
                    float a, b, cc;
                    remc_qrotate(coef, rx[m0 + k], ry[m0 + k], rz[m0 + k], &a, &b, &cc);
                    rx[m0 + k] = a; ry[m0 + k] = b; rz[m0 + k] = cc;
                }
                double qo[4] = {quat[4 * d.body], quat[4 * d.body + 1], quat[4 * d.body + 2], quat[4 * d.body + 3]};
                remc_qmul(qmv, qo, &quat[4 * d.body]);
            } else {
                cx[d.body] = ncx; cy[d.body] = ncy; cz[d.body] = ncz;
            }
            E = e_new;
            cnt[dec == 1 ? 0 : 1]++;
        } else {
            cnt[2]++;
            // undo + MoleculeDataToDevice again (Replica.cpp:468-471)
            HIP_TRY(hipMemcpyAsync(c->d_px + bo + m0, x.data() + m0, 4 * nm, hipMemcpyHostToDevice, st));
            HIP_TRY(hipMemcpyAsync(c->d_py + bo + m0, y.data() + m0, 4 * nm, hipMemcpyHostToDevice, st));
            HIP_TRY(hipMemcpyAsync(c->d_pz + bo + m0, z.data() + m0, 4 * nm, hipMemcpyHostToDevice, st));
        }
    }
    HIP_TRY(hipMemcpyAsync(c->d_rx + bo, rx.data(), 4 * N, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_ry + bo, ry.data(), 4 * N, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_rz + bo, rz.data(), 4 * N, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_cx + mo, cx.data(), 4 * M, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_cy + mo, cy.data(), 4 * M, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_cz + mo, cz.data(), 4 * M, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_quat + 4 * mo, quat.data(), 32 * M, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_E + replica, &E, 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_step + replica, &step, 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_cnt + 3 * replica, cnt, 24, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    c->synced_copies = 1;
    return REMC_OK;
}

int remc_exchange(remc_ctx *c, int *n_swaps)
{
    CHECK_CTX(c);
    FINALIZE(c);
    const int S = (int)c->systems.size();
    std::vector<unsigned long long> before(S), after(S);
    if (n_swaps)
        HIP_TRY(hipMemcpyAsync(before.data(), c->d_exswaps, 8 * S, hipMemcpyDeviceToHost, c->stream));
    {
        TimedLaunch t(c, 3);
        HIP_TRY(launch_exchange(c->params(), S, c->stream));
    }
    if (n_swaps) {
        HIP_TRY(hipMemcpyAsync(after.data(), c->d_exswaps, 8 * S, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        unsigned long long n = 0;
        for (int i = 0; i < S; ++i) n += after[i] - before[i];
        *n_swaps = (int)n;
    }
    return REMC_OK;
}

int remc_run(remc_ctx *c, int rounds, int steps_per_round)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (rounds < 0 || steps_per_round < 0) return fail(REMC_ERR_ARG, "negative rounds or steps");
    if (rounds == 0) return REMC_OK;
    int rc = ensure_mc_config(c);
    if (rc != REMC_OK) return rc;
    rc = sync_copies(c, c->mc_coop);
    if (rc != REMC_OK) return rc;
    if (!c->graph || c->graph_steps != steps_per_round) {
        destroy_graph(c);
        KParams p = c->params();
        fill_mc_lds(c, p, c->mc_coop, c->mc_block);
        int per_cu = 0; // also sets the kernel's dynamic-LDS attribute before capture
        HIP_TRY(mc_max_active_per_cu(p, c->mc_block, &per_cu));
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e1 = launch_mc(p, (int)c->reps.size(), steps_per_round, c->mc_block, c->mc_coop, c->stream);
        // with the refresh, the exchange reads full energies, as the reference's does
        if (e1 == hipSuccess && c->run_refresh)
            e1 = launch_energy_full(p, (int)c->reps.size(), c->max_tiles, c->d_full_partials, nullptr, 0, 1,
                                    c->full_split, c->stream);
        hipError_t e2 = launch_exchange(p, (int)c->systems.size(), c->stream);
        hipError_t e3 = hipStreamEndCapture(c->stream, &g);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(REMC_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3));
        }
        hipError_t e4 = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e4 != hipSuccess) {
            c->graph = nullptr;
            return fail(REMC_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e4));
        }
        c->graph_steps = steps_per_round;
    }
    for (int i = 0; i < rounds; ++i) {
        TimedLaunch t(c, 0);
        HIP_TRY(hipGraphLaunch(c->graph, c->stream));
    }
    return REMC_OK;
}

int remc_set_run_refresh(remc_ctx *c, int enable)
{
    CHECK_CTX(c);
    const bool on = enable != 0;
    if (on != c->run_refresh) destroy_graph(c); // the captured rounds change
    c->run_refresh = on;
    return REMC_OK;
}

int remc_set_recording(remc_ctx *c, int max_steps)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (max_steps < 0) return fail(REMC_ERR_ARG, "negative recording length");
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->d_rec) (void)hipFree(c->d_rec);
    c->d_rec = nullptr;
    c->rec_max = 0;
    destroy_graph(c);
    if (max_steps == 0) return REMC_OK;
    const int R = (int)c->reps.size();
    HIP_TRY(hipMalloc(&c->d_rec, sizeof(remc_step_record) * (size_t)R * max_steps));
    HIP_TRY(hipMemsetAsync(c->d_rec, 0xff, sizeof(remc_step_record) * (size_t)R * max_steps, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_rec_start, c->d_step, sizeof(unsigned long long) * R, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->rec_max = max_steps;
    return REMC_OK;
}

int remc_get_records(remc_ctx *c, int replica, remc_step_record *out, int n)
{
    CHECK_CTX(c);
    FINALIZE(c);
    CHECK_REP(c, replica);
    if (!out || n < 0 || n > c->rec_max) return fail(REMC_ERR_ARG, "bad record request");
    HIP_TRY(hipMemcpyAsync(out, c->d_rec + (size_t)replica * c->rec_max, sizeof(remc_step_record) * n,
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_get_energies(remc_ctx *c, double *out)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (!out) return fail(REMC_ERR_ARG, "null output");
    HIP_TRY(hipMemcpyAsync(out, c->d_E, sizeof(double) * c->reps.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return check_err(c);
}

int remc_get_temperatures(remc_ctx *c, double *out)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (!out) return fail(REMC_ERR_ARG, "null output");
    HIP_TRY(hipMemcpyAsync(out, c->d_T, sizeof(double) * c->reps.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_get_counters(remc_ctx *c, remc_counters *out)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (!out) return fail(REMC_ERR_ARG, "null output");
    const int R = (int)c->reps.size();
    std::vector<unsigned long long> cnt(3 * R), st(R);
    HIP_TRY(hipMemcpyAsync(cnt.data(), c->d_cnt, 24 * R, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(st.data(), c->d_step, 8 * R, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int r = 0; r < R; ++r) {
        out[r].accept = cnt[3 * r];
        out[r].accept_boltz = cnt[3 * r + 1];
        out[r].reject = cnt[3 * r + 2];
        out[r].steps = st[r];
    }
    return REMC_OK;
}

int remc_get_ladder(remc_ctx *c, int sys_id, int *slot_replica)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (sys_id < 0 || sys_id >= (int)c->systems.size() || !slot_replica) return fail(REMC_ERR_ARG, "bad ladder request");
    int off = 0;
    for (int i = 0; i < sys_id; ++i) off += (int)c->systems[i].replicas.size();
    HIP_TRY(hipMemcpyAsync(slot_replica, c->d_slot + off, sizeof(int) * c->systems[sys_id].replicas.size(),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_get_state(remc_ctx *c, int replica, float *x, float *y, float *z, float *rx, float *ry, float *rz,
                   float *cx, float *cy, float *cz, double *quat)
{
    CHECK_CTX(c);
    FINALIZE(c);
    CHECK_REP(c, replica);
    const HostSys &s = c->systems[c->reps[replica].sys];
    const size_t nb = s.n_beads, nm = s.n_bodies;
    const int bo = c->rep_bead_off[replica], mo = c->rep_body_off[replica];
    hipStream_t st = c->stream;
    float *dst_b[6] = {x, y, z, rx, ry, rz};
    float *src_b[6] = {c->d_px, c->d_py, c->d_pz, c->d_rx, c->d_ry, c->d_rz};
    for (int i = 0; i < 6; ++i)
        if (dst_b[i]) HIP_TRY(hipMemcpyAsync(dst_b[i], src_b[i] + bo, 4 * nb, hipMemcpyDeviceToHost, st));
    float *dst_m[3] = {cx, cy, cz};
    float *src_m[3] = {c->d_cx, c->d_cy, c->d_cz};
    for (int i = 0; i < 3; ++i)
        if (dst_m[i]) HIP_TRY(hipMemcpyAsync(dst_m[i], src_m[i] + mo, 4 * nm, hipMemcpyDeviceToHost, st));
    if (quat) HIP_TRY(hipMemcpyAsync(quat, c->d_quat + 4 * (size_t)mo, 32 * nm, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return REMC_OK;
}

int remc_set_state(remc_ctx *c, int replica, const float *x, const float *y, const float *z, const float *rx,
                   const float *ry, const float *rz, const float *cx, const float *cy, const float *cz,
                   const double *quat)
{
    CHECK_CTX(c);
    FINALIZE(c);
    CHECK_REP(c, replica);
    const HostSys &s = c->systems[c->reps[replica].sys];
    const size_t nb = s.n_beads, nm = s.n_bodies;
    const int bo = c->rep_bead_off[replica], mo = c->rep_body_off[replica];
    hipStream_t st = c->stream;
    const float *src_b[6] = {x, y, z, rx, ry, rz};
    const float *src_c[3] = {cx, cy, cz};
    for (int i = 0; i < 6; ++i)
        if (src_b[i])
            for (size_t k = 0; k < nb; ++k)
                if (!std::isfinite(src_b[i][k])) return fail(REMC_ERR_ARG, "non-finite bead coordinate");
    for (int i = 0; i < 3; ++i)
        if (src_c[i])
            for (size_t k = 0; k < nm; ++k)
                if (!std::isfinite(src_c[i][k])) return fail(REMC_ERR_ARG, "non-finite body centre");
    if (quat)
        for (size_t k = 0; k < 4 * nm; ++k)
            if (!std::isfinite(quat[k])) return fail(REMC_ERR_ARG, "non-finite rotation");
    float *dst_b[6] = {c->d_px, c->d_py, c->d_pz, c->d_rx, c->d_ry, c->d_rz};
    for (int i = 0; i < 6; ++i)
        if (src_b[i]) HIP_TRY(hipMemcpyAsync(dst_b[i] + bo, src_b[i], 4 * nb, hipMemcpyHostToDevice, st));
    const float *src_m[3] = {cx, cy, cz};
    float *dst_m[3] = {c->d_cx, c->d_cy, c->d_cz};
    for (int i = 0; i < 3; ++i)
        if (src_m[i]) HIP_TRY(hipMemcpyAsync(dst_m[i] + mo, src_m[i], 4 * nm, hipMemcpyHostToDevice, st));
    if (quat) HIP_TRY(hipMemcpyAsync(c->d_quat + 4 * (size_t)mo, quat, 32 * nm, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    c->synced_copies = 1;
    return REMC_OK;
}

int remc_set_energies(remc_ctx *c, const double *E)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (!E) return fail(REMC_ERR_ARG, "null input");
    const int R = (int)c->reps.size();
    for (int r = 0; r < R; ++r)
        if (!std::isfinite(E[r])) return fail(REMC_ERR_ARG, "non-finite energy for replica %d", r);
    HIP_TRY(hipMemcpyAsync(c->d_E, E, sizeof(double) * R, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_set_counters(remc_ctx *c, const remc_counters *in)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (!in) return fail(REMC_ERR_ARG, "null input");
    const int R = (int)c->reps.size();
    std::vector<unsigned long long> cnt(3 * R), st(R);
    for (int r = 0; r < R; ++r) {
        cnt[3 * r] = in[r].accept;
        cnt[3 * r + 1] = in[r].accept_boltz;
        cnt[3 * r + 2] = in[r].reject;
        st[r] = in[r].steps;
    }
    HIP_TRY(hipMemcpyAsync(c->d_cnt, cnt.data(), 24 * R, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_step, st.data(), 8 * R, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

static int slot_offset(remc_ctx *c, int sys_id)
{
    int off = 0;
    for (int i = 0; i < sys_id; ++i) off += (int)c->systems[i].replicas.size();
    return off;
}

int remc_set_ladder(remc_ctx *c, int sys_id, const int *slot_replica)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (sys_id < 0 || sys_id >= (int)c->systems.size() || !slot_replica) return fail(REMC_ERR_ARG, "bad ladder");
    const HostSys &s = c->systems[sys_id];
    const int n = (int)s.replicas.size();
    std::vector<int> seen(c->reps.size(), 0);
    for (int k = 0; k < n; ++k) {
        const int r = slot_replica[k];
        if (r < 0 || r >= (int)c->reps.size() || c->reps[r].sys != sys_id || seen[r]++)
            return fail(REMC_ERR_ARG, "ladder entry %d (replica %d) is not a permutation of system %d's replicas", k, r,
                        sys_id);
    }
    const int R = (int)c->reps.size();
    std::vector<double> T(R);
    HIP_TRY(hipMemcpyAsync(T.data(), c->d_T, sizeof(double) * R, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int k = 0; k < n; ++k) T[slot_replica[k]] = s.ladder[k];
    HIP_TRY(hipMemcpyAsync(c->d_slot + slot_offset(c, sys_id), slot_replica, sizeof(int) * n, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_T, T.data(), sizeof(double) * R, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_get_exchange_state(remc_ctx *c, int sys_id, uint64_t *round, int *offset, uint64_t *swaps)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (sys_id < 0 || sys_id >= (int)c->systems.size()) return fail(REMC_ERR_ARG, "bad system %d", sys_id);
    unsigned long long rd = 0, sw = 0;
    int of = 0;
    HIP_TRY(hipMemcpyAsync(&rd, c->d_exround + sys_id, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&of, c->d_exoff + sys_id, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&sw, c->d_exswaps + sys_id, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (round) *round = rd;
    if (offset) *offset = of;
    if (swaps) *swaps = sw;
    return REMC_OK;
}

int remc_set_exchange_state(remc_ctx *c, int sys_id, uint64_t round, int offset, uint64_t swaps)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (sys_id < 0 || sys_id >= (int)c->systems.size()) return fail(REMC_ERR_ARG, "bad system %d", sys_id);
    if (offset != 0 && offset != 1) return fail(REMC_ERR_ARG, "exchange offset must be 0 or 1");
    const unsigned long long rd = round, sw = swaps;
    HIP_TRY(hipMemcpyAsync(c->d_exround + sys_id, &rd, 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_exoff + sys_id, &offset, 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_exswaps + sys_id, &sw, 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

// Grows the analysis scratch to at least `bytes`.
static int ana_reserve(remc_ctx *c, size_t bytes)
{
    if (c->ana_bytes >= bytes) return REMC_OK;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->d_ana) (void)hipFree(c->d_ana);
    c->d_ana = nullptr;
    c->ana_bytes = 0;
    HIP_TRY(hipMalloc(&c->d_ana, bytes));
    c->ana_bytes = bytes;
    return REMC_OK;
}

int remc_drms(remc_ctx *c, int sys_id, const float *ref_x, const float *ref_y, const float *ref_z, double *out)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (sys_id < 0 || sys_id >= (int)c->systems.size()) return fail(REMC_ERR_ARG, "bad system %d", sys_id);
    if (!ref_x || !ref_y || !ref_z || !out) return fail(REMC_ERR_ARG, "null argument");
    const HostSys &s = c->systems[sys_id];
    const int N = s.n_beads, R = (int)s.replicas.size();
    for (int b = 0; b < N; ++b)
        if (!std::isfinite(ref_x[b]) || !std::isfinite(ref_y[b]) || !std::isfinite(ref_z[b]))
            return fail(REMC_ERR_ARG, "non-finite reference coordinate");
    double pairs = 0.0, beads_before = 0.0; // pairs of beads of different non-crowder bodies
    for (int m = 0; m < s.n_bodies; ++m) {
        if (s.crowder[m]) continue;
        const double nm = s.body_start[m + 1] - s.body_start[m];
        pairs += nm * beads_before;
        beads_before += nm;
    }
    if (pairs <= 0.0) return fail(REMC_ERR_ARG, "system %d has fewer than two non-crowder bodies", sys_id);
    const int tiles = (N + 255) / 256;
    const size_t fbytes = ((3 * (size_t)N * sizeof(float) + 255) / 256) * 256;
    int rc = ana_reserve(c, fbytes + sizeof(double) * ((size_t)R * tiles + R));
    if (rc != REMC_OK) return rc;
    float *rx = (float *)c->d_ana, *ry = rx + N, *rz = ry + N;
    double *part = (double *)((char *)c->d_ana + fbytes), *dout = part + (size_t)R * tiles;
    HIP_TRY(hipMemcpyAsync(rx, ref_x, sizeof(float) * N, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(ry, ref_y, sizeof(float) * N, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(rz, ref_z, sizeof(float) * N, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_drms(c->params(), sys_id, N, R, rx, ry, rz, part, 1.0 / pairs, dout, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, sizeof(double) * R, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_centre_histogram(remc_ctx *c, int sys_id, const int *body_type, int n_types, int nbin, double lo, double hi,
                          uint64_t *counts)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (sys_id < 0 || sys_id >= (int)c->systems.size()) return fail(REMC_ERR_ARG, "bad system %d", sys_id);
    if (!body_type || !counts) return fail(REMC_ERR_ARG, "null argument");
    if (n_types < 1 || nbin < 1 || !(hi > lo) || !std::isfinite(lo) || !std::isfinite(hi))
        return fail(REMC_ERR_ARG, "need n_types >= 1, nbin >= 1 and lo < hi");
    const size_t nh = (size_t)n_types * n_types * nbin;
    if (nh * sizeof(unsigned int) > 48 * 1024)
        return fail(REMC_ERR_LIMIT, "histogram of %zu bins exceeds the 12288-bin workgroup limit", nh);
    const HostSys &s = c->systems[sys_id];
    const int M = s.n_bodies, R = (int)s.replicas.size();
    for (int m = 0; m < M; ++m)
        if (body_type[m] < -1 || body_type[m] >= n_types) return fail(REMC_ERR_ARG, "body %d has type %d", m, body_type[m]);
    const size_t tbytes = ((sizeof(int) * (size_t)M + 255) / 256) * 256;
    int rc = ana_reserve(c, tbytes + sizeof(unsigned long long) * nh);
    if (rc != REMC_OK) return rc;
    int *dtype = (int *)c->d_ana;
    unsigned long long *dcounts = (unsigned long long *)((char *)c->d_ana + tbytes);
    HIP_TRY(hipMemcpyAsync(dtype, body_type, sizeof(int) * M, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemsetAsync(dcounts, 0, sizeof(unsigned long long) * nh, c->stream));
    HIP_TRY(launch_centre_hist(c->params(), sys_id, M, R, dtype, nullptr, n_types, nbin, lo, (hi - lo) / nbin, dcounts,
                               c->stream));
    HIP_TRY(hipMemcpyAsync(counts, dcounts, sizeof(unsigned long long) * nh, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_radial_distribution(remc_ctx *c, int sys_id, const int *body_type, const int *site_bead, int n_types,
                             int nbin, double lo, double hi, double *gr)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (sys_id < 0 || sys_id >= (int)c->systems.size()) return fail(REMC_ERR_ARG, "bad system %d", sys_id);
    if (!body_type || !gr) return fail(REMC_ERR_ARG, "null argument");
    if (n_types < 1 || nbin < 1 || !(hi > lo) || !std::isfinite(lo) || !std::isfinite(hi))
        return fail(REMC_ERR_ARG, "need n_types >= 1, nbin >= 1 and lo < hi");
    const HostSys &s = c->systems[sys_id];
    if (!(s.box > 0.f)) return fail(REMC_ERR_ARG, "g(r) needs a periodic box (the volume normalises it)");
    const size_t nh = (size_t)n_types * n_types * nbin;
    if (nh * sizeof(unsigned int) > 48 * 1024)
        return fail(REMC_ERR_LIMIT, "histogram of %zu bins exceeds the 12288-bin workgroup limit", nh);
    const int M = s.n_bodies, R = (int)s.replicas.size();
    std::vector<int> type_count(n_types, 0);
    for (int m = 0; m < M; ++m) {
        if (body_type[m] < -1 || body_type[m] >= n_types) return fail(REMC_ERR_ARG, "body %d has type %d", m, body_type[m]);
        if (body_type[m] >= 0) ++type_count[body_type[m]];
        if (site_bead && body_type[m] >= 0 &&
            (site_bead[m] < s.body_start[m] || site_bead[m] >= s.body_start[m + 1]))
            return fail(REMC_ERR_ARG, "site bead %d is not a bead of body %d", site_bead[m], m);
    }
    auto up = [](size_t b) { return ((b + 255) / 256) * 256; };
    const size_t tb = up(sizeof(int) * (size_t)M), sb = up(sizeof(int) * (size_t)M), cb = up(sizeof(int) * n_types);
    const size_t hb = up(sizeof(unsigned long long) * nh);
    int rc = ana_reserve(c, tb + sb + cb + hb + sizeof(double) * nh);
    if (rc != REMC_OK) return rc;
    char *base = (char *)c->d_ana;
    int *dtype = (int *)base, *dsite = (int *)(base + tb), *dcount = (int *)(base + tb + sb);
    unsigned long long *dhist = (unsigned long long *)(base + tb + sb + cb);
    double *dgr = (double *)(base + tb + sb + cb + hb);
    HIP_TRY(hipMemcpyAsync(dtype, body_type, sizeof(int) * M, hipMemcpyHostToDevice, c->stream));
    if (site_bead) HIP_TRY(hipMemcpyAsync(dsite, site_bead, sizeof(int) * M, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dcount, type_count.data(), sizeof(int) * n_types, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemsetAsync(dhist, 0, sizeof(unsigned long long) * nh, c->stream));
    const double dr = (hi - lo) / nbin, L = (double)s.box;
    HIP_TRY(launch_centre_hist(c->params(), sys_id, M, R, dtype, site_bead ? dsite : nullptr, n_types, nbin, lo, dr,
                               dhist, c->stream));
    HIP_TRY(launch_gr_normalise(dhist, dcount, n_types, nbin, lo, dr, (double)R / (L * L * L), dgr, c->stream));
    HIP_TRY(hipMemcpyAsync(gr, dgr, sizeof(double) * nh, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return REMC_OK;
}

int remc_adjoint(remc_ctx *c, int replica, double *force_torque, double *dparams)
{
    CHECK_CTX(c);
    FINALIZE(c);
    CHECK_REP(c, replica);
    if (!force_torque && !dparams) return fail(REMC_ERR_ARG, "null outputs");
    const HostSys &s = c->systems[c->reps[replica].sys];
    return adjoint_range(c, replica, 1, s.n_beads, s.n_bodies, c->rep_bead_off[replica], (size_t)s.n_beads,
                         c->rep_body_off[replica], s.n_bodies, force_torque, dparams);
}

int remc_adjoint_all(remc_ctx *c, double *force_torque, double *dparams)
{
    CHECK_CTX(c);
    FINALIZE(c);
    if (!force_torque && !dparams) return fail(REMC_ERR_ARG, "null outputs");
    if (c->adj_target == 0) {
        const int rc = autotune_adjoint(c);
        if (rc != REMC_OK) return rc;
    }
    return adjoint_range(c, 0, (int)c->reps.size(), c->max_beads, c->max_bodies, 0, (size_t)c->total_rep_beads, 0,
                         c->total_rep_bodies, force_torque, dparams, c->adj_target);
}

int remc_timing_enable(remc_ctx *c, int enable)
{
    CHECK_CTX(c);
    int rc = collect_timing(c);
    if (rc != REMC_OK) return rc;
    for (int i = 0; i < 5; ++i) {
        c->t_total[i] = 0.0;
        c->t_count[i] = 0;
    }
    c->timing = enable != 0;
    return REMC_OK;
}

int remc_timing_get(remc_ctx *c, int kernel_class, double *total_ms, int64_t *launches)
{
    CHECK_CTX(c);
    if (kernel_class < 0 || kernel_class > 4) return fail(REMC_ERR_ARG, "kernel class out of range");
    int rc = collect_timing(c);
    if (rc != REMC_OK) return rc;
    if (total_ms) *total_ms = c->t_total[kernel_class];
    if (launches) *launches = c->t_count[kernel_class];
    return REMC_OK;
}

int remc_synchronize(remc_ctx *c)
{
    CHECK_CTX(c);
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->finalized) return check_err(c);
    return REMC_OK;
}

int remc_set_mc_block(remc_ctx *c, int block)
{
    CHECK_CTX(c);
    if (block != 0 && block != 256 && block != 512 && block != 1024)
        return fail(REMC_ERR_ARG, "MC block must be 0, 256, 512 or 1024");
    c->mc_block = block;
    destroy_graph(c);
    return REMC_OK;
}

int remc_set_mc_coop(remc_ctx *c, int coop)
{
    CHECK_CTX(c);
    if (coop != 0 && coop != 1 && coop != 2 && coop != 4 && coop != 8)
        return fail(REMC_ERR_ARG, "workgroups per replica must be 0, 1, 2, 4 or 8");
    c->mc_coop = coop;
    destroy_graph(c);
    return REMC_OK;
}

int remc_get_mc_coop(remc_ctx *c, int *coop)
{
    CHECK_CTX(c);
    if (!coop) return fail(REMC_ERR_ARG, "null output");
    *coop = c->mc_coop;
    return REMC_OK;
}

int remc_set_mc_tuning(remc_ctx *c, int lds_cache, int split_min_pairs)
{
    CHECK_CTX(c);
    if (split_min_pairs < 0) return fail(REMC_ERR_ARG, "split threshold must be >= 0");
    c->lds_cache = lds_cache != 0;
    if (split_min_pairs > 0) c->coop_min_pairs = split_min_pairs;
    destroy_graph(c);
    return REMC_OK;
}

int remc_set_mc_window(remc_ctx *c, int enable)
{
    CHECK_CTX(c);
    c->mc_window = enable != 0;
    destroy_graph(c);
    return REMC_OK;
}

int remc_get_fb_block(remc_ctx *c, int *block)
{
    CHECK_CTX(c);
    if (!block) return fail(REMC_ERR_ARG, "null output");
    *block = c->fb_block;
    return REMC_OK;
}

int remc_get_mc_block(remc_ctx *c, int *block)
{
    CHECK_CTX(c);
    if (!block) return fail(REMC_ERR_ARG, "null output");
    *block = c->mc_block;
    return REMC_OK;
}

} // extern "C"
