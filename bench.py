#!/usr/bin/env python3
"""Benchmark of the REMC hot path on one MI355X.

A "step" of this benchmark is one exchange round: --steps-per-launch (default
1000) MC steps of every replica of the workload (proposal, dE of the moved
body, Metropolis) followed by one temperature-exchange round over the ladder,
run as one hipGraph replay of (persistent MC kernel + exchange kernel).
--steps counts timed rounds and --warmup untimed ones, so the default
`bench.py --gpus 1 --steps 20 --warmup 5` times 20,000 MC steps per replica.
Inputs are resident in HBM before the timed region.

Default workload: config 4 (10 x 100-residue chains + 300 crowders = 1300
beads in 310 rigid bodies, 32 replicas 288-480 K), the configuration the
throughput target is stated on.  Reported beside it:

- the plain path on the same workload: the reference's per-step protocol
  with the reference-shaped tiled kernel, and that kernel's own rate;
- the force-biased (adjoint-driven) MC kernel, the batched adjoint and the
  batched full energy;
- the CPU baseline: the FP64 C restatement of the reference CPU path (full
  recomputation per step, replicas block-partitioned over pthreads), on all
  host threads and on one, on a bounded sample, plus config 1 in full
  (8 replicas x 10,000 steps with exchanges);
- the largest |dE_GPU - dE_FP64| in kT on every configuration 1-5.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

METRIC = "MC steps/sec & pair-interactions/sec, batched vs plain, |ΔE|<1e-4 kT vs FP64"
FLOPS_PER_PAIR = 41          # fixed counting rule per pair term (FMA = 2, sqrt/rcp/exp/rint = 1)
FLOPS_PER_PAIR_DERIV = 71    # energy and derivative of one pair (41 + 30)
PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector peak (datasheet)
PEAK_FP32_FMA_TFLOPS = 78.6  # non-packed FP32 FMA ceiling
SEED = 0x52454D43


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed rounds; a round is --steps-per-launch MC steps of every replica + one exchange")
    ap.add_argument("--warmup", type=int, default=2, help="untimed rounds before the timed ones (at least 1 is run)")
    ap.add_argument("--steps-per-launch", type=int, default=1000, help="MC steps per round (between exchanges)")
    ap.add_argument("--config", type=int, default=4, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--plain-steps", type=int, default=300, help="MC steps of the plain protocol (replica 0)")
    ap.add_argument("--cpu-steps", type=int, default=200,
                    help="MC steps per replica of the CPU baseline; the GPU records the same first steps")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the CPU baseline (0 = min(replicas, cores))")
    ap.add_argument("--cpu-single-replicas", type=int, default=4,
                    help="replicas of the 1-thread CPU figure (each --cpu-steps steps; 0 = skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-config1-full", action="store_true",
                    help="skip config 1 in full (8 replicas x 10,000 steps) on the GPU and the CPU")
    ap.add_argument("--accuracy-steps", type=int, default=100,
                    help="recorded MC steps per replica compared with FP64 on every config (0 = skip)")
    ap.add_argument("--block", type=int, default=0, help="MC workgroup size (0 = autotune)")
    ap.add_argument("--coop", type=int, default=0, help="MC workgroups per replica (0 = autotune)")
    ap.add_argument("--fb-steps", type=int, default=2000,
                    help="force-biased MC steps per replica timed after the main run (0 = skip)")
    ap.add_argument("--adjoint-calls", type=int, default=20,
                    help="adjoint calls (all replicas each) timed after the main run (0 = skip)")
    ap.add_argument("--energy-calls", type=int, default=20,
                    help="full-energy refreshes of every replica timed after the main run (0 = skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (HBM counter calibration) and exit")
    return ap.parse_args(argv)


def plan(args):
    """Checks the arguments and returns the run plan.  Raises SystemExit
    (status 2, with a message) for anything this benchmark cannot run."""
    if args.gpus != 1:
        raise SystemExit("bench.py measures one GPU: every replica of the workload runs on a single MI355X; "
                         "--gpus %d is not supported (runs over several GPUs are out of scope)" % args.gpus)
    if args.steps < 1:
        raise SystemExit("--steps must be at least 1 (timed rounds)")
    if args.warmup < 0 or args.steps_per_launch < 1:
        raise SystemExit("--warmup must be >= 0 and --steps-per-launch >= 1")
    if args.cpu_steps < 1 or args.plain_steps < 1:
        raise SystemExit("--cpu-steps and --plain-steps must be at least 1")
    return {
        "rounds": args.steps,
        "warm_rounds": max(args.warmup, 1),
        "spl": args.steps_per_launch,
        "timed_mc_steps_per_replica": args.steps * args.steps_per_launch,
        "record_steps": args.cpu_steps,
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def workload_label(cid, wl):
    sa = wl.systems[0]
    labels = {
        1: "config1: 76+24-residue chains, 150 A box",
        2: "config2: 76+24-residue chains + 200 crowders, 300 A box",
        3: "config3: 64 uneven systems (2-12 chains of 30-300 residues, 0-500 crowders)",
        4: "config4: 10 x 100-residue chains + 300 crowders",
        5: "config5: 60 x 150-residue chains + 2000 crowders, 600 A box",
    }
    return "%s; %d replicas; %s" % (labels[cid], wl.n_replicas,
                                    "%d beads / %d bodies / box %.1f A" % (sa.n_beads, sa.n_bodies, sa.box)
                                    if len(wl.systems) == 1 else "%d systems" % len(wl.systems))


def read_traffic(kernel, workload):
    path = os.path.join(REPO, "profiles", "traffic_%s_%s.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get("hbm_bytes_per_launch")


def compare_records(gpu_rec, orc_rec, temperatures, n):
    """Largest |dE_GPU - dE_FP64| in kT over the steps where both chains
    still agree (they part at a decision inside the FP32 rounding of a
    threshold), with counts."""
    from remcdockingsimulator_amd import forcefield as ffm

    worst, worst_rel, agree, total = 0.0, 0.0, 0, 0
    worst_at = None
    for r in range(len(orc_rec)):
        g, o = gpu_rec[r][:n], orc_rec[r]
        kt = ffm.kT(temperatures[r])
        for k in range(n):
            total += 1
            if g["body"][k] != o["body"][k] or g["accepted"][k] != o["accepted"][k]:
                break
            err = abs(g["delta"][k] - o["delta"][k]) / kt
            if err > worst:
                worst = err
                worst_at = {"replica": r, "step": k, "dE_fp64_kT": float(o["delta"][k] / kt)}
            worst_rel = max(worst_rel, abs(g["delta"][k] - o["delta"][k]) / max(abs(o["delta"][k]), 1e-30))
            agree += 1
    return {"max_abs_dE_err_kT": worst, "max_rel_dE_err": worst_rel, "worst": worst_at,
            "steps_compared": total, "steps_equal": agree}


def oracle_replicas(wl, s, first_rng_id):
    from oracle import pyoracle

    osys = pyoracle.OracleSystem(wl.systems[s], wl.ff)
    lad = wl.temperatures[s]
    reps = [pyoracle.OracleReplica(osys, lad[k], wl.translate_step, wl.rotate_step, first_rng_id + k)
            for k in range(len(lad))]
    return osys, reps


def cpu_baseline(wl, nsteps, threads, single_reps, gpu_records):
    """The oracle (FP64 C restatement of the reference CPU path, full energy
    recomputation per step as Replica.cpp:435 does) timed on this host with
    replicas block-partitioned over pthreads (main.cpp:879-889), and on one
    thread; plus the |dE| comparison of the GPU's recorded steps against the
    FP64 values of the same steps."""
    from oracle import pyoracle

    lad = wl.temperatures[0]
    osys, reps = oracle_replicas(wl, 0, 0)
    secs, rec = pyoracle.mc_threaded(osys, reps, SEED, nsteps, 1, threads, record=True)
    n_steps = len(reps) * nsteps
    pairs = n_steps * wl.pairs_full()
    out = {
        "value": pairs / secs,
        "unit": "pair-interactions/s",
        "mc_steps_per_s": n_steps / secs,
        "cores": threads,
        "cpu_model": cpu_model(),
        "host_cores_visible": os.cpu_count(),
        "kind": "port",
        "sample": "%d replicas x %d MC steps of the same workload, full-energy recomputation per step "
                  "(%d pair terms each), FP64, %d pthreads" % (len(reps), nsteps, wl.pairs_full(), threads),
        "seconds": secs,
    }
    out.update(compare_records(gpu_records, rec, lad, nsteps))
    if single_reps > 0:
        k = min(single_reps, len(lad))
        osys1, reps1 = oracle_replicas(wl, 0, 0)
        s1, _ = pyoracle.mc_threaded(osys1, reps1[:k], SEED, nsteps, 1, 1, record=False)
        out["single_thread"] = {"value": k * nsteps * wl.pairs_full() / s1, "unit": "pair-interactions/s",
                                "mc_steps_per_s": k * nsteps / s1, "cores": 1, "seconds": s1,
                                "sample": "%d replicas x %d MC steps, 1 thread" % (k, nsteps)}
    return out


def energy_check(e, wl, replicas):
    """Full energies of the device state after the run: the device's full
    evaluation and its tracked (sum of accepted dE) energy against the
    oracle's FP64 full energy of the same conformations."""
    from oracle import pyoracle
    from remcdockingsimulator_amd import forcefield as ffm

    osys = pyoracle.OracleSystem(wl.systems[0], wl.ff)
    full = e.energy_all()
    tracked = e.energies()
    temps = e.temperatures()
    rel, trk = 0.0, 0.0
    for r in replicas:
        want = osys.energy(e.state(r)["pos"])
        rel = max(rel, abs(full[r] - want) / max(abs(want), 1e-12))
        trk = max(trk, abs(tracked[r] - want) / ffm.kT(temps[r]))
    return {"replicas": list(replicas), "max_rel_E_err_full": rel, "max_abs_tracked_E_err_kT": trk}


def config1_full(threads):
    """Config 1 as its workload states it: 8 replicas x 10,000 MC steps with an
    exchange round every 1,000, on the GPU (10 graph rounds) and on the CPU
    oracle (FP64, full recomputation per step, exchanges in between)."""
    from oracle import pyoracle
    from remcdockingsimulator_amd import synth
    from remcdockingsimulator_amd.engine import Engine

    wl = synth.config(1)
    R, every, total = wl.n_replicas, wl.exchange_every, wl.steps
    rounds = total // every
    e = Engine(0)
    synth.load_engine(e, wl, seed=SEED)
    e.mc_steps(1)          # autotune outside the timed region
    e.synchronize()
    e2 = Engine(0)
    synth.load_engine(e2, wl, seed=SEED)
    e2.set_mc_block(e.mc_block())
    e2.set_mc_coop(e.mc_coop())
    e.close()
    e2.synchronize()
    t0 = time.perf_counter()
    e2.run(rounds, every)
    e2.synchronize()
    gw = time.perf_counter() - t0
    e2.close()

    osys, reps = oracle_replicas(wl, 0, 0)
    lad = wl.temperatures[0]
    slots = np.arange(R, dtype=np.int32)
    t0 = time.perf_counter()
    csecs = 0.0
    for rd in range(rounds):
        s, _ = pyoracle.mc_threaded(osys, reps, SEED, every, 1, threads, record=False)
        csecs += s
        energies = np.array([r.energy for r in reps])
        slots, _, _ = pyoracle.exchange_round(lad, slots, energies, rd % 2, rd, 0, SEED)
        for k, rid in enumerate(slots):
            reps[rid].r.temperature = lad[k]
    cw = time.perf_counter() - t0
    pairs = R * total * wl.pairs_full()
    return {"workload": "config1 in full: 8 replicas x %d MC steps, exchange every %d" % (total, every),
            "gpu": {"seconds": gw, "mc_steps_per_s": R * total / gw,
                    "pair_interactions_per_s_full_equivalent": pairs / gw,
                    "form": "batched dE kernel, %d hipGraph rounds (first replay includes capture)" % rounds},
            "cpu": {"seconds": cw, "mc_seconds": csecs, "mc_steps_per_s": R * total / cw,
                    "pair_interactions_per_s": pairs / cw, "cores": min(threads, R),
                    "kind": "port (FP64 full recomputation per step)"}}


def accuracy_all_configs(nsteps):
    """Largest |dE_GPU - dE_FP64| (kT) on configs 1-5: every replica runs
    nsteps recorded MC steps on the GPU (batched kernel, autotuned) and the
    oracle replays the same counters in FP64 (dE of the moved body)."""
    from oracle import pyoracle
    from remcdockingsimulator_amd import synth
    from remcdockingsimulator_amd.engine import Engine

    out = {}
    for cid in (1, 2, 3, 4, 5):
        wl = synth.config(cid)
        e = Engine(0)
        synth.load_engine(e, wl, seed=SEED)
        e.set_recording(nsteps)
        e.mc_steps(nsteps)
        e.synchronize()
        res = {"max_abs_dE_err_kT": 0.0, "max_rel_dE_err": 0.0, "steps_compared": 0, "steps_equal": 0,
               "worst": None}
        rid = 0
        for s in range(len(wl.systems)):
            n_rep = len(wl.temperatures[s])
            if cid == 3 and s % 8:
                rid += n_rep
                continue  # config 3: every 8th system (8 of 64, 64 replicas)
            osys, reps = oracle_replicas(wl, s, rid)
            _, rec = pyoracle.mc_threaded(osys, reps, SEED, nsteps, 0, os.cpu_count() or 1, record=True)
            grec = [e.records(rid + k) for k in range(n_rep)]
            c = compare_records(grec, rec, wl.temperatures[s], nsteps)
            if c["max_abs_dE_err_kT"] >= res["max_abs_dE_err_kT"]:
                res["worst"] = c["worst"] and dict(c["worst"], replica=rid + c["worst"]["replica"])
            res["max_abs_dE_err_kT"] = max(res["max_abs_dE_err_kT"], c["max_abs_dE_err_kT"])
            res["max_rel_dE_err"] = max(res["max_rel_dE_err"], c["max_rel_dE_err"])
            res["steps_compared"] += c["steps_compared"]
            res["steps_equal"] += c["steps_equal"]
            rid += n_rep
        e.close()
        out["config%d" % cid] = res
        log("accuracy config %d: %s" % (cid, res))
    return out


def main(argv=None):
    args = parse_args(argv)
    try:
        pl = plan(args)
    except SystemExit as ex:
        print(ex, file=sys.stderr)
        sys.exit(2)

    from remcdockingsimulator_amd import rng as rngm
    from remcdockingsimulator_amd import synth
    from remcdockingsimulator_amd.engine import Engine

    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))
        return

    cid = args.config
    wl = synth.config(cid)
    R = wl.n_replicas
    spl, rounds, warm = pl["spl"], pl["rounds"], pl["warm_rounds"]
    single = len(wl.systems) == 1
    sizes = [np.diff(sa.body_start) for sa in wl.systems]

    e = Engine(0)
    synth.load_engine(e, wl, seed=SEED)
    if args.block:
        e.set_mc_block(args.block)
    if args.coop:
        e.set_mc_coop(args.coop)
    log("workload:", workload_label(cid, wl))

    # recorded steps (compared with the CPU oracle's FP64 dE of the same steps)
    rec_n = pl["record_steps"]
    e.set_recording(rec_n)
    e.mc_steps(rec_n)
    e.synchronize()
    gpu_records = [e.records(r) for r in range(R)]
    e.set_recording(0)
    step_cost = None
    if single:
        sizes0 = sizes[0]
        allrec = np.concatenate(gpu_records)
        nm = sizes0[allrec["body"]]
        small = nm * (sizes0.sum() - nm) < 8192
        us = allrec["ticks"] / 100.0  # 100 MHz device clock
        step_cost = {"small_move_us": float(np.mean(us[small])) if small.any() else None,
                     "large_move_us": float(np.mean(us[~small])) if (~small).any() else None,
                     "large_move_fraction": float(np.mean(~small))}
        log("per-step cost (recorded):", step_cost)

    # untimed rounds: graph capture and first replays
    e.run(warm, spl)
    e.synchronize()
    first_timed = rec_n + warm * spl
    mc_block, mc_coop = e.mc_block(), e.mc_coop()
    log("mc block:", mc_block, "workgroups per replica:", mc_coop)

    # ---------------------------------------------------------- timed ----
    e.timing_enable(True)
    e.synchronize()
    t0 = time.perf_counter()
    e.run(rounds, spl)
    e.synchronize()
    t1 = time.perf_counter()
    kms, nl = e.timing(Engine.MC)
    wall = t1 - t0

    # exact pair terms evaluated in the timed region (counter-based RNG replayed on the host)
    pairs = 0
    rid = 0
    for s, sa in enumerate(wl.systems):
        n_rep = len(wl.temperatures[s])
        pairs += rngm.delta_pairs(SEED, range(rid, rid + n_rep), first_timed, pl["timed_mc_steps_per_replica"],
                                  sizes[s])
        rid += n_rep
    value = pairs / wall
    steps_total = R * pl["timed_mc_steps_per_replica"]
    kernel_pairs_per_s = pairs / (kms / 1e3)
    achieved = kernel_pairs_per_s * FLOPS_PER_PAIR / 1e12
    traffic = read_traffic("k_mc", "config%d" % cid)
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                "frac_of_fma_ceiling": achieved / PEAK_FP32_FMA_TFLOPS,
                "kernel": "k_mc (persistent batched MC, dE of the moved body)",
                "kernel_avg_ms": kms / max(nl, 1), "launches": nl,
                "timing": "HIP events on the library stream around every graph replay (k_mc + k_exchange)",
                "flops_per_pair": FLOPS_PER_PAIR, "pairs_per_launch": pairs / max(nl, 1)}
    log("timed: %.3f s, %d rounds, kernel %.3f ms/round" % (wall, nl, kms / max(nl, 1)))

    # ---------------------------------------------------------- plain ----
    plain = None
    if not args.no_plain and single:
        e.mc_steps_plain(0, 3)  # warm
        e.timing_enable(True)
        e.synchronize()
        t2 = time.perf_counter()
        e.mc_steps_plain(0, args.plain_steps)
        t3 = time.perf_counter()
        pms, pnl = e.timing(Engine.PLAIN)
        pw = t3 - t2
        pf = wl.pairs_full()
        pk = pms / max(pnl, 1)
        plain = {"pair_interactions_per_s": args.plain_steps * pf / pw,
                 "mc_steps_per_s": args.plain_steps / pw,
                 "kernel_avg_ms": pk,
                 "kernel_pair_interactions_per_s": pf / (pk / 1e3),
                 "kernel_roofline_frac": pf / (pk / 1e3) * FLOPS_PER_PAIR / 1e12 / PEAK_FP32_TFLOPS,
                 "protocol": "reference per-step protocol: host proposal, H2D moved body, tiled kernel "
                             "(lower triangle culled, FP32 partials), D2H partials, host FP64 sum and Metropolis",
                 "steps": args.plain_steps}
        log("plain: %.3f s for %d steps, kernel %.4f ms" % (pw, args.plain_steps, pk))

    # ---------------------------------------------------- force-biased ----
    fb = None
    if args.fb_steps > 0 and single:
        e.mc_steps_fb(8)  # warm (launch configuration)
        e.synchronize()
        fb_first = int(e.counters()[0, 3])
        e.timing_enable(True)
        e.synchronize()
        t4 = time.perf_counter()
        e.mc_steps_fb(args.fb_steps)
        e.synchronize()
        t5 = time.perf_counter()
        fms, fnl = e.timing(Engine.MC)
        fw = t5 - t4
        fpairs = rngm.delta_pairs(SEED, range(R), fb_first, args.fb_steps, sizes[0], tag=rngm.TAG_FB)
        fk = fpairs / (fms / 1e3)
        fach = fk * FLOPS_PER_PAIR_DERIV / 1e12
        fb = {"mc_steps_per_s": R * args.fb_steps / fw,
              "dE_pair_interactions_per_s": fpairs / fw,
              "kernel_dE_pair_interactions_per_s": fk,
              "pair_evaluations_per_s": 2.0 * fpairs / fw,
              "roofline": {"bound": "valu", "achieved": fach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                           "frac": fach / PEAK_FP32_TFLOPS, "flops_per_pair": FLOPS_PER_PAIR_DERIV,
                           "counting": "every (moved bead, partner) pair at the old and the new position gets "
                                       "its energy and its derivative: 41 + 30 flops"},
              "kernel": "k_mc_fb (force-biased MC)",
              "kernel_avg_ms": fms / max(fnl, 1), "launches": fnl, "steps": args.fb_steps,
              "work_per_step": "force at the old position, energy and force at the new, energy at the old: "
                               "2 n_m (N - n_m) pair energies and 2 n_m (N - n_m) pair derivatives"}
        log("fb: %.3f s for %d steps" % (fw, args.fb_steps))

    # -------------------------------------------------------- adjoint ----
    adj = None
    if args.adjoint_calls > 0 and single:
        e.adjoint_all()  # warm
        e.timing_enable(True)
        e.synchronize()
        t6 = time.perf_counter()
        for k in range(args.adjoint_calls):
            e.adjoint_all()
        t7 = time.perf_counter()
        ams, anl = e.timing(Engine.ADJOINT)
        dpairs = 2 * wl.pairs_full() * R
        ak = ams / max(anl, 1)
        aach = dpairs / (ak / 1e3) * FLOPS_PER_PAIR_DERIV / 1e12
        adj = {"calls": args.adjoint_calls, "kernel_avg_ms": ak,
               "pair_derivatives_per_s": dpairs / (ak / 1e3),
               "roofline": {"bound": "valu", "achieved": aach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                            "frac": aach / PEAK_FP32_TFLOPS, "flops_per_pair": FLOPS_PER_PAIR_DERIV},
               "calls_per_s_wall": args.adjoint_calls / (t7 - t6),
               "kernel": "k_adjoint_rows + k_adjoint_bodies: adjoint of every replica in one launch (forces, torques, dE/d{lambda,e0,xi})",
               "work_per_call": "2 P_full ordered pair derivatives per replica (P_full = %d), %d replicas"
                                % (wl.pairs_full(), R)}
        log("adjoint: %.3f ms per call of all replicas (kernels)" % ak)

    # ---------------------------------------------------- full energy ----
    full = None
    if args.energy_calls > 0:
        e.refresh_energies()  # warm
        e.timing_enable(True)
        e.synchronize()
        t8 = time.perf_counter()
        for k in range(args.energy_calls):
            e.refresh_energies()
        t9 = time.perf_counter()
        ems, enl = e.timing(Engine.FULL)
        terms = sum(len(wl.temperatures[s]) * wl.pairs_full(s) for s in range(len(wl.systems)))
        ek = ems / max(enl, 1)
        full = {"calls": args.energy_calls, "kernel_avg_ms": ek,
                "pair_terms_per_s": terms / (ek / 1e3),
                "roofline_frac": terms / (ek / 1e3) * FLOPS_PER_PAIR / 1e12 / PEAK_FP32_TFLOPS,
                "calls_per_s_wall": args.energy_calls / (t9 - t8),
                "kernel": "k_energy_full + k_sum_partials (full FP32-Kahan energy of every replica, FP64 totals)",
                "work_per_call": "%d inter-body pair terms over all replicas" % terms}
        log("full energy: %.3f ms per refresh of all replicas (kernels)" % ek)

    # ------------------------------------------------------------ cpu ----
    cpu = None
    threads = args.cpu_threads or min(R, os.cpu_count() or 1)
    if not args.no_cpu and single:
        cpu = cpu_baseline(wl, rec_n, threads, args.cpu_single_replicas, gpu_records)
        log("cpu: %.2f s (%d threads)" % (cpu["seconds"], threads))
        cpu["energy_check"] = energy_check(e, wl, range(0, R, max(1, R // 4)))
        log("energy check:", cpu["energy_check"])
    e.close()

    c1 = None
    if not args.no_config1_full and not args.no_cpu:
        c1 = config1_full(args.cpu_threads or os.cpu_count() or 1)
        log("config1 full:", c1)
    acc = accuracy_all_configs(args.accuracy_steps) if args.accuracy_steps > 0 else None

    speedup = None
    if plain:
        speedup = {
            "kernel_vs_kernel_pair_interactions": kernel_pairs_per_s / plain["kernel_pair_interactions_per_s"],
            "force_biased_kernel_vs_plain_kernel_dE_pairs":
                (fb["kernel_dE_pair_interactions_per_s"] / plain["kernel_pair_interactions_per_s"]) if fb else None,
            "force_biased_kernel_vs_plain_kernel_pair_evaluations":
                (2.0 * fb["kernel_dE_pair_interactions_per_s"] / plain["kernel_pair_interactions_per_s"])
                if fb else None,
            "protocol_inclusive_pair_interactions": value / plain["pair_interactions_per_s"],
            "protocol_inclusive_mc_steps": (steps_total / wall) / plain["mc_steps_per_s"],
            "note": "kernel_vs_kernel: k_mc pair terms per second of kernel time over the plain tiled kernel's; "
                    "protocol_inclusive: whole-job rates, the plain one including its host proposal, copies and "
                    "sync every step",
        }

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / rounds,
        "step": "one round: %d MC steps of every replica + one exchange round (one hipGraph replay)" % spl,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": "f32 pair terms, Kahan per thread, f64 totals/Metropolis/exchange",
        "data": "synthetic (seeded self-avoiding chains, uniform residue types, seeded [-0.6,0] kT contact table)",
        "config": {"workload": workload_label(cid, wl), "replicas": R, "steps_per_exchange": spl,
                   "timed_mc_steps_per_replica": pl["timed_mc_steps_per_replica"],
                   "mc_block": mc_block, "mc_workgroups_per_replica": mc_coop, "form": "batched persistent dE kernel, hipGraph rounds"},
        "mc_steps_per_s": steps_total / wall,
        "pair_interactions_per_s": value,
        "kernel_pair_interactions_per_s": kernel_pairs_per_s,
        "step_cost": step_cost,
        "roofline": roofline,
        "plain": plain,
        "speedup_vs_plain": speedup,
        "cpu_baseline": cpu,
        "config1_full": c1,
        "accuracy": acc,
        "force_biased": fb,
        "adjoint": adj,
        "full_energy": full,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
