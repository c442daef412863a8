#!/usr/bin/env python3
"""Benchmark of the REMC hot path on one MI355X.

A "step" is one MC step of every replica of the workload (proposal, dE of the
moved body, Metropolis), with a temperature-exchange round every
--steps-per-launch steps; the batched path runs them as hipGraph replays of
(persistent MC kernel + exchange kernel).  Inputs are resident in HBM before
the timed region.

Default workload: config 4 (10 x 100-residue chains + 300 crowders = 1300
beads in 310 rigid bodies, 32 replicas 288-480 K), the configuration the
throughput target is stated on.  Reported beside it: the plain path (the
reference's per-step protocol with the reference-shaped tiled kernel) on the
same workload, and the CPU baseline (the FP64 C restatement of the reference
CPU path, full recomputation per step, replicas block-partitioned over
pthreads) on a bounded sample.

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

from remcdockingsimulator_amd import forcefield as ffm  # noqa: E402
from remcdockingsimulator_amd import rng as rngm  # noqa: E402
from remcdockingsimulator_amd import synth  # noqa: E402
from remcdockingsimulator_amd.engine import Engine  # noqa: E402

METRIC = "MC steps/sec & pair-interactions/sec, batched vs plain, |ΔE|<1e-4 kT vs FP64"
FLOPS_PER_PAIR = 41          # fixed counting rule per pair term (FMA = 2, sqrt/rcp/exp/rint = 1)
PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector peak (datasheet)
SEED = 0x52454D43


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def cpu_baseline(wl, nsteps, threads, gpu_records):
    """The oracle (FP64 C restatement of the reference CPU path, full energy
    recomputation per step, Replica.cpp:435) timed on this host, plus the
    |dE| comparison of the GPU's recorded steps against its FP64 values."""
    from oracle import pyoracle

    sa = wl.systems[0]
    lad = wl.temperatures[0]
    osys = pyoracle.OracleSystem(sa, wl.ff)
    reps = [pyoracle.OracleReplica(osys, lad[r], wl.translate_step, wl.rotate_step, r) for r in range(len(lad))]
    secs, rec = pyoracle.mc_threaded(osys, reps, SEED, nsteps, 1, threads, record=True)
    n_steps = len(reps) * nsteps
    pairs = n_steps * wl.pairs_full()
    worst = 0.0
    agree = 0
    total = 0
    for r in range(len(lad)):
        g = gpu_records[r][:nsteps]
        o = rec[r]
        kt = ffm.kT(lad[r])
        for k in range(nsteps):
            total += 1
            if g["body"][k] != o["body"][k] or g["accepted"][k] != o["accepted"][k]:
                break  # chains diverged at a near-threshold decision: stop comparing
            worst = max(worst, abs(g["delta"][k] - o["delta"][k]) / kt)
            agree += 1
    return {
        "value": pairs / secs,
        "unit": "pair-interactions/s",
        "mc_steps_per_s": n_steps / secs,
        "cores": threads,
        "kind": "port",
        "sample": "%d replicas x %d MC steps of the same workload, full-energy recomputation per step "
                  "(%d pair terms each), FP64, %d pthreads" % (len(reps), nsteps, wl.pairs_full(), threads),
        "seconds": secs,
        "max_abs_dE_err_kT": worst,
        "decisions_compared": total,
        "decisions_equal": agree,
    }


def energy_check(e, wl, replicas):
    """Full energies of the device state after the run: the device's full
    evaluation and its tracked (sum of accepted dE) energy against the
    oracle's FP64 full energy of the same conformations."""
    from oracle import pyoracle

    sa = wl.systems[0]
    osys = pyoracle.OracleSystem(sa, wl.ff)
    full = e.energy_all()
    tracked = e.energies()
    temps = e.temperatures()
    rel, trk = 0.0, 0.0
    for r in replicas:
        want = osys.energy(e.state(r)["pos"])
        rel = max(rel, abs(full[r] - want) / max(abs(want), 1e-12))
        trk = max(trk, abs(tracked[r] - want) / ffm.kT(temps[r]))
    return {"replicas": list(replicas), "max_rel_E_err_full": rel, "max_abs_tracked_E_err_kT": trk}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000, help="timed MC steps per replica")
    ap.add_argument("--warmup", type=int, default=1000, help="untimed MC steps per replica (recorded)")
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--steps-per-launch", type=int, default=1000, help="MC steps between exchanges")
    ap.add_argument("--plain-steps", type=int, default=300)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--block", type=int, default=0, help="MC workgroup size (0 = autotune)")
    ap.add_argument("--coop", type=int, default=0, help="MC workgroups per replica (0 = autotune)")
    ap.add_argument("--fb-steps", type=int, default=2000,
                    help="force-biased MC steps per replica timed after the main run (0 = skip)")
    ap.add_argument("--adjoint-calls", type=int, default=20,
                    help="adjoint (gradient) calls timed after the main run (0 = skip)")
    ap.add_argument("--energy-calls", type=int, default=20,
                    help="full-energy refreshes of every replica timed after the main run (0 = skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: every replica of the workload runs on a single MI355X; "
              "--gpus %d is not supported (multi-GPU runs are out of scope)" % args.gpus, file=sys.stderr)
        sys.exit(2)

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
    spl = args.steps_per_launch
    if args.steps % spl:
        raise SystemExit("--steps must be a multiple of --steps-per-launch")
    rounds = args.steps // spl
    sizes = [np.diff(sa.body_start) for sa in wl.systems]

    e = Engine(0)
    synth.load_engine(e, wl, seed=SEED)
    if args.block:
        e.set_mc_block(args.block)
    if args.coop:
        e.set_mc_coop(args.coop)
    rec_n = max(args.warmup, 1)
    e.set_recording(rec_n)
    log("workload:", workload_label(cid, wl))
    e.mc_steps(args.warmup)
    e.synchronize()
    gpu_records = [e.records(r) for r in range(R)] if len(wl.systems) == 1 else None
    step_cost = None
    if gpu_records is not None:
        sizes0 = sizes[0]
        allrec = np.concatenate(gpu_records)
        nm = sizes0[allrec["body"]]
        small = nm * (sizes0.sum() - nm) < 8192
        us = allrec["ticks"] / 100.0  # 100 MHz device clock
        step_cost = {"small_move_us": float(np.mean(us[small])) if small.any() else None,
                     "large_move_us": float(np.mean(us[~small])) if (~small).any() else None,
                     "large_move_fraction": float(np.mean(~small))}
        log("per-step cost (warmup, recorded):", step_cost)
    e.set_recording(0)
    e.run(1, spl)          # capture + one untimed replay
    e.synchronize()
    warm_steps = args.warmup + spl
    log("mc block:", e.mc_block(), "workgroups per replica:", e.mc_coop())

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
        pairs += rngm.delta_pairs(SEED, range(rid, rid + n_rep), warm_steps, args.steps, sizes[s])
        rid += n_rep
    value = pairs / wall
    steps_total = R * args.steps
    achieved = pairs * FLOPS_PER_PAIR / (kms / 1e3) / 1e12
    traffic = read_traffic("k_mc", "config%d" % cid)
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                "kernel": "k_mc (persistent batched MC, dE of the moved body)",
                "kernel_avg_ms": kms / max(nl, 1), "launches": nl,
                "flops_per_pair": FLOPS_PER_PAIR, "pairs_per_launch": pairs / max(nl, 1)}
    energies = e.energies()
    log("timed: %.3f s, %d launches, kernel %.3f ms/launch" % (wall, nl, kms / max(nl, 1)))

    # ---------------------------------------------------------- plain ----
    plain = None
    if not args.no_plain and len(wl.systems) == 1:
        e.timing_enable(True)
        e.synchronize()
        t2 = time.perf_counter()
        e.mc_steps_plain(0, args.plain_steps)
        t3 = time.perf_counter()
        pms, pnl = e.timing(Engine.PLAIN)
        pw = t3 - t2
        pf = wl.pairs_full()
        plain = {"pair_interactions_per_s": args.plain_steps * pf / pw,
                 "mc_steps_per_s": args.plain_steps / pw,
                 "kernel_avg_ms": pms / max(pnl, 1),
                 "kernel_pair_interactions_per_s": pf / (pms / max(pnl, 1) / 1e3),
                 "protocol": "reference per-step protocol: host proposal, H2D moved body, tiled kernel "
                             "(lower triangle culled, FP32 partials), D2H partials, host FP64 sum and Metropolis",
                 "steps": args.plain_steps}
        log("plain: %.3f s for %d steps" % (pw, args.plain_steps))

    # ---------------------------------------------------- force-biased ----
    fb = None
    if args.fb_steps > 0 and len(wl.systems) == 1:
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
        fb = {"mc_steps_per_s": R * args.fb_steps / fw,
              "dE_pair_interactions_per_s": fpairs / fw,
              "pair_evaluations_per_s": 2.0 * fpairs / fw,
              "kernel": "k_mc_fb (force-biased MC, 512-thread workgroups, as many per replica as the MC kernel; "
                        "large moves split over them)",
              "kernel_avg_ms": fms / max(fnl, 1), "launches": fnl, "steps": args.fb_steps,
              "work_per_step": "force at the old position, energy and force at the new, energy at the old: "
                               "2 n_m (N - n_m) pair energies and 2 n_m (N - n_m) pair derivatives"}
        log("fb: %.3f s for %d steps" % (fw, args.fb_steps))

    # -------------------------------------------------------- adjoint ----
    adj = None
    if args.adjoint_calls > 0 and len(wl.systems) == 1:
        e.adjoint(0)  # warm
        e.timing_enable(True)
        e.synchronize()
        t6 = time.perf_counter()
        for k in range(args.adjoint_calls):
            e.adjoint(k % R)
        t7 = time.perf_counter()
        ams, anl = e.timing(Engine.ADJOINT)
        # every ordered inter-body bead pair once: 2 P_full pair derivatives per call
        dpairs = 2 * wl.pairs_full()
        adj = {"calls": args.adjoint_calls, "kernel_avg_ms": ams / max(anl, 1),
               "pair_derivatives_per_s": dpairs / (ams / max(anl, 1) / 1e3),
               "calls_per_s_wall": args.adjoint_calls / (t7 - t6),
               "kernel": "k_adjoint_pairs + k_adjoint_reduce + k_adjoint_bodies (forces, torques, dE/d{lambda,e0,xi} of one replica)",
               "work_per_call": "2 P_full ordered pair derivatives (P_full = %d)" % wl.pairs_full()}
        log("adjoint: %.3f ms per call (kernels)" % adj["kernel_avg_ms"])

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
        full = {"calls": args.energy_calls, "kernel_avg_ms": ems / max(enl, 1),
                "pair_terms_per_s": terms / (ems / max(enl, 1) / 1e3),
                "calls_per_s_wall": args.energy_calls / (t9 - t8),
                "kernel": "k_energy_full + k_sum_partials (full FP32-Kahan energy of every replica, FP64 totals)",
                "work_per_call": "%d inter-body pair terms over all replicas" % terms}
        log("full energy: %.3f ms per refresh of all replicas (kernels)" % full["kernel_avg_ms"])

    # ------------------------------------------------------------ cpu ----
    cpu = None
    if not args.no_cpu and len(wl.systems) == 1 and gpu_records is not None:
        threads = args.cpu_threads or min(R, os.cpu_count() or 1)
        cpu = cpu_baseline(wl, min(args.cpu_steps, rec_n), threads, gpu_records)
        log("cpu: %.2f s" % cpu["seconds"])
        cpu["energy_check"] = energy_check(e, wl, range(0, R, max(1, R // 4)))
        log("energy check:", cpu["energy_check"])

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair terms, Kahan per thread, f64 totals/Metropolis",
        "data": "synthetic (seeded self-avoiding chains, uniform residue types, seeded [-0.6,0] kT contact table)",
        "config": {"workload": workload_label(cid, wl), "replicas": R, "steps_per_exchange": spl,
                   "mc_block": e.mc_block(), "mc_workgroups_per_replica": e.mc_coop(),
                   "form": "batched persistent dE kernel, hipGraph rounds"},
        "mc_steps_per_s": steps_total / wall,
        "pair_interactions_per_s": value,
        "step_cost": step_cost,
        "roofline": roofline,
        "plain": plain,
        "speedup_vs_plain": {"pair_interactions": value / plain["pair_interactions_per_s"],
                             "mc_steps": (steps_total / wall) / plain["mc_steps_per_s"],
                             "force_biased_pair_evaluations": (fb["pair_evaluations_per_s"]
                                                               / plain["pair_interactions_per_s"]) if fb else None}
        if plain else None,
        "cpu_baseline": cpu,
        "force_biased": fb,
        "adjoint": adj,
        "full_energy": full,
        "energy_mean": float(np.mean(energies)),
    }
    e.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
