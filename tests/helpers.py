"""Shared helpers of the GPU parity tests: build an Engine and the matching
oracle replicas from the same workload."""
import sys

import numpy as np

from remcdockingsimulator_amd import forcefield as ffm
from remcdockingsimulator_amd import synth
from remcdockingsimulator_amd.engine import Engine


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def engine_for(wl, seed=0x1234, block=None):
    e = Engine(0)
    synth.load_engine(e, wl, seed=seed)
    if block:
        e.set_mc_block(block)
    return e


def kT(T):
    return ffm.R_GAS * T / 4184.0


def compare_decisions(gpu_rec, orc_rec, T, guard_kT=1e-4, delta_tol_kT=1e-4, rel_tol=0.0):
    """Per-step decision parity.  Returns (n_equal_prefix, diverged).  Bodies and
    move types must match at every step up to a divergence; a divergence is
    allowed only at a step whose oracle decision margin lies inside the guard
    band (the FP32-vs-FP64 error band).  rel_tol widens both bars by that
    fraction of |dE| (steep contacts, where FP32 pair terms carry a relative
    error of a few 1e-6)."""
    n = len(orc_rec)
    for k in range(n):
        assert gpu_rec["body"][k] == orc_rec["body"][k], ("body", k)
        assert gpu_rec["move"][k] == orc_rec["move"][k], ("move", k)
        band = rel_tol * abs(orc_rec["delta"][k]) / kT(T)
        d_err = abs(gpu_rec["delta"][k] - orc_rec["delta"][k]) / kT(T)
        assert d_err < delta_tol_kT + band, ("delta", k, gpu_rec["delta"][k], orc_rec["delta"][k])
        if gpu_rec["accepted"][k] != orc_rec["accepted"][k]:
            assert orc_rec["margin"][k] < guard_kT + band, ("decision", k, orc_rec["margin"][k])
            return k, True
    return n, False
