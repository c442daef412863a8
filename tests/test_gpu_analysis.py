"""Device analysis reductions through the C ABI against numpy restatements
of the reference's post-processing: DRMS ("post processor/pdbgensrc/
drms.cpp:7-58") and the centroid g(r) histogram ("post processor/gr.py:
93-140").  Bars: DRMS within 1e-6 relative (FP32 distances, the same
operation sequence); histogram counts exact (FP64 distances with the same
operation sequence as the numpy restatement)."""
import numpy as np
import pytest

from helpers import engine_for, say
from remcdockingsimulator_amd import synth
from remcdockingsimulator_amd._native import RemcError

pytestmark = pytest.mark.gpu


def drms_numpy(pos, ref, body, crowder):
    """Mean over bead pairs of different non-crowder bodies of |d - d_ref|,
    distances in float32 as Vector3f::magnitude computes them."""
    keep = ~crowder[body].astype(bool)
    p, q, b = pos[keep].astype(np.float32), ref[keep].astype(np.float32), body[keep]
    tot, n = 0.0, 0
    for i in range(len(p) - 1):
        dp = p[i + 1:] - p[i]
        dq = q[i + 1:] - q[i]
        d = np.sqrt((dp[:, 0] * dp[:, 0] + dp[:, 1] * dp[:, 1]) + dp[:, 2] * dp[:, 2]).astype(np.float32)
        e = np.sqrt((dq[:, 0] * dq[:, 0] + dq[:, 1] * dq[:, 1]) + dq[:, 2] * dq[:, 2]).astype(np.float32)
        other = b[i + 1:] != b[i]
        tot += float(np.sum(np.abs((d - e).astype(np.float32))[other].astype(np.float64)))
        n += int(np.sum(other))
    return tot / n


def round_half_away(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def hist_numpy(cen, btype, n_types, nbin, lo, hi, box):
    c = cen.astype(np.float64)
    dr = (hi - lo) / nbin
    out = np.zeros((n_types, n_types, nbin), dtype=np.uint64)
    M = len(c)
    for a in range(M):
        ta = btype[a]
        if ta < 0:
            continue
        for b in range(M):
            tb = btype[b]
            if b == a or tb < 0 or tb < ta:
                continue
            d = c[a] - c[b]
            if box > 0:
                d = d - box * round_half_away(d / box)
            r = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
            f = np.floor((r - lo) / dr)
            if 0 <= f < nbin:
                out[ta, tb, int(f)] += 1
    return out


def test_drms_against_numpy():
    wl = synth.config(2)
    sa = wl.systems[0]
    e = engine_for(wl, seed=4)
    e.mc_steps(300)
    e.synchronize()
    body = np.repeat(np.arange(sa.n_bodies), np.diff(sa.body_start))
    got = e.drms(0, sa.pos)
    for r in (0, 5, 19):
        say("drms replica", r)
        want = drms_numpy(e.state(r)["pos"], sa.pos, body, sa.crowder)
        assert abs(got[r] - want) <= 1e-6 * max(want, 1e-3), (r, got[r], want)
    # the starting conformation against itself
    f = engine_for(wl, seed=4)
    assert np.all(f.drms(0, sa.pos) == 0.0)
    e.close()
    f.close()


def test_centre_histogram_against_numpy():
    wl = synth.config(2)
    sa = wl.systems[0]
    e = engine_for(wl, seed=6)
    e.mc_steps(200)
    e.synchronize()
    # types: the two chains, crowders; one crowder left out
    btype = np.where(sa.crowder.astype(bool), 2, np.arange(sa.n_bodies) % 2).astype(np.int32)
    btype[5] = -1
    n_types, nbin, lo, hi = 3, 75, 0.0, sa.box / 2.0 - 1.0
    got = e.centre_histogram(0, btype, n_types, nbin, lo, hi)
    want = np.zeros_like(got)
    for r in range(wl.n_replicas):
        want += hist_numpy(e.state(r)["center"], btype, n_types, nbin, lo, hi, sa.box)
    assert np.array_equal(got, want)
    assert got.sum() > 0 and np.all(got[1, 0] == 0) and np.all(got[2, 1] == 0)
    with pytest.raises(RemcError):
        e.centre_histogram(0, btype, n_types, 5000, lo, hi)  # beyond the bin limit
    e.close()
