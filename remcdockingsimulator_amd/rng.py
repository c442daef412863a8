"""Vectorised host mirror of the device RNG draw layout (numpy).

The kernels draw per MC step two Philox4x32-10 blocks (Salmon et al., SC'11)
at counter (step, replica, 'MC'+0/1) under the context seed.  This mirror lets
host code predict the body/move sequence of any replica -- used to count the
exact pair terms a run evaluates.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
TAG_MC = 0x4D430000
TAG_FB = 0x46420000
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint32) for c in (c0, c1, c2, c3))
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    c0, c1, c2, c3 = c0.copy(), c1.copy(), c2.copy(), c3.copy()
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            n0 = (p1 >> np.uint64(32)).astype(np.uint32) ^ c1 ^ k0
            n2 = (p0 >> np.uint64(32)).astype(np.uint32) ^ c3 ^ k1
            c1 = (p1 & MASK).astype(np.uint32)
            c3 = (p0 & MASK).astype(np.uint32)
            c0, c2 = n0, n2
            k0 = np.uint32((int(k0) + int(W0)) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(W1)) & 0xFFFFFFFF)
    return c0, c1, c2, c3


def step_bodies(seed, rng_id, first_step, nsteps, n_bodies, tag=TAG_MC):
    """Body index and move type (1 = rotate) of steps [first_step, first_step+nsteps).
    tag=TAG_MC: the uniform-proposal MC (body from word 0, move from word 1 of
    the first block); tag=TAG_FB: the force-biased MC (the same two words)."""
    s = np.arange(first_step, first_step + nsteps, dtype=np.uint64)
    lo = (s & MASK).astype(np.uint32)
    hi = (s >> np.uint64(32)).astype(np.uint32)
    w0, w1, _, _ = philox4x32_10(lo, hi, np.uint32(rng_id), np.uint32(tag), seed & 0xFFFFFFFF,
                                 (seed >> 32) & 0xFFFFFFFF)
    body = ((w0.astype(np.uint64) * np.uint64(n_bodies)) >> np.uint64(32)).astype(np.int64)
    rotate = (w1.astype(np.float64) * (1.0 / 4294967296.0)) < 0.5
    return body, rotate


def delta_pairs(seed, rng_ids, first_step, nsteps, body_sizes, tag=TAG_MC):
    """Pair terms evaluated by dE kernels (old and new positions) over a run:
    sum over replicas and steps of 2*n_m*(N - n_m)."""
    sizes = np.asarray(body_sizes, dtype=np.int64)
    n = int(sizes.sum())
    total = 0
    for r in rng_ids:
        body, _ = step_bodies(seed, r, first_step, nsteps, len(sizes), tag)
        nm = sizes[body]
        total += int(np.sum(2 * nm * (n - nm)))
    return total
