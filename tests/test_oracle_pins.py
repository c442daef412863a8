"""Pins for the oracle functions that only the GPU tests used to read:

- orc_energy_bodies (Replica::E(Molecule*,Molecule*), Replica.cpp:793-828):
  the energy is additive over body pairs, so the sum over every a < b of the
  pair-of-bodies energy equals the full energy (itself pinned by the
  reference's fixtures and a long-double brute force);
- orc_energy_noncrowder (the E_complex sum of Replica::sample,
  Replica.cpp:1723-1743): the full energy of the same conformation with the
  crowder bodies removed, and independent of where the crowders sit;
- orc_body_force (the force-biased proposal's force and torque): the per-body
  rows of orc_gradient, which FP64 central finite differences pin
  (test_oracle_adjoint.py), also for a body displaced from the positions
  passed for the rest of the system.

Each pin is shown to fail for a plausible slip planted in a copy of the
oracle's value on the CPU (crowder pairs included, a flipped torque, a body
range off by one)."""
import ctypes
import dataclasses

import numpy as np
import pytest

from conftest import conformation
from remcdockingsimulator_amd import forcefield as ffm
from remcdockingsimulator_amd import synth


def crowded_system(seed):
    """A small periodic system whose crowders touch the chains (non-zero
    crowder terms), so that including a crowder pair changes the result."""
    sa = synth.small_system(seed, n_chains=3, chain_len=(10, 24), n_crowders=8, box=45.0)
    return sa


def without_crowders(sa):
    keep = [m for m in range(sa.n_bodies) if not sa.crowder[m]]
    idx = np.concatenate([np.arange(sa.body_start[m], sa.body_start[m + 1]) for m in keep])
    sizes = [sa.body_start[m + 1] - sa.body_start[m] for m in keep]
    return dataclasses.replace(
        sa, body_start=np.cumsum([0] + sizes).astype(np.int32), type=sa.type[idx], charge=sa.charge[idx],
        sigma=sa.sigma[idx], crowder=sa.crowder[keep], pos=sa.pos[idx], rel=sa.rel[idx],
        center=sa.center[keep], quat=sa.quat[keep])


def crowder_energy(osys, sa):
    """Energy of every pair that involves a crowder body (full energy minus the
    crowder-free system's)."""
    return osys.energy(sa.pos) - sum(osys.energy_bodies(sa.pos, a, b)
                                     for a in range(sa.n_bodies) for b in range(a + 1, sa.n_bodies)
                                     if not sa.crowder[a] and not sa.crowder[b])


# ------------------------------------------------------ energy_bodies ----
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_energy_additive_over_body_pairs_synthetic(oracle, seed):
    sa = crowded_system(seed)
    osys = oracle.OracleSystem(sa, ffm.ForceField.synthetic(seed))
    full = osys.energy(sa.pos)
    total = sum(osys.energy_bodies(sa.pos, a, b) for a in range(sa.n_bodies) for b in range(a + 1, sa.n_bodies))
    assert abs(total - full) <= 1e-12 * max(abs(full), 1.0), (total, full)
    # each unordered pair once: E(a,b) = E(b,a), E(a,a) is not part of the sum
    assert abs(osys.energy_bodies(sa.pos, 0, 1) - osys.energy_bodies(sa.pos, 1, 0)) <= 1e-12 * (
        abs(osys.energy_bodies(sa.pos, 0, 1)) + 1e-12)


@pytest.mark.parametrize("idx", [1, 6])
def test_energy_additive_over_body_pairs_fixture(oracle, aas, mj_ff, idx):
    """The reference's UIM/ubiquitin conformations under MJ1996: two bodies,
    so the single pair's energy is the whole energy."""
    sa = conformation(idx, aas, box=300.0)
    osys = oracle.OracleSystem(sa, mj_ff)
    full = osys.energy(sa.pos)
    assert abs(osys.energy_bodies(sa.pos, 0, 1) - full) <= 1e-12 * abs(full)


def test_planted_body_range_slip_fails_additivity(oracle):
    """A pair energy over body b's range shifted by one bead (a plausible
    indexing slip, planted by shifting the positions it reads) breaks the sum."""
    sa = crowded_system(4)
    osys = oracle.OracleSystem(sa, ffm.ForceField.synthetic(4))
    shifted = np.roll(sa.pos, -1, axis=0)  # bead j reads bead j+1's position
    total_slip = 0.0
    for a in range(sa.n_bodies):
        for b in range(a + 1, sa.n_bodies):
            p = sa.pos.copy()
            b0, b1 = sa.body_start[b], sa.body_start[b + 1]
            p[b0:b1] = shifted[b0:b1]
            total_slip += osys.energy_bodies(p, a, b)
    full = osys.energy(sa.pos)
    assert abs(total_slip - full) > 1e-6 * max(abs(full), 1.0)


# -------------------------------------------------- energy_noncrowder ----
def test_noncrowder_equals_full_without_crowders(oracle):
    sa = synth.small_system(7, n_chains=4, chain_len=(8, 20), n_crowders=0, box=50.0)
    osys = oracle.OracleSystem(sa, ffm.ForceField.synthetic(7))
    full = osys.energy(sa.pos)
    assert abs(osys.energy_noncrowder(sa.pos) - full) <= 1e-12 * max(abs(full), 1.0)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_noncrowder_equals_full_energy_of_the_chains_alone(oracle, seed):
    sa = crowded_system(seed)
    ff = ffm.ForceField.synthetic(seed)
    osys = oracle.OracleSystem(sa, ff)
    sub = without_crowders(sa)
    want = oracle.OracleSystem(sub, ff).energy(sub.pos)
    got = osys.energy_noncrowder(sa.pos)
    assert abs(got - want) <= 1e-12 * max(abs(want), 1.0), (got, want)
    # the crowders do interact here, so a sum that included their pairs would differ
    planted = got + crowder_energy(osys, sa)
    assert abs(crowder_energy(osys, sa)) > 1e-3
    assert abs(planted - want) > 1e-6 * max(abs(want), 1.0)


def test_noncrowder_ignores_crowder_positions(oracle):
    sa = crowded_system(5)
    osys = oracle.OracleSystem(sa, ffm.ForceField.synthetic(5))
    ref = osys.energy_noncrowder(sa.pos)
    rng = np.random.default_rng(5)
    p = sa.pos.copy()
    for m in range(sa.n_bodies):
        if sa.crowder[m]:
            b0, b1 = sa.body_start[m], sa.body_start[m + 1]
            p[b0:b1] += rng.uniform(-5, 5, 3).astype(np.float32)
    assert osys.energy_noncrowder(p) == ref
    assert osys.energy(p) != osys.energy(sa.pos)


# ------------------------------------------------------- body_force ----
def body_force(oracle, osys, pos, m, bpos, brel):
    L = oracle.lib()
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    cols = [f32(pos[:, k]) for k in range(3)] + [f32(bpos[:, k]) for k in range(3)] + [f32(brel[:, k])
                                                                                      for k in range(3)]
    ft = np.zeros(6)
    fp = ctypes.POINTER(ctypes.c_float)
    L.orc_body_force(ctypes.byref(osys.s), ctypes.byref(osys.ff), cols[0].ctypes.data_as(fp),
                     cols[1].ctypes.data_as(fp), cols[2].ctypes.data_as(fp), m,
                     *[c.ctypes.data_as(fp) for c in cols[3:]], ft.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return ft


def body_of(sa):
    return np.repeat(np.arange(sa.n_bodies), np.diff(sa.body_start))


def check_body_forces(oracle, osys, sa, pos, bodies):
    """orc_body_force against orc_gradient's force/torque rows (FP32 positions
    used for both; rel = position - centre)."""
    pos32 = np.asarray(pos, dtype=np.float32)
    # lever arms as the FP32 values orc_body_force reads
    rel = (pos32.astype(np.float64) - sa.center[body_of(sa)].astype(np.float64)).astype(np.float32).astype(np.float64)
    g, ft_grad, _ = osys.gradient(pos32.astype(np.float64), rel)
    scale = np.max(np.abs(ft_grad)) + 1e-300
    out = []
    for m in bodies:
        b0, b1 = sa.body_start[m], sa.body_start[m + 1]
        ft = body_force(oracle, osys, pos32, m, pos32[b0:b1], rel[b0:b1])
        assert np.max(np.abs(ft - ft_grad[m])) <= 1e-12 * scale, (m, ft, ft_grad[m])
        # and the bead gradients sum to the force: F = -sum_i dE/dx_i
        assert np.max(np.abs(ft[:3] + g[b0:b1].sum(0))) <= 1e-12 * scale
        out.append((ft, ft_grad[m]))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_body_force_equals_fd_pinned_gradient_synthetic(oracle, seed):
    sa = crowded_system(seed)
    osys = oracle.OracleSystem(sa, ffm.ForceField.synthetic(seed))
    check_body_forces(oracle, osys, sa, sa.pos, range(sa.n_bodies))


@pytest.mark.parametrize("idx", [3, 8])
def test_body_force_equals_fd_pinned_gradient_fixture_mj(oracle, aas, mj_ff, idx):
    sa = conformation(idx, aas, box=300.0)
    osys = oracle.OracleSystem(sa, mj_ff)
    check_body_forces(oracle, osys, sa, sa.pos, [0, 1])


def test_body_force_at_displaced_pose_ignores_stale_self_positions(oracle):
    """The force-biased move evaluates the force at the proposed pose while
    the system arrays still hold the old one: only the passed body beads are
    read for body m."""
    sa = crowded_system(6)
    osys = oracle.OracleSystem(sa, ffm.ForceField.synthetic(6))
    m = 0
    b0, b1 = sa.body_start[m], sa.body_start[m + 1]
    moved = sa.pos.copy()
    moved[b0:b1] += np.float32([1.5, -0.75, 0.5])
    rel = (moved.astype(np.float64) - sa.center[body_of(sa)].astype(np.float64)).astype(np.float32).astype(np.float64)
    _, ft_grad, _ = osys.gradient(moved.astype(np.float64), rel)
    ft = body_force(oracle, osys, sa.pos, m, moved[b0:b1], rel[b0:b1])  # stale body m in the system arrays
    assert np.max(np.abs(ft - ft_grad[m])) <= 1e-12 * (np.max(np.abs(ft_grad)) + 1e-300)


def test_planted_torque_sign_fails(oracle):
    sa = crowded_system(2)
    osys = oracle.OracleSystem(sa, ffm.ForceField.synthetic(2))
    (ft, want), = check_body_forces(oracle, osys, sa, sa.pos, [0])
    planted = ft.copy()
    planted[3:] = -planted[3:]
    assert np.max(np.abs(want[3:])) > 1e-6
    assert np.max(np.abs(planted - want)) > 1e-6 * np.max(np.abs(want))
