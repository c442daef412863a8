"""GPU energies through the C ABI against the FP64 oracle.

Bar: 1e-5 relative on per-replica energies (absolute floor 1e-7 kcal/mol for
near-zero totals), the tolerance this path states for FP32 pair terms with
compensated sums and FP64 totals."""
import numpy as np
import pytest

from conftest import conformation
from helpers import engine_for, say
from remcdockingsimulator_amd import model, synth
from remcdockingsimulator_amd.engine import Engine

pytestmark = pytest.mark.gpu

REL = 1e-5
ABS = 1e-7


def _engine_single(sa, ff, n_rep=1, temps=None):
    e = Engine(0)
    e.set_forcefield(ff)
    e.set_seed(1)
    sid = e.add_system(sa)
    temps = np.linspace(300.0, 400.0, n_rep) if temps is None else temps
    e.add_replicas(sid, temps, 2.0, 0.2)
    return e


@pytest.mark.parametrize("box", [0.0, 300.0])
def test_fixture_conformations_full_and_plain(aas, mj_ff, oracle, box):
    for idx in range(1, 11):
        sa = conformation(idx, aas, box=box)
        say("conformation", idx, "box", box)
        want = oracle.OracleSystem(sa, mj_ff).energy(sa.pos)
        e = _engine_single(sa, mj_ff)
        full = e.energy_all()[0]
        tracked = e.energies()[0]
        plain = e.energy_plain(0)
        for got in (full, tracked, plain):
            assert abs(got - want) <= REL * abs(want) + ABS, (idx, got, want)
        e.close()


@pytest.mark.parametrize("cid", [1, 2, 4])
def test_synthetic_configs_full_and_plain(oracle, cid):
    wl = synth.config(cid)
    sa = wl.systems[0]
    say("config", cid, "beads", sa.n_beads)
    want = oracle.OracleSystem(sa, wl.ff).energy(sa.pos)
    e = engine_for(wl)
    full = e.energy_all()
    assert np.all(np.abs(full - want) <= REL * abs(want) + ABS), (full[:4], want)
    for block in (32, 64, 128, 256):
        plain = e.energy_plain(0, block)
        assert abs(plain - want) <= REL * abs(want) + ABS, (block, plain, want)
    e.close()


def test_config5_full_energy(oracle):
    wl = synth.config(5)
    sa = wl.systems[0]
    say("config 5 beads", sa.n_beads)
    want = oracle.OracleSystem(sa, wl.ff).energy(sa.pos)
    e = engine_for(wl)
    full = e.energy_all()
    assert np.all(np.abs(full - want) <= REL * abs(want) + ABS)
    assert abs(e.energy_plain(0, 64) - want) <= REL * abs(want) + ABS
    e.close()


def test_full_energy_partner_cuts(oracle):
    """The config-4 system (6 bead tiles, 24 work units per replica) with 32,
    48 and 96 replicas: the launch cuts each unit's partner tile into 4, 2 and
    1 sub-ranges (fewer than 1024, 2048 and at least 2048 units).  Every
    replica matches the oracle and every cut gives the same totals."""
    wl = synth.config(4)
    sa = wl.systems[0]
    want = oracle.OracleSystem(sa, wl.ff).energy(sa.pos)
    totals = []
    for n_rep in (32, 48, 96):
        say("config 4 system, replicas", n_rep)
        e = _engine_single(sa, wl.ff, n_rep)
        full = e.energy_all()
        assert np.all(np.abs(full - want) <= REL * abs(want) + ABS), (n_rep, full[:4], want)
        assert np.all(full == full[0]), n_rep  # identical conformations, fixed summation order
        totals.append(full[0])
        e.close()
    assert max(totals) - min(totals) <= REL * abs(want) + ABS, totals


def test_ragged_systems_in_one_launch(oracle):
    """Several systems of different sizes (uneven systems of config 3's kind)
    evaluated in one batched launch."""
    from remcdockingsimulator_amd.forcefield import ForceField

    ff = ForceField.synthetic(21)
    systems = [synth.small_system(s, n_chains=2 + s % 4, chain_len=(10, 60), n_crowders=s * 7, box=90.0 + 10 * s)
               for s in range(6)]
    e = Engine(0)
    e.set_forcefield(ff)
    e.set_seed(3)
    for sa in systems:
        sid = e.add_system(sa)
        e.add_replicas(sid, [300.0, 320.0, 350.0], 2.0, 0.2)
    full = e.energy_all()
    for k, sa in enumerate(systems):
        want = oracle.OracleSystem(sa, ff).energy(sa.pos)
        say("system", k, "beads", sa.n_beads, "E", want)
        assert np.all(np.abs(full[3 * k:3 * k + 3] - want) <= REL * abs(want) + ABS)
    e.close()


def test_noncrowder_energy(oracle):
    wl = synth.config(2)
    sa = wl.systems[0]
    want = oracle.OracleSystem(sa, wl.ff).energy_noncrowder(sa.pos)
    e = engine_for(wl)
    got = e.energy_noncrowder_all()
    assert np.all(np.abs(got - want) <= REL * abs(want) + ABS)
    e.close()


def test_largest_body_and_limit(oracle):
    from remcdockingsimulator_amd.forcefield import ForceField

    rng = np.random.default_rng(2)
    ff = ForceField.synthetic(2)
    body = synth.Body(rng.uniform(-60, 60, (1024, 3)), rng.integers(0, 20, 1024), False)
    sa = synth.build_system([body, synth.crowder_body(), synth.chain_body(20, rng)], 400.0, rng)
    want = oracle.OracleSystem(sa, ff).energy(sa.pos)
    e = _engine_single(sa, ff, n_rep=2)
    assert np.all(np.abs(e.energy_all() - want) <= REL * abs(want) + ABS)
    e.mc_steps(20)
    e.synchronize()
    e.close()
    big = synth.Body(rng.uniform(-60, 60, (1025, 3)), rng.integers(0, 20, 1025), False)
    sa2 = synth.build_system([big, synth.crowder_body()], 500.0, rng)
    e2 = Engine(0)
    e2.set_forcefield(ff)
    with pytest.raises(RuntimeError, match="limit"):
        e2.add_system(sa2)
    e2.close()


def test_overlapping_beads_give_inf_not_crash(oracle):
    """Coincident beads of different bodies: the pair term overflows (the
    reference's r = 0 + EPS gives inf); the kernels cap each pair at 1e30 so
    compensated sums stay finite, and nothing faults."""
    from remcdockingsimulator_amd.forcefield import ForceField

    ff = ForceField.synthetic(4)
    sa = synth.small_system(4, n_chains=2, chain_len=(5, 5), n_crowders=0, box=50.0)
    sa.pos[5] = sa.pos[0]
    e = _engine_single(sa, ff)
    got = e.energy_all()[0]
    assert np.isinf(got) or got > 1e20
    e.close()
