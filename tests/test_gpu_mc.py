"""GPU Monte Carlo and replica exchange through the C ABI against the oracle
under identical Philox streams.

Bars: bodies and move types identical at every step (RNG parity); per-step dE
within 1e-4 kT of the FP64 oracle; accept/reject decisions identical, except
at a step whose FP64 decision margin is inside the 1e-4 kT guard band (then
the chains may diverge from that step on and the comparison stops there);
positions bit-identical while decisions agree (the move arithmetic is shared
IEEE code); tracked energies within 1e-5 relative of the oracle's full
energy of the GPU's own final conformation."""
import numpy as np
import pytest

from helpers import compare_decisions, engine_for, kT, say
from remcdockingsimulator_amd import synth

pytestmark = pytest.mark.gpu

SEED = 0xC0FFEE


def run_parity(oracle, cid, nsteps, block=None, replicas=None, coop=None, window=True, wl=None):
    wl = synth.config(cid) if wl is None else wl
    sa = wl.systems[0]
    lad = wl.temperatures[0]
    e = engine_for(wl, seed=SEED, block=block)
    if coop:
        e.set_mc_coop(coop)
    if not window:
        e.set_mc_window(False)
    e.set_recording(nsteps)
    e.mc_steps(nsteps)
    e.synchronize()
    o = oracle.OracleSystem(sa, wl.ff)
    energies = e.energies()
    full = e.energy_all()
    reps = range(len(lad)) if replicas is None else replicas
    diverged = 0
    for r in reps:
        say("config", cid, "replica", r, "block", e.mc_block())
        grec = e.records(r)
        orep = oracle.OracleReplica(o, lad[r], wl.translate_step, wl.rotate_step, r)
        orec = orep.mc(seed=SEED, nsteps=nsteps, mode=0)
        n_eq, div = compare_decisions(grec, orec, lad[r])
        st = e.state(r)
        if not div:
            assert np.array_equal(st["pos"], orep.pos), r
            assert np.array_equal(st["rel"], orep.rel), r
            assert np.array_equal(st["center"], orep.center), r
            assert np.array_equal(st["quat"], orep.quat), r
            assert abs(energies[r] - orep.energy) < 1e-4 * kT(lad[r])
        else:
            diverged += 1
        want = o.energy(st["pos"])
        assert abs(full[r] - want) <= 1e-5 * abs(want) + 1e-7
        assert abs(energies[r] - want) <= 1e-5 * abs(want) + 1e-4 * kT(lad[r])
    e.close()
    return diverged


def test_mc_parity_config1(oracle):
    assert run_parity(oracle, 1, 2000) <= 1


def test_mc_parity_config2(oracle):
    assert run_parity(oracle, 2, 1500) <= 1


def test_mc_parity_config4(oracle):
    assert run_parity(oracle, 4, 300, replicas=range(0, 32, 4)) <= 1


@pytest.mark.parametrize("mode", ["per_type", "per_bead"])
def test_mc_sigma_per_type_and_per_bead(oracle, mode):
    """Config 4 with bead sigmas that differ between residue types (the
    large-move pass then reads sigma_ij^2 from the per-system type table) and
    with sigmas that differ bead by bead within a type (the pass forms
    sigma_ij^2 from the two beads' sigmas): both against the oracle."""
    wl = synth.config(4)
    sa = wl.systems[0]
    if mode == "per_type":
        sa.sigma = (5.0 + 0.25 * (sa.type % 7)).astype(np.float32)
    else:
        rng = np.random.default_rng(5)
        sa.sigma = (6.0 + rng.uniform(-0.5, 0.5, len(sa.sigma))).astype(np.float32)
    assert run_parity(oracle, 4, 300, replicas=[0, 13, 31], wl=wl) <= 1


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_mc_block_sizes_agree(oracle, block):
    assert run_parity(oracle, 2, 400, block=block, replicas=[0, 7, 19], coop=1) <= 1


@pytest.mark.parametrize("coop", [2, 4, 8])
def test_mc_cooperative_workgroups(oracle, coop):
    """Several workgroups per replica, each on a slice of the partner beads,
    exchanging partial dE every step: same decisions and states."""
    assert run_parity(oracle, 4, 300, block=512, replicas=range(0, 32, 8), coop=coop) <= 1
    assert run_parity(oracle, 1, 500, block=256, coop=coop) <= 1


@pytest.mark.parametrize("cid", [2, 4])
def test_mc_windows_off_matches_oracle(oracle, cid):
    """The one-step-at-a-time kernel (speculative windows switched off)."""
    assert run_parity(oracle, cid, 600 if cid == 2 else 300, block=1024, replicas=[0, 7, 19] if cid == 2 else [0, 31],
                      window=False) <= 1


def repeats_in_window(bodies, width):
    """Fraction of steps whose body was drawn exactly once in the previous
    width - 1 steps (a slot chained to an earlier slot of its window), and
    more than once (a slot that stands only if those are all rejected)."""
    once = twice = 0
    for k in range(len(bodies)):
        n = int(np.sum(bodies[max(0, k - width + 1):k] == bodies[k]))
        once += n == 1
        twice += n > 1
    return once / len(bodies), twice / len(bodies)


@pytest.mark.parametrize("block,coop", [(256, 1), (1024, 1), (1024, 2)])
def test_mc_windows_chained_repeats(oracle, block, coop):
    """Six small bodies in a dilute box, nearly every move accepted: a window
    often draws a body twice, and the later slot is staged from the earlier
    slot's proposal (chained), so the window goes on past it when that
    proposal is accepted; a third draw of the body falls back to ending the
    window.  Decisions, dE and the bit-exact states must match the oracle's
    one step at a time."""
    from remcdockingsimulator_amd.engine import Engine
    from remcdockingsimulator_amd.forcefield import ForceField

    rng = np.random.default_rng(5)
    ff = ForceField.synthetic(5)
    bodies = [synth.chain_body(6, rng), synth.chain_body(9, rng)] + [synth.crowder_body() for _ in range(4)]
    sa = synth.build_system(bodies, 70.0, rng)
    lad = [300.0, 420.0, 600.0]
    nsteps = 1200
    e = Engine(0)
    e.set_forcefield(ff)
    e.set_seed(17)
    sid = e.add_system(sa)
    e.add_replicas(sid, lad, 1.5, 0.3)
    e.set_mc_block(block)
    e.set_mc_coop(coop)
    e.set_recording(nsteps)
    e.mc_steps(nsteps)
    e.synchronize()
    o = oracle.OracleSystem(sa, ff)
    diverged = 0
    for r in range(3):
        say("chained windows block", block, "coop", coop, "replica", r)
        orep = oracle.OracleReplica(o, lad[r], 1.5, 0.3, r)
        orec = orep.mc(seed=17, nsteps=nsteps, mode=0)
        # over the first 8 slots of a window (windows hold 2 slots per wave)
        once, twice = repeats_in_window(orec["body"], 8)
        assert once > 0.15 and twice > 0.05, (once, twice)  # the draws exercise both kinds
        assert np.mean(orec["accepted"] != 0) > 0.8
        # the chains meet in steep contacts (dE of tens of kT), where FP32
        # pair terms differ from FP64 by a few 1e-6 relative
        n_eq, div = compare_decisions(e.records(r), orec, lad[r], rel_tol=1e-5)
        if div:
            diverged += 1
            continue
        st = e.state(r)
        for key, want in (("pos", orep.pos), ("rel", orep.rel), ("center", orep.center), ("quat", orep.quat)):
            assert np.array_equal(st[key], want), (r, key)
        assert abs(e.energies()[r] - orep.energy) < 1e-4 * kT(lad[r])
    assert diverged <= 1
    e.close()


def test_mc_windows_with_frequent_conflicts(oracle):
    """Three small bodies at high temperatures: most speculative windows end
    early, on a body drawn a third time, on a rejected move a later slot was
    chained to, or on an accepted move of an interacting body; most
    corrections between slots are non-zero."""
    from remcdockingsimulator_amd.engine import Engine
    from remcdockingsimulator_amd.forcefield import ForceField

    rng = np.random.default_rng(21)
    ff = ForceField.synthetic(21)
    sa = synth.build_system([synth.chain_body(5, rng), synth.chain_body(12, rng), synth.crowder_body()], 30.0, rng)
    lad = [300.0, 600.0, 1200.0]
    nsteps = 1500
    for block in (256, 1024):
        e = Engine(0)
        e.set_forcefield(ff)
        e.set_seed(9)
        sid = e.add_system(sa)
        e.add_replicas(sid, lad, 1.0, 0.3)
        e.set_mc_block(block)
        e.set_mc_coop(1)
        e.set_recording(nsteps)
        e.mc_steps(nsteps)
        e.synchronize()
        o = oracle.OracleSystem(sa, ff)
        for r in range(3):
            say("conflict windows block", block, "replica", r)
            orep = oracle.OracleReplica(o, lad[r], 1.0, 0.3, r)
            orec = orep.mc(seed=9, nsteps=nsteps, mode=0)
            n_eq, div = compare_decisions(e.records(r), orec, lad[r])
            if not div:
                st = e.state(r)
                assert np.array_equal(st["pos"], orep.pos) and np.array_equal(st["quat"], orep.quat)
                assert abs(e.energies()[r] - orep.energy) < 1e-4 * kT(lad[r])
        assert np.sum(e.counters()[:, :2]) > nsteps  # plenty of accepted moves
        e.close()


@pytest.mark.parametrize("lds_cache", [False, True])
def test_mc_partner_cache_modes(oracle, lds_cache):
    """Partner beads read from HBM or from the LDS cache: same decisions."""
    wl = synth.config(2)
    sa = wl.systems[0]
    lad = wl.temperatures[0]
    e = engine_for(wl, seed=SEED, block=512)
    e.set_mc_coop(2)
    e.set_mc_tuning(lds_cache=lds_cache)
    e.set_recording(400)
    e.mc_steps(400)
    e.synchronize()
    o = oracle.OracleSystem(sa, wl.ff)
    for r in (0, 19):
        orep = oracle.OracleReplica(o, lad[r], wl.translate_step, wl.rotate_step, r)
        n_eq, div = compare_decisions(e.records(r), orep.mc(seed=SEED, nsteps=400, mode=0), lad[r])
        if not div:
            assert np.array_equal(e.state(r)["pos"], orep.pos)
    e.close()


def test_mc_config5_member_slices(oracle):
    """Config 5 (11000 beads): every move is split over 4 members, each caching
    only its partner slice."""
    wl = synth.config(5)
    sa = wl.systems[0]
    lad = wl.temperatures[0]
    e = engine_for(wl, seed=SEED, block=512)
    e.set_mc_coop(4)
    e.set_recording(60)
    e.mc_steps(60)
    e.synchronize()
    o = oracle.OracleSystem(sa, wl.ff)
    for r in (0, 63):
        say("config 5 replica", r)
        orep = oracle.OracleReplica(o, lad[r], wl.translate_step, wl.rotate_step, r)
        n_eq, div = compare_decisions(e.records(r), orep.mc(seed=SEED, nsteps=60, mode=0), lad[r])
        if not div:
            assert np.array_equal(e.state(r)["pos"], orep.pos)
    e.close()


def test_cooperative_state_resync_after_host_changes(oracle):
    """Plain (host-protocol) steps and set_state change copy 0 only; the next
    cooperative launch must refresh the other copies first."""
    wl = synth.config(2)
    sa = wl.systems[0]
    lad = wl.temperatures[0]
    e = engine_for(wl, seed=SEED, block=256)
    e.set_mc_coop(4)
    e.mc_steps(100)
    e.mc_steps_plain(3, 50)
    e.set_recording(200)
    e.mc_steps(200)
    e.synchronize()
    o = oracle.OracleSystem(sa, wl.ff)
    orep = oracle.OracleReplica(o, lad[3], wl.translate_step, wl.rotate_step, 3)
    orec = orep.mc(seed=SEED, nsteps=150, mode=0)
    orep2 = oracle.OracleReplica(o, lad[3], wl.translate_step, wl.rotate_step, 3, pos=orep.pos, rel=orep.rel,
                                 center=orep.center, quat=orep.quat, energy=orep.energy, step=150)
    orec2 = orep2.mc(seed=SEED, nsteps=200, mode=0)
    n_eq, div = compare_decisions(e.records(3), orec2, lad[3])
    if not div:
        assert np.array_equal(e.state(3)["pos"], orep2.pos)
    e.close()


def test_plain_protocol_matches_oracle(oracle):
    """remc_mc_steps_plain: the reference's per-step host protocol with the
    reference-shaped kernel takes the oracle's decisions."""
    wl = synth.config(2)
    sa = wl.systems[0]
    lad = wl.temperatures[0]
    e = engine_for(wl, seed=SEED)
    n = 300
    o = oracle.OracleSystem(sa, wl.ff)
    for r in (0, 10):
        say("plain replica", r)
        e.mc_steps_plain(r, n)
        orep = oracle.OracleReplica(o, lad[r], wl.translate_step, wl.rotate_step, r)
        orec = orep.mc(seed=SEED, nsteps=n, mode=1)
        st = e.state(r)
        cnt = e.counters()[r]
        assert cnt[3] == n
        if cnt[2] == np.sum(orec["accepted"] == 0) and np.array_equal(st["pos"], orep.pos):
            continue
        # a divergence must come from a near-threshold decision
        assert np.min(orec["margin"]) < 1e-4
    e.close()


def test_exchange_matches_oracle(oracle):
    wl = synth.config(2)
    lad = wl.temperatures[0]
    e = engine_for(wl, seed=SEED)
    slots = np.arange(len(lad))
    for rnd in range(12):
        e.mc_steps(50)
        en = e.energies()
        n = e.exchange()
        want, nw, _ = oracle.exchange_round(lad, slots, en, rnd % 2, rnd, 0, SEED)
        got = e.ladder(0)
        say("exchange round", rnd, "swaps", n)
        assert n == nw
        assert np.array_equal(got, want)
        slots = want
        T = e.temperatures()
        assert np.allclose(T[got], lad)
    e.close()


def test_exchange_every_system_ladder(oracle):
    """Config 3: one exchange call runs the rounds of all 64 temperature
    ladders (one per system); each ladder's swaps equal the oracle's round
    for that system (the slot-replica map holds global replica indices)."""
    wl = synth.config(3)
    e = engine_for(wl, seed=SEED)
    first = np.cumsum([0] + [len(t) for t in wl.temperatures])
    slots = [np.arange(len(t)) + first[s] for s, t in enumerate(wl.temperatures)]
    for rnd in range(6):
        e.mc_steps(20)
        en = e.energies()
        n = e.exchange()
        total = 0
        for s, lad in enumerate(wl.temperatures):
            loc = slots[s] - first[s]
            want, nw, _ = oracle.exchange_round(lad, loc, en[first[s]:first[s + 1]], rnd % 2, rnd, s, SEED)
            total += nw
            got = e.ladder(s)
            assert np.array_equal(got, want + first[s]), (rnd, s, got, want)
            slots[s] = want + first[s]
        say("exchange round", rnd, "swaps", n)
        assert n == total
    e.close()


def test_graph_run_equals_eager(oracle):
    """remc_run (hipGraph replay of MC + exchange rounds) gives bit-identical
    state to the same calls made eagerly, over several replays."""
    wl = synth.config(2)
    a = engine_for(wl, seed=SEED, block=256)
    b = engine_for(wl, seed=SEED, block=256)
    a.set_mc_coop(4)
    b.set_mc_coop(4)
    for k in range(3):
        a.run(4, 100)
        for _ in range(4):
            b.mc_steps(100)
            b.exchange()
        a.synchronize()
        b.synchronize()
        say("graph replay batch", k)
        assert np.array_equal(a.energies(), b.energies())
        assert np.array_equal(a.temperatures(), b.temperatures())
        assert np.array_equal(a.ladder(0), b.ladder(0))
        for r in (0, 5, 19):
            assert np.array_equal(a.state(r)["pos"], b.state(r)["pos"])
    a.close()
    b.close()


def test_graph_run_with_refresh_equals_eager(oracle):
    """remc_run with the energy refresh before every exchange
    (remc_set_run_refresh) equals the eager sequence MC steps, refresh,
    exchange bit for bit, and leaves the tracked energies equal to the full
    evaluation of the conformations."""
    wl = synth.config(4)
    a = engine_for(wl, seed=SEED, block=1024)
    b = engine_for(wl, seed=SEED, block=1024)
    a.set_mc_coop(8)
    b.set_mc_coop(8)
    a.set_run_refresh(True)
    for k in range(2):
        a.run(3, 300)
        for _ in range(3):
            b.mc_steps(300)
            b.refresh_energies()
            b.exchange()
        a.synchronize()
        b.synchronize()
        say("refreshing graph replay batch", k)
        assert np.array_equal(a.energies(), b.energies())
        assert np.array_equal(a.temperatures(), b.temperatures())
        assert np.array_equal(a.ladder(0), b.ladder(0))
        for r in (0, 13, 31):
            assert np.array_equal(a.state(r)["pos"], b.state(r)["pos"])
    assert np.array_equal(a.energies(), a.energy_all())  # the last round's exchange read refreshed energies
    a.set_run_refresh(False)  # recaptured without the refresh: back to the plain rounds
    a.run(1, 300)
    a.synchronize()
    a.close()
    b.close()


def test_energy_split_agrees_across_contexts(oracle):
    """The full-energy split is timed once per process and grid size: two
    contexts of one size evaluate the same bits, and both match the oracle."""
    wl = synth.config(2)
    a = engine_for(wl, seed=SEED)
    b = engine_for(wl, seed=SEED + 1)
    ea, eb = a.energy_all(), b.energy_all()
    assert np.array_equal(ea, eb)  # same starting conformations
    osys = oracle.OracleSystem(wl.systems[0], wl.ff)
    want = osys.energy(a.state(0)["pos"])
    assert abs(ea[0] - want) <= 1e-5 * abs(want)
    a.close()
    b.close()


def test_single_body_and_all_crowder_systems(oracle):
    """Edge systems: one body (dE = 0, every move accepted) and crowders only."""
    from remcdockingsimulator_amd.engine import Engine
    from remcdockingsimulator_amd.forcefield import ForceField

    rng = np.random.default_rng(8)
    ff = ForceField.synthetic(8)
    one = synth.build_system([synth.chain_body(30, rng)], 100.0, rng)
    crow = synth.build_system([synth.crowder_body() for _ in range(40)], 60.0, rng)
    e = Engine(0)
    e.set_forcefield(ff)
    e.set_seed(5)
    s1 = e.add_system(one)
    e.add_replicas(s1, [300.0, 310.0], 2.0, 0.2)
    s2 = e.add_system(crow)
    e.add_replicas(s2, [300.0, 320.0, 340.0], 2.0, 0.2)
    e.set_recording(200)
    e.mc_steps(200)
    e.synchronize()
    cnt = e.counters()
    assert np.all(cnt[:2, 2] == 0)  # single body: no rejections
    assert np.all(e.energies()[:2] == 0.0)
    o = oracle.OracleSystem(crow, ff)
    for r in range(3):
        orep = oracle.OracleReplica(o, [300.0, 320.0, 340.0][r], 2.0, 0.2, 2 + r)
        orec = orep.mc(seed=5, nsteps=200, mode=0)
        compare_decisions(e.records(2 + r), orec, [300.0, 320.0, 340.0][r])
    e.close()


def test_mc_uneven_systems_in_one_launch(oracle):
    """Config 3: 64 systems of uneven bead, body and crowder counts (8 replicas
    each) in one persistent launch; replicas of several systems, from the
    first to the last, follow the oracle step for step."""
    wl = synth.config(3)
    nsteps = 200
    e = engine_for(wl, seed=SEED)
    e.set_recording(nsteps)
    e.mc_steps(nsteps)
    e.synchronize()
    energies = e.energies()
    full = e.energy_all()
    diverged = 0
    first = np.cumsum([0] + [len(t) for t in wl.temperatures])
    for s in (0, 17, 40, 63):
        sa = wl.systems[s]
        o = oracle.OracleSystem(sa, wl.ff)
        for k in (0, 7):
            r = int(first[s] + k)
            T = wl.temperatures[s][k]
            say("config 3 system", s, "beads", sa.n_beads, "replica", r)
            orep = oracle.OracleReplica(o, T, wl.translate_step, wl.rotate_step, r)
            orec = orep.mc(seed=SEED, nsteps=nsteps, mode=0)
            n_eq, div = compare_decisions(e.records(r), orec, T)
            st = e.state(r)
            if not div:
                assert np.array_equal(st["pos"], orep.pos), r
                assert np.array_equal(st["quat"], orep.quat), r
                assert abs(energies[r] - orep.energy) < 1e-4 * kT(T) + 1e-5 * abs(orep.energy)
            else:
                diverged += 1
            want = o.energy(st["pos"])
            assert abs(full[r] - want) <= 1e-5 * abs(want) + 1e-7
    e.close()
    assert diverged <= 1


def _quat_rotate(q, v):
    """Rotation of row vectors v by unit quaternions q (w, x, y, z), FP64."""
    w, x, y, z = q[:, 0:1], q[:, 1:2], q[:, 2:3], q[:, 3:4]
    u = np.concatenate([x, y, z], 1)
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def test_mc_full_size_rigid_body_invariants(oracle):
    """Config 5 at full size (11000 beads, 64 replicas) after 3000 steps of
    windows, split chain moves and exchanges: size-independent properties
    of rigid-body MC hold for every replica -- positions are relative
    positions plus centres (bit for bit), centres lie in [0, L), rotations
    stay unit quaternions, relative positions are the initial ones rotated by
    the accumulated rotation, and the tracked energy follows the full one."""
    wl = synth.config(5)
    sa = wl.systems[0]
    e = engine_for(wl, seed=SEED)
    e.run(3, 1000)
    e.synchronize()
    body = np.repeat(np.arange(sa.n_bodies), np.diff(sa.body_start))
    q0 = np.asarray(sa.quat, np.float64)
    q0_inv = q0 * np.array([1.0, -1.0, -1.0, -1.0])
    full = e.energy_all()
    tracked = e.energies()
    temps = e.temperatures()
    moved = 0
    for r in range(wl.n_replicas):
        st = e.state(r)
        assert np.array_equal(st["pos"], (st["rel"] + st["center"][body]).astype(np.float32)), r
        assert np.all(st["center"] >= 0.0) and np.all(st["center"] < sa.box), r
        q = st["quat"]
        assert np.max(np.abs(np.sum(q * q, 1) - 1.0)) < 1e-12, r
        # accumulated rotation q q0^-1 applied to the initial relative positions
        w1, x1, y1, z1 = q.T
        w2, x2, y2, z2 = q0_inv.T
        dq = np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                       w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], 1)
        want = _quat_rotate(dq[body], sa.rel.astype(np.float64))
        assert np.max(np.abs(st["rel"] - want)) < 2e-3, (r, np.max(np.abs(st["rel"] - want)))
        moved += int(np.sum(np.any(st["center"] != sa.center, 1)))
        assert abs(tracked[r] - full[r]) <= 1e-5 * abs(full[r]) + 1e-4 * kT(temps[r]) * 10, r
    assert moved > wl.n_replicas * 100  # the runs moved many bodies
    # the full energies themselves against the FP64 oracle on two replicas
    o = oracle.OracleSystem(sa, wl.ff)
    for r in (0, wl.n_replicas - 1):
        want = o.energy(e.state(r)["pos"])
        assert abs(full[r] - want) <= 1e-5 * abs(want)
    e.close()
