"""remc_energy_all_noalloc (remc_energy_all with its partials in the
library's static device scratch instead of context buffers) against
remc_energy_all through the C ABI: identical bits on two calls in a row, the
second with a smaller input, and a clear error where the scratch is too small
or the context is not set up."""
import numpy as np
import pytest

from remcdockingsimulator_amd import _native as nat
from remcdockingsimulator_amd import forcefield as ffm
from remcdockingsimulator_amd import model, synth
from remcdockingsimulator_amd.engine import Engine

pytestmark = pytest.mark.gpu


def engine_with(wl, seed):
    e = Engine(0)
    synth.load_engine(e, wl, seed=seed)
    return e


def test_noalloc_equals_energy_all_two_calls_second_smaller():
    # first call: config 4 (32 replicas x 1300 beads) after some MC steps
    e1 = engine_with(synth.config(4), 7)
    e1.mc_steps(100)
    e1.synchronize()
    want1 = e1.energy_all()
    got1 = e1.energy_all_noalloc()
    assert np.array_equal(got1, want1), np.max(np.abs(got1 - want1))
    # second call, smaller input: config 1 (8 replicas x 100 beads) in another
    # context on the same device, so on the same scratch
    e2 = engine_with(synth.config(1), 8)
    e2.mc_steps(50)
    e2.synchronize()
    want2 = e2.energy_all()
    got2 = e2.energy_all_noalloc()
    assert np.array_equal(got2, want2), np.max(np.abs(got2 - want2))
    # the first context again, after the smaller call left its contents
    assert np.array_equal(e1.energy_all_noalloc(), want1)
    e2.close()
    e1.close()


def test_noalloc_needs_a_set_up_context():
    e = engine_with(synth.config(1), 9)
    with pytest.raises(nat.RemcError, match="not set up"):
        e.energy_all_noalloc()
    e.energies()  # sets the context up
    assert np.array_equal(e.energy_all_noalloc(), e.energy_all())
    e.close()


def test_noalloc_rejects_more_than_its_scratch():
    """16 replicas of 40,000 single-bead bodies need 16 (4 x 157 x 79 + 1) =
    793,808 doubles of scratch, more than its 524,288."""
    side = 35
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)[:40000]
    n = len(g)
    pos = (g * 7.0 + 3.0).astype(np.float32)
    sa = model.SystemArrays(
        body_start=np.arange(n + 1, dtype=np.int32), type=np.zeros(n, np.int32), charge=np.zeros(n, np.float32),
        sigma=np.full(n, 6.0, np.float32), crowder=np.ones(n, np.uint8), box=float(np.float32(side * 7.0)),
        pos=pos, rel=np.zeros((n, 3), np.float32), center=pos.copy(),
        quat=np.tile(np.array([1.0, 0.0, 0.0, 0.0]), (n, 1)))
    e = Engine(0)
    e.set_forcefield(ffm.ForceField.synthetic(3))
    sid = e.add_system(sa)
    e.add_replicas(sid, [300.0 + k for k in range(16)], 2.0, 0.2)
    e.energies()
    with pytest.raises(nat.RemcError, match="scratch"):
        e.energy_all_noalloc()
    e.close()
