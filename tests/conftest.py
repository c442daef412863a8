import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle

    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def aas():
    from remcdockingsimulator_amd import forcefield

    return forcefield.load_amino_acids(os.path.join(GOLDEN, "AminoAcids"))


@pytest.fixture(scope="session")
def mj_ff():
    from remcdockingsimulator_amd import forcefield

    return forcefield.ForceField.mj1996(GOLDEN)


def conformation(idx, aas, box=0.0):
    """The reference's conformation fixture idx (1..10), chains split by segID."""
    from remcdockingsimulator_amd import model

    mols = model.read_ca_pdb(os.path.join(GOLDEN, "conf", "%dx.pdb" % idx), aas, split_segid=True)
    return model.pack_molecules(mols, box)


@pytest.fixture(scope="session")
def gpu_engine_factory():
    """Factory of Engines on cuda:0; skips nothing: a missing library raises."""
    from remcdockingsimulator_amd.engine import Engine

    made = []

    def make():
        e = Engine(0)
        made.append(e)
        return e

    yield make
    for e in made:
        e.close()
