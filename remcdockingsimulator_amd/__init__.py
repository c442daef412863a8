"""MI355X-native replica-exchange Monte Carlo hot path of a coarse-grained
rigid-body protein docking simulator (Kim-Hummer residue model: MJ1996 contact
energies, Debye-Hueckel electrostatics, repulsive crowders, periodic box).

The compute path is libremc.so (HIP kernels for gfx950 behind the C ABI of
include/remc.h); this package holds its Python bindings, the data layer
(force field tables, Calpha PDB loading, rigid-body model), synthetic workload
generators and the Replica/Molecule facade.
"""
from . import forcefield, model  # noqa: F401

__all__ = ["forcefield", "model", "engine", "synth"]
