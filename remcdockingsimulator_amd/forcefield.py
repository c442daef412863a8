"""Force-field tables and constants.

Mirrors the reference's data layer:
  load_amino_acids      AminoAcids::loadAminoAcidData   (AminoAcids.cpp:25-71)
  amino_acid_index      AminoAcids::getAminoAcidIndex   (AminoAcids.cpp:146-195)
  load_pair_table       AminoAcids::loadLJPotentialData (AminoAcids.cpp:73-139)
  reference_constants   definitions.h:11-35,61,89 (float expressions, widened)

Reference paths are relative to "simulation engine/src/".  The data files
(AminoAcids, pair_potentials_MJ1996) ship in remcdockingsimulator_amd/data.
"""
import os
from dataclasses import dataclass, field

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
NTYPES = 20


@dataclass
class AminoAcid:
    index: int
    name: str
    sname: str
    symbol: str
    vdw_radius: float
    charge: float


def load_amino_acids(path=None):
    """Parse the reference's AminoAcids file: blocks of name / 3-letter name /
    1-letter symbol / vdW radius / charge, '#' lines and blank lines skipped.
    Raises FileNotFoundError where the reference calls exit(0)."""
    path = path or os.path.join(DATA_DIR, "AminoAcids")
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    out = []
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln != "" and not ln.startswith("#"):
            if i + 4 >= len(lines):
                raise ValueError("truncated amino-acid record at line %d of %s" % (i + 1, path))
            name = ln[:19]
            sname = lines[i + 1][:3]
            symbol = lines[i + 2][:1]
            vdw = float(np.float32(float(lines[i + 3].split()[0])))
            charge = float(np.float32(float(lines[i + 4].split()[0])))
            out.append(AminoAcid(len(out), name, sname, symbol, vdw, charge))
            i += 5
        else:
            i += 1
    if len(out) != NTYPES:
        raise ValueError("expected %d amino acids in %s, found %d" % (NTYPES, path, len(out)))
    return out


def amino_acid_index(aas, name):
    """Index of a residue by 3-letter name (case-normalised), 1-letter symbol or
    full name, searched from the end of the table as the reference does.
    Unknown names raise KeyError (the reference returns -1 and indexes with it)."""
    if len(name) < 1:
        raise KeyError(name)
    if len(name) == 3:
        key = name[0].upper() + name[1:].lower()
        for aa in reversed(aas):
            if aa.sname == key:
                return aa.index
    elif len(name) == 1:
        key = name[0].upper()
        for aa in reversed(aas):
            if aa.symbol == key:
                return aa.index
    else:
        key = name[0].upper() + name[1:]
        for aa in reversed(aas):
            if aa.name == key:
                return aa.index
    raise KeyError("unknown residue name %r" % name)


def load_pair_table(aas, path=None):
    """Upper-triangular 20x20 table with a label row, symmetrised, reordered to
    the amino-acid table's indices.  Values kept as float32 (the reference's
    float LJpotentials[20][20])."""
    path = path or os.path.join(DATA_DIR, "pair_potentials_MJ1996")
    table = np.zeros((NTYPES, NTYPES), dtype=np.float32)
    labels = None
    row = 0
    with open(path) as f:
        for ln in f:
            ln = ln.rstrip("\n")
            if ln == "" or ln.startswith("#"):
                continue
            toks = ln.split()
            if labels is None:
                labels = [amino_acid_index(aas, t) for t in toks]
                continue
            vals = [np.float32(float(v)) for v in toks[1:]]
            for k, j in enumerate(range(row, NTYPES)):
                a, b = labels[row], labels[j]
                table[a, b] = vals[k]
                table[b, a] = vals[k]
            row += 1
    if row != NTYPES:
        raise ValueError("pair table %s has %d rows, expected %d" % (path, row, NTYPES))
    return table


def reference_constants():
    """Float-evaluated constants of definitions.h, widened to double."""
    f = np.float32
    kbt = f(1.0) / (f(294.0) * f(8.314472) / f(4184.0))
    dh = f(0.948) * f(1.602176487) * f(1.602176487)
    return dict(
        lambda_=float(f(0.159)),
        e0=float(f(-2.27)),
        xi=10.0,
        lj_cf=float(f(0.3507221006079)),
        dh_cf=float(dh),
        kbt_cf=float(kbt),
        crowder_sigma=6.0,
        crowder_cutoff=15.0,
        eps_r=float(f(1e-38)),
    )


R_GAS = 8.314472


def kT(temperature):
    """kT in kcal/mol at temperature K (the Metropolis unit, Replica.cpp:451)."""
    return R_GAS * temperature / 4184.0


@dataclass
class ForceField:
    table: np.ndarray  # (20,20) float32, symmetric
    lambda_: float
    e0: float
    xi: float
    lj_cf: float
    dh_cf: float
    kbt_cf: float
    crowder_sigma: float = 6.0
    crowder_cutoff: float = 15.0
    eps_r: float = float(np.float32(1e-38))

    @classmethod
    def mj1996(cls, data_dir=None):
        aas = load_amino_acids(os.path.join(data_dir, "AminoAcids") if data_dir else None)
        table = load_pair_table(aas, os.path.join(data_dir, "pair_potentials_MJ1996") if data_dir else None)
        c = reference_constants()
        return cls(table=table, lambda_=c["lambda_"], e0=c["e0"], xi=c["xi"], lj_cf=c["lj_cf"],
                   dh_cf=c["dh_cf"], kbt_cf=c["kbt_cf"])

    @classmethod
    def synthetic(cls, seed, low=-0.6, high=0.0):
        """Seeded symmetric 20x20 table uniform in [low, high] kT units, used
        directly as E_ij (lambda = 1, e0 = 0); reference unit conversions."""
        rng = np.random.default_rng(seed)
        u = rng.uniform(low, high, size=(NTYPES, NTYPES)).astype(np.float32)
        table = np.triu(u) + np.triu(u, 1).T
        c = reference_constants()
        return cls(table=table.astype(np.float32), lambda_=1.0, e0=0.0, xi=c["xi"], lj_cf=c["lj_cf"],
                   dh_cf=c["dh_cf"], kbt_cf=c["kbt_cf"])

    def eij(self):
        """E_ij = lambda*(table - e0) in kT units (float64)."""
        return self.lambda_ * (self.table.astype(np.float64) - self.e0)
