"""Rigid-body model: Calpha PDB loading, molecules, and packing a system for
the C ABI.

Mirrors (paths relative to "simulation engine/src/"):
  read_ca_pdb         Molecule::initFromPDB       Molecule.cpp:430-830
  Molecule.translate  Molecule::translate         Molecule.cpp:146-171
  Molecule.set_position  Molecule::setPosition    Molecule.cpp:247-260
  Molecule.rotate_q   Molecule::rotateQ           Molecule.cpp:288-333
  Molecule.set_rotation  Molecule::setRotation    Molecule.cpp:335-346
  Molecule.set_role   Molecule::setMoleculeRoleIdentifier  Molecule.cpp:118-139
  write_single_pdb    Replica::saveAsSinglePDB    Replica.cpp:495-530

All arithmetic follows the reference's precision: positions, centres and
relative positions are float32; rotations are applied in float64 and stored
as float32 (numpy performs no fused multiply-add, so each operation rounds
once, as the reference's scalar C++ does).
"""
from dataclasses import dataclass, field

import numpy as np

from . import forcefield as ffm

CROWDER = -1.0  # CROWDER_IDENTIFIER (definitions.h:92)
PROTEIN = 0.0

f32 = np.float32


@dataclass
class PdbResidue:
    name: str
    chain_id: str
    res_seq: int
    seg_id: str
    x: float
    y: float
    z: float


def parse_ca_records(path):
    """Calpha ATOM records in file order.  Columns per the PDB format the
    reference parses (Molecule.cpp:625-719): name 13-16, resName 18-20,
    chainID 22, resSeq 23-26, x/y/z 31-54, segID 73-76.  As in the reference,
    END does not stop the scan of later lines."""
    out = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            toks = line.split()
            if not toks or toks[0] != "ATOM":
                continue
            name = line[12:16].strip()
            if name != "CA":
                continue
            res_name = line[17:20].strip()
            chain = line[21:22] if len(line) > 21 else " "
            try:
                res_seq = int(line[22:26])
            except ValueError:
                res_seq = 0
            x = float(f32(float(line[30:38])))
            y = float(f32(float(line[38:46])))
            z = float(f32(float(line[46:54])))
            seg = line[72:76].strip() if len(line) > 72 else ""
            out.append(PdbResidue(res_name, chain, res_seq, seg, x, y, z))
    return out


def quat_mul(a, b):
    """Quaternion::multiply (Quaternion.cpp:35-47), (w,x,y,z), float64."""
    w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3]
    x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2]
    y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1]
    z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]
    return np.array([w, x, y, z], dtype=np.float64)


def quat_rotate(q, v):
    """Quaternion::rotateVector (Quaternion.cpp:84-101): float64 arithmetic on
    float32 vectors v (n,3), rounded to float32."""
    w, x, y, z = (float(t) for t in q)
    w2, x2, y2, z2 = w * w, x * x, y * y, z * z
    wx, wy, wz, xy, xz, yz = w * x, w * y, w * z, x * y, x * z, y * z
    vx = v[:, 0].astype(np.float64)
    vy = v[:, 1].astype(np.float64)
    vz = v[:, 2].astype(np.float64)
    ox = vx * (w2 + x2 - y2 - z2) + (vy * (xy - wz) + vz * (wy + xz)) * 2.0
    oy = (vx * (wz + xy) + vz * (yz - wx)) * 2.0 + vy * (w2 - x2 + y2 - z2)
    oz = (vx * (xz - wy) + vy * (yz + wx)) * 2.0 + vz * (w2 - x2 - y2 + z2)
    return np.stack([ox, oy, oz], axis=1).astype(np.float32)


def normalize_axis(axis):
    """Vector3double::normalizeInPlace (vector3f.cpp:36-43)."""
    a = np.asarray(axis, dtype=np.float64)
    mag = np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    return a / mag


class Molecule:
    """One rigid body (Molecule.h:77-95)."""

    def __init__(self, types, charges, vdw, positions, name="", chain_ids=None, res_seq=None):
        self.types = np.asarray(types, dtype=np.int32)
        self.charges = np.asarray(charges, dtype=np.float32)
        self.vdw = np.asarray(vdw, dtype=np.float32)
        self.position = np.asarray(positions, dtype=np.float32).reshape(-1, 3).copy()
        n = len(self.types)
        self.chain_ids = list(chain_ids) if chain_ids is not None else [" "] * n
        self.res_seq = list(res_seq) if res_seq is not None else list(range(1, n + 1))
        self.name = name
        self.role = PROTEIN
        self.rotation = np.array([1.0, 0.0, 0.0, 0.0], dtype=np.float64)
        # centre = float sum / count (Molecule.cpp:798-808), relative positions (811-816)
        acc = np.zeros(3, dtype=np.float32)
        for k in range(n):
            acc = (acc + self.position[k]).astype(np.float32)
        self.center = (acc / f32(n)).astype(np.float32)
        self.relative = (self.position - self.center).astype(np.float32)

    @property
    def residue_count(self):
        return len(self.types)

    def translate(self, v):
        v = np.asarray(v, dtype=np.float32)
        self.center = (self.center + v).astype(np.float32)
        self.position = (self.position + v).astype(np.float32)

    def set_position(self, v):
        self.center = np.asarray(v, dtype=np.float32).copy()
        self.position = (self.relative + self.center).astype(np.float32)

    def rotate_q(self, axis, angle):
        """axis must be a unit vector (the caller normalises, Replica.cpp:259)."""
        sa, ca = np.sin(angle / 2.0), np.cos(angle / 2.0)
        q = np.array([ca, sa * axis[0], sa * axis[1], sa * axis[2]], dtype=np.float64)
        q = q / np.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3])
        self.rotation = quat_mul(q, self.rotation)
        self.relative = quat_rotate(q, self.relative)
        self.position = (self.relative + self.center).astype(np.float32)

    def set_rotation(self, q):
        q = np.asarray(q, dtype=np.float64)
        q = q / np.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3])
        self.rotation = quat_mul(q, self.rotation)
        self.relative = quat_rotate(q, self.relative)
        self.position = (self.relative + self.center).astype(np.float32)

    def set_role(self, role):
        self.role = float(role)

    @property
    def is_crowder(self):
        return self.role == CROWDER

    @classmethod
    def from_residues(cls, residues, aas, name=""):
        types, charges, vdw, pos = [], [], [], []
        for r in residues:
            idx = ffm.amino_acid_index(aas, r.name)
            types.append(idx)
            charges.append(aas[idx].charge)
            vdw.append(aas[idx].vdw_radius)
            pos.append((r.x, r.y, r.z))
        return cls(types, charges, vdw, pos, name=name, chain_ids=[r.chain_id for r in residues],
                   res_seq=[r.res_seq for r in residues])


def read_ca_pdb(path, aas=None, split_segid=False):
    """Molecule::initFromPDB.  Returns one Molecule, or one per segment ID when
    split_segid is set (the reference's conformation fixtures keep both chains
    in one file, told apart by segID)."""
    aas = aas or ffm.load_amino_acids()
    recs = parse_ca_records(path)
    if not recs:
        raise ValueError("no CA atoms in %s" % path)
    if not split_segid:
        return Molecule.from_residues(recs, aas, name=path)
    groups = {}
    order = []
    for r in recs:
        if r.seg_id not in groups:
            groups[r.seg_id] = []
            order.append(r.seg_id)
        groups[r.seg_id].append(r)
    return [Molecule.from_residues(groups[s], aas, name="%s:%s" % (path, s)) for s in order]


@dataclass
class SystemArrays:
    """A system packed for the C ABI (remc_system_desc)."""
    body_start: np.ndarray
    type: np.ndarray
    charge: np.ndarray
    sigma: np.ndarray
    crowder: np.ndarray
    box: float
    pos: np.ndarray       # (N,3) float32
    rel: np.ndarray       # (N,3) float32
    center: np.ndarray    # (M,3) float32
    quat: np.ndarray      # (M,4) float64

    @property
    def n_beads(self):
        return int(self.body_start[-1])

    @property
    def n_bodies(self):
        return len(self.body_start) - 1


def pack_molecules(molecules, box):
    """Concatenate molecules in order (non-crowders first by convention, as the
    reference's config loader does).  Crowder residues keep their type; the
    reference replaces it by -1 (Molecule.cpp:121-128) but never reads it for
    crowder pairs."""
    starts = [0]
    for m in molecules:
        starts.append(starts[-1] + m.residue_count)
    cat = lambda attr: np.concatenate([getattr(m, attr) for m in molecules])
    return SystemArrays(
        body_start=np.asarray(starts, dtype=np.int32),
        type=cat("types").astype(np.int32),
        charge=cat("charges").astype(np.float32),
        sigma=cat("vdw").astype(np.float32),
        crowder=np.asarray([1 if m.is_crowder else 0 for m in molecules], dtype=np.uint8),
        box=float(np.float32(box)),
        pos=np.concatenate([m.position for m in molecules]).astype(np.float32),
        rel=np.concatenate([m.relative for m in molecules]).astype(np.float32),
        center=np.stack([m.center for m in molecules]).astype(np.float32),
        quat=np.stack([m.rotation for m in molecules]).astype(np.float64),
    )


def write_single_pdb(path, molecules, aas, potential, temperature):
    """Replica::saveAsSinglePDB record layout (Replica.cpp:495-530)."""
    with open(path, "w") as f:
        f.write("REMARK %s \n" % path)
        f.write("REMARK potential: %0.10f \n" % float(np.float32(potential)))
        f.write("REMARK temperature: %5.1f \n" % temperature)
        for i, m in enumerate(molecules):
            f.write("REMARK Molecule: %d\n" % i)
            f.write("REMARK Rotation relative to input Q(w,x,y,z): %f %f %f %f\n" % tuple(m.rotation))
            f.write("REMARK Centriod position P(x,y,z): %f %f %f\n" % tuple(float(c) for c in m.center))
        chain = ord("A")
        item = 0
        for m in molecules:
            last = 0
            for i in range(m.residue_count):
                item += 1
                f.write("ATOM  %5d %4s%c%3s %c%4d%c  %8.3f%8.3f%8.3f%6.2f%6.2f\n" % (
                    item, "CA", " ", aas[m.types[i]].sname, chr(chain), m.res_seq[i], " ",
                    m.position[i, 0], m.position[i, 1], m.position[i, 2], 1.0, 1.0))
                last = m.res_seq[i]
            f.write("TER   %5d      %3s %c%4d \n" % (item, aas[m.types[m.residue_count - 1]].sname, chr(chain), last))
            chain += 1
        f.write("END \n")
