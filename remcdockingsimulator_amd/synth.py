"""Seeded synthetic workloads for the five benchmark configurations.

  config 1  two rigid chains (76 + 24 residues), 150 A box, 8 replicas 288-450 K
  config 2  config 1 + 200 crowder beads, 300 A box, 20 replicas 288-480 K
  config 3  64 independent systems (2-12 chains of 30-300 residues, 0-500
            crowders, 500-4000 beads), 8 replicas each (512 instances)
  config 4  10 x 100-residue chains + 300 crowders, 32 replicas
  config 5  60 x 150-residue identical chains + 2000 crowders, 600 A box,
            64 replicas 288-500 K

Chains are self-avoiding random walks (3.8 A Calpha bonds, >= 4.0 A between
non-bonded beads) grown inside a sphere sized for a radius of gyration of
12*(n/76)^(1/3) A.  Residue types are uniform over the 20 amino acids, charges
+1 for K/R, -1 for D/E, 0 otherwise; sigma = 6 A for every residue.  The
contact table is a seeded symmetric 20x20 uniform in [-0.6, 0] kT used
directly as E_ij.  A crowder is one repulsive bead.  Temperatures follow the
reference's geometric ladder (main.cpp:1161-1169).  Move magnitudes are fixed
per replica, as in the reference; the configurations' "at most 2 A / 0.2 rad"
is taken as those fixed magnitudes.
"""
from dataclasses import dataclass, field

import numpy as np

from . import forcefield as ffm
from .model import SystemArrays, quat_rotate

BASE_SEED = 0x52454D43
ARG, ASP, GLU, LYS = 1, 3, 5, 9  # indices in the AminoAcids table order


def charges_for(types):
    q = np.zeros(len(types), dtype=np.float32)
    q[(types == ARG) | (types == LYS)] = 1.0
    q[(types == ASP) | (types == GLU)] = -1.0
    return q


def target_rg(n):
    return 12.0 * (n / 76.0) ** (1.0 / 3.0)


def saw_chain(n, rng, bond=3.8, excl=4.0, rg=None, max_restarts=200):
    """Self-avoiding walk grown inside a sphere of radius sqrt(5/3)*rg,
    centred at the origin on return."""
    rg = target_rg(n) if rg is None else rg
    R = np.sqrt(5.0 / 3.0) * rg
    for _ in range(max_restarts):
        pos = np.zeros((n, 3))
        pos[0] = rng.uniform(-0.2 * R, 0.2 * R, 3)
        ok = True
        for i in range(1, n):
            placed = False
            for _attempt in range(300):
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                p = pos[i - 1] + bond * d
                if np.dot(p, p) > R * R:
                    continue
                if i >= 2:
                    dd = pos[: i - 1] - p
                    if np.min(np.einsum("ij,ij->i", dd, dd)) < excl * excl:
                        continue
                pos[i] = p
                placed = True
                break
            if not placed:
                ok = False
                break
        if ok:
            pos -= pos.mean(axis=0)
            return pos
    raise RuntimeError("could not grow a self-avoiding chain of %d beads" % n)


def random_quat(rng):
    """Uniform random unit quaternion (Shoemake 1992)."""
    u1, u2, u3 = rng.uniform(size=3)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.array([b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3)], dtype=np.float64)


def min_image_dist2(a, b, box):
    d = a[:, None, :] - b[None, :, :]
    if box > 0:
        d -= box * np.round(d / box)
    return np.einsum("ijk,ijk->ij", d, d)


@dataclass
class Body:
    local: np.ndarray       # (n,3) float64 coordinates about the centre
    types: np.ndarray       # (n,) int
    crowder: bool


def place_bodies(bodies, box, rng, min_gap_chain=4.5, min_gap_crowder=8.0, max_tries=5000):
    """Random centres and orientations without close contacts (periodic)."""
    placed_pos = np.zeros((0, 3))
    out = []
    for b in bodies:
        gap = min_gap_crowder if b.crowder else min_gap_chain
        for _ in range(max_tries):
            q = np.array([1.0, 0, 0, 0]) if b.crowder else random_quat(rng)
            c = rng.uniform(0.0, box, 3).astype(np.float32)
            rel = quat_rotate(q, b.local.astype(np.float32))
            pos = (rel + c).astype(np.float32)
            if len(placed_pos):
                if min_image_dist2(pos.astype(np.float64), placed_pos, box).min() < gap * gap:
                    continue
            placed_pos = np.concatenate([placed_pos, pos.astype(np.float64)])
            out.append((pos, rel, c, q))
            break
        else:
            raise RuntimeError("could not place body of %d beads in a %.1f A box" % (len(b.types), box))
    return out


def build_system(bodies, box, rng):
    placed = place_bodies(bodies, box, rng)
    starts = np.cumsum([0] + [len(b.types) for b in bodies]).astype(np.int32)
    types = np.concatenate([b.types for b in bodies]).astype(np.int32)
    return SystemArrays(
        body_start=starts,
        type=types,
        charge=np.concatenate([np.zeros(len(b.types), np.float32) if b.crowder else charges_for(b.types)
                               for b in bodies]),
        sigma=np.full(len(types), 6.0, dtype=np.float32),
        crowder=np.array([1 if b.crowder else 0 for b in bodies], dtype=np.uint8),
        box=float(np.float32(box)),
        pos=np.concatenate([p[0] for p in placed]),
        rel=np.concatenate([p[1] for p in placed]),
        center=np.stack([p[2] for p in placed]),
        quat=np.stack([p[3] for p in placed]),
    )


def chain_body(n, rng):
    return Body(saw_chain(n, rng), rng.integers(0, 20, n), False)


def crowder_body():
    return Body(np.zeros((1, 3)), np.zeros(1, dtype=np.int64), True)


def geometric_ladder(tmin, tmax, n):
    if n == 1:
        return np.array([tmin], dtype=np.float64)
    g = (tmax / tmin) ** (1.0 / (n - 1))
    return np.array([tmin * g ** i for i in range(n)], dtype=np.float64)


def box_for_fraction(n_chain_beads_list, n_crowders, phi, crowder_radius=14.0):
    vol = sum(4.0 / 3.0 * np.pi * (np.sqrt(5.0 / 3.0) * target_rg(n)) ** 3 for n in n_chain_beads_list)
    vol += n_crowders * 4.0 / 3.0 * np.pi * crowder_radius ** 3
    return float((vol / phi) ** (1.0 / 3.0))


@dataclass
class Workload:
    name: str
    systems: list               # SystemArrays
    temperatures: list          # per system ladder
    translate_step: float
    rotate_step: float
    ff: ffm.ForceField
    steps: int = 0              # reference per-replica step count where the config states it
    exchange_every: int = 0
    notes: dict = field(default_factory=dict)

    @property
    def n_replicas(self):
        return sum(len(t) for t in self.temperatures)

    def pairs_full(self, s=0):
        """Inter-body pair terms of one full evaluation of system s."""
        sa = self.systems[s]
        sizes = np.diff(sa.body_start).astype(np.int64)
        n = sizes.sum()
        return int((n * n - (sizes * sizes).sum()) // 2)

    def pairs_delta_mean(self, s=0):
        """Mean pair terms of one dE evaluation (old and new positions) over
        uniformly chosen bodies."""
        sa = self.systems[s]
        sizes = np.diff(sa.body_start).astype(np.int64)
        n = sizes.sum()
        return float(np.mean(2 * sizes * (n - sizes)))


def config(cid, seed=None):
    seed = BASE_SEED + cid if seed is None else seed
    rng = np.random.default_rng(seed)
    ff = ffm.ForceField.synthetic(seed)
    if cid == 1 or cid == 2:
        bodies = [chain_body(76, rng), chain_body(24, rng)]
        if cid == 2:
            bodies += [crowder_body() for _ in range(200)]
        box = 150.0 if cid == 1 else 300.0
        sa = build_system(bodies, box, rng)
        lad = geometric_ladder(288.0, 450.0, 8) if cid == 1 else geometric_ladder(288.0, 480.0, 20)
        return Workload("config%d" % cid, [sa], [lad], 2.0, 0.2, ff,
                        steps=10000 if cid == 1 else 100000, exchange_every=1000)
    if cid == 3:
        systems, ladders = [], []
        for _ in range(64):
            while True:
                nch = int(rng.integers(2, 13))
                lens = [int(rng.integers(30, 301)) for _ in range(nch)]
                ncr = int(rng.integers(0, 501))
                if 500 <= sum(lens) + ncr <= 4000:
                    break
            bodies = [chain_body(n, rng) for n in lens] + [crowder_body() for _ in range(ncr)]
            box = box_for_fraction(lens, ncr, 0.2)
            systems.append(build_system(bodies, box, rng))
            ladders.append(geometric_ladder(288.0, 450.0, 8))
        return Workload("config3", systems, ladders, 2.0, 0.2, ff, exchange_every=1000)
    if cid == 4:
        bodies = [chain_body(100, rng) for _ in range(10)] + [crowder_body() for _ in range(300)]
        box = box_for_fraction([100] * 10, 300, 0.2)
        sa = build_system(bodies, box, rng)
        return Workload("config4", [sa], [geometric_ladder(288.0, 480.0, 32)], 2.0, 0.2, ff,
                        exchange_every=1000, notes={"box_rule": "packing fraction 0.2"})
    if cid == 5:
        proto = saw_chain(150, rng)
        bodies = [Body(proto, rng.integers(0, 20, 150), False)]
        bodies = [Body(proto, bodies[0].types, False) for _ in range(60)] + [crowder_body() for _ in range(2000)]
        sa = build_system(bodies, 600.0, rng)
        return Workload("config5", [sa], [geometric_ladder(288.0, 500.0, 64)], 2.0, 0.2, ff, exchange_every=1000)
    raise ValueError("config id must be 1..5")


def small_system(seed, n_chains=3, chain_len=(8, 20), n_crowders=4, box=60.0, ff=None):
    """Small randomized system for parity tests."""
    rng = np.random.default_rng(seed)
    bodies = [chain_body(int(rng.integers(chain_len[0], chain_len[1] + 1)), rng) for _ in range(n_chains)]
    bodies += [crowder_body() for _ in range(n_crowders)]
    return build_system(bodies, box, rng)


def load_engine(engine, wl, seed=0x1234):
    """Add a workload's systems and replicas to an Engine."""
    engine.set_forcefield(wl.ff)
    engine.set_seed(seed)
    for sa, lad in zip(wl.systems, wl.temperatures):
        sid = engine.add_system(sa)
        engine.add_replicas(sid, lad, wl.translate_step, wl.rotate_step)
    return engine
