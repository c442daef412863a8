"""Radial distribution g(r) of chains from PDB frames, as "post processor/gr.py"
computes it, with the histogram and its normalisation on the device
(remc_radial_distribution).

    python -m remcdockingsimulator_amd.radial <boxlen> frame1.pdb [frame2.pdb ...] [--outdir d]

Follows gr.py:
- chains of the first frame with the same sequence form a type, keyed by its
  first chain (gr.py:14-60; types ordered by key here, where gr.py takes
  its dictionary's order);
- a type's site is the residue (by number) nearest the centroid of its key
  chain in the first frame (gr.py:41-60), and a chain's point in every frame
  is the position of its residue with that number (gr.py:62-88; the last
  record with the number, as gr.py's loop leaves it);
- pairs of chains Ichain != Jchain over types I <= J, minimum image in the
  cubic box, lo = 0, 75 bins up to box/2 - 1 (gr.py:93-138);
- g = counts / (n_frames 4 pi r^2 dr n_I n_J / box^3) at the bin centre r,
  written to gr_<I>_<J>.dat as "%8.3f %8.3f" lines (gr.py:140-155).

ATOM records are read with gr.py's columns (residue name 17:20, chain 21,
residue number 22:26, x/y/z 29:37/37:45/45:53), which are those of the
reference's PDB writer (Replica.cpp:519, Molecule.cpp:236).
"""
import argparse
import os
import sys

import numpy as np

from .engine import Engine
from .forcefield import ForceField
from .model import SystemArrays


def read_frame(path):
    """{chain: [(resname, resnum, (x, y, z)), ...]} in file order."""
    chains = {}
    with open(path) as f:
        for a in f:
            if a[0:6] != "ATOM  ":
                continue
            chains.setdefault(a[21], []).append(
                (a[17:20], int(a[22:26]), (float(a[29:37]), float(a[37:45]), float(a[45:53]))))
    return chains


def chain_types(frame):
    """(type keys, {key: [chains]}, {key: site residue number}) from the first frame."""
    types, site = {}, {}
    for c in sorted(frame):
        seq = [r[0] for r in frame[c]]
        for k in types:
            if [r[0] for r in frame[k]] == seq:
                types[k].append(c)
                break
        else:
            types[c] = [c]
            xyz = np.array([r[2] for r in frame[c]], dtype=np.float64)
            d = np.sqrt(((xyz - xyz.mean(0)) ** 2).sum(1))
            site[c] = frame[c][int(np.argmin(d))][1]  # first residue at the smallest distance
    return sorted(types), types, site


def radial_distribution(boxlen, frames, nbin=75, lo=0.0, hi=None):
    """(type keys, bin centres r, g[I, J, bin] for I <= J) of PDB frames."""
    if hi is None:
        hi = boxlen / 2.0 - 1.0
    first = read_frame(frames[0])
    keys, types, site = chain_types(first)
    chain_list = sorted(first)
    sizes = [len(first[c]) for c in chain_list]
    start = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(start[-1])
    body_type = np.full(len(chain_list), -1, dtype=np.int32)
    site_bead = np.zeros(len(chain_list), dtype=np.int32)
    for m, c in enumerate(chain_list):
        k = next(k for k in keys if c in types[k])
        body_type[m] = keys.index(k)
        hits = [i for i, r in enumerate(first[c]) if r[1] == site[k]]
        if not hits:
            raise ValueError("chain %s has no residue %d" % (c, site[k]))
        site_bead[m] = start[m] + hits[-1]
    pos = []
    for f in frames:
        fr = read_frame(f)
        if sorted(fr) != chain_list or any(len(fr[c]) != len(first[c]) for c in chain_list):
            raise ValueError("%s does not hold the first frame's chains" % f)
        pos.append(np.array([r[2] for c in chain_list for r in fr[c]], dtype=np.float32))
    # one system (chains as bodies), one replica per frame; only positions are read
    centre = np.stack([pos[0][start[m]:start[m + 1]].mean(0) for m in range(len(chain_list))]).astype(np.float32)
    body = np.repeat(np.arange(len(chain_list)), sizes)
    sa = SystemArrays(body_start=start, type=np.zeros(n, np.int32), charge=np.zeros(n, np.float32),
                      sigma=np.full(n, 6.0, np.float32), crowder=np.zeros(len(chain_list), np.uint8),
                      box=float(np.float32(boxlen)), pos=pos[0], rel=(pos[0] - centre[body]).astype(np.float32),
                      center=centre, quat=np.tile([1.0, 0.0, 0.0, 0.0], (len(chain_list), 1)))
    e = Engine(0)
    try:
        e.set_forcefield(ForceField.synthetic(0))
        sid = e.add_system(sa)
        e.add_replicas(sid, [300.0 + 1e-3 * k for k in range(len(frames))], 2.0, 0.2)
        for k, p in enumerate(pos):
            e.set_state(k, pos=p)
        g = e.radial_distribution(sid, body_type, len(keys), nbin, lo, hi, site_bead=site_bead)
    finally:
        e.close()
    dr = (hi - lo) / nbin
    r = lo + (0.5 + np.arange(nbin)) * dr
    return keys, r, g


def write_files(outdir, keys, r, g):
    paths = []
    for i in range(len(keys)):
        for j in range(i, len(keys)):
            path = os.path.join(outdir, "gr_%s_%s.dat" % (keys[i], keys[j]))
            with open(path, "w") as f:
                for b in range(len(r)):
                    f.write("%8.3f %8.3f\n" % (r[b], g[i, j, b]))
            paths.append(path)
    return paths


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("boxlen", type=float)
    ap.add_argument("frames", nargs="+")
    ap.add_argument("--outdir", default=".")
    args = ap.parse_args(argv)
    keys, r, g = radial_distribution(args.boxlen, args.frames)
    for p in write_files(args.outdir, keys, r, g):
        print(p)


if __name__ == "__main__":
    sys.exit(main())
