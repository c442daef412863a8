"""Regenerate the golden fixtures in this directory from the reference checkout.

The fixtures are the reference's own validation data (no reference code is
copied or run):

  conf/{1..10}x.pdb                  "post processor/data/structures.tgz"
                                     (ten UIM (segID A, 24 Calpha) + ubiquitin
                                     (segID B, 76 Calpha) conformations)
  energy.txt                         "simulation engine/data/structures/energy.txt"
                                     (the reference's E_tot/E_vdw/E_elec beside
                                     the Kim-Hummer values)
  ref_ene_nocut_noim_c35a2.dat       "simulation engine/data/structures/..."
                                     (Kim-Hummer E_tot/E_elec/E_vdw)
  pairs/pairs_{1..10}.dat            "simulation engine/data/pairinteractions/"
                                     (i j E_tot E_elec E_vdw per residue pair)
  AminoAcids, pair_potentials_MJ1996 "simulation engine/data/" (force field)

Usage:  python tests/golden/make_golden.py [/root/reference]
"""
import os
import shutil
import sys
import tarfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref="/root/reference"):
    eng = os.path.join(ref, "simulation engine", "data")
    post = os.path.join(ref, "post processor", "data")
    os.makedirs(os.path.join(HERE, "conf"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "pairs"), exist_ok=True)
    with tarfile.open(os.path.join(post, "structures.tgz")) as tf:
        for m in tf.getmembers():
            name = os.path.basename(m.name)
            if m.isfile() and name.endswith("x.pdb"):
                data = tf.extractfile(m).read()
                with open(os.path.join(HERE, "conf", name), "wb") as f:
                    f.write(data)
    for n in ("energy.txt", "ref_ene_nocut_noim_c35a2.dat"):
        shutil.copyfile(os.path.join(eng, "structures", n), os.path.join(HERE, n))
    for i in range(1, 11):
        shutil.copyfile(os.path.join(eng, "pairinteractions", "pairs_%d.dat" % i),
                        os.path.join(HERE, "pairs", "pairs_%d.dat" % i))
    for n in ("AminoAcids", "pair_potentials_MJ1996"):
        shutil.copyfile(os.path.join(eng, n), os.path.join(HERE, n))


if __name__ == "__main__":
    main(*sys.argv[1:])
