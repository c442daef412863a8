"""The facade's Replica/Molecule calls of the reference's API that need no
device: Molecule::saveBeforeStateChange / undoStateChange (Molecule.cpp:
88-102) restore a moved molecule bit for bit, and Replica::copy
(Replica.cpp:161-216) copies the state but never a device context."""
import subprocess

from facade_api_harness import DATA, build
from facade_helpers import split_chains


def test_undo_and_copy(tmp_path):
    tmp = str(tmp_path)
    exe = build(tmp)
    paths = split_chains(1, tmp)
    r = subprocess.run([exe, "cpu", DATA, paths["A"], paths["B"], tmp], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = dict(ln.split() for ln in r.stdout.strip().split("\n"))
    assert got == {"moved": "1", "undo": "1", "copy": "1", "copyctor": "1"}, r.stdout
