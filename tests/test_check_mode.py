"""remc_sim -c, the reference's self-check (main.cpp:1652-1707): the ten
example conformations data/conf<i>/1a.pdb + 1b.pdb in a 118.4 A box, each
summed by two device paths.  The reference compares its host Replica::E()
with EonDevice; the drop-in compares the reference-shaped plain kernel
(EonDevice) with the batched full-energy kernel, and prints one
"PLAIN,BATCHED,err: %24.20f %24.20f %24.20f" line per conformation.

The GPU test checks both columns against the oracle at the same box (the
reference's own fixtures), the err column against their difference, and the
start-up consistency lines of a normal run (main.cpp:1767-1789)."""
import os
import re

import pytest

from conftest import GOLDEN, conformation
from facade_helpers import run_sim, split_chains, write_config
from helpers import say

BOX = 118.4
LINE = re.compile(r"PLAIN,BATCHED,err: +(-?\d+\.\d{20}) +(-?\d+\.\d{20}) +(\d+\.\d{20})$")


def write_examples(root):
    """<root>/conf<i>/1a.pdb (UIM) and 1b.pdb (ubiquitin) of fixture i."""
    for i in range(1, 11):
        d = os.path.join(root, "conf%d" % i)
        os.makedirs(d)
        segs = split_chains(i, d)
        os.rename(segs["A"], os.path.join(d, "1a.pdb"))
        os.rename(segs["B"], os.path.join(d, "1b.pdb"))


def test_check_without_examples_fails_before_the_device(tmp_path):
    r = run_sim(["-c", "--check-dir", str(tmp_path / "nothing")], cwd=str(tmp_path))
    assert r.returncode == 1
    assert "performing check..." in r.stdout
    assert "cannot load example conformation 1" in r.stdout


def test_help_lists_check():
    r = run_sim(["-h"], cwd=GOLDEN)
    assert r.returncode == 0 and "\t-c:" in r.stdout


@pytest.mark.gpu
def test_check_mode_plain_against_batched(tmp_path, aas, mj_ff, oracle):
    write_examples(str(tmp_path / "data"))
    r = run_sim(["-c"], cwd=str(tmp_path))  # the reference's layout: data/conf<i> under the working directory
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [LINE.match(ln) for ln in r.stdout.splitlines() if ln.startswith("PLAIN,BATCHED")]
    assert len(rows) == 10 and all(rows), r.stdout
    for i, m in enumerate(rows, start=1):
        plain, batched, err = (float(x) for x in m.groups())
        sa = conformation(i, aas, box=BOX)
        want = oracle.OracleSystem(sa, mj_ff).energy(sa.pos)
        say("conformation %d: plain %.9f batched %.9f oracle %.9f" % (i, plain, batched, want))
        for got in (plain, batched):
            assert abs(got - want) <= 1e-5 * abs(want) + 1e-6, (i, got, want)
        assert abs(err - abs(plain - batched)) <= 1e-18 + 1e-12 * abs(plain)


@pytest.mark.gpu
def test_start_up_consistency_lines(tmp_path):
    split_chains(1, str(tmp_path))
    cfg = os.path.join(str(tmp_path), "cfg")
    write_config(cfg, {"mcsteps": 100, "resteps": 1, "replicas": 2, "samplefrequency": 100, "sampleafter": 0,
                       "boundry": 120, "temperaturemax": 320, "temperaturemin": 300, "prefix": "c"})
    r = run_sim(["-f", cfg, "--outdir", "o", "--pid", "5", "--seed", "7"], cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    gpu = re.search(r"GPU initial energy value: +(-?\d+\.\d{23}) kcal/mol", r.stdout)
    ab = re.search(r"Absolute error: +(\d+\.\d{23}) kcal/mol", r.stdout)
    rel = re.search(r"Relative error: +(\d+\.\d{23}) kcal/mol", r.stdout)
    assert gpu and ab and rel, r.stdout
    assert float(rel.group(1)) < 1e-5 and "WARNING: Inconsistency" not in r.stdout
