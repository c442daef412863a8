"""remc_sim end to end on the device: a short REMC run of the reference's
UIM/ubiquitin complex (conformation fixture 1, both chains left in their
docked pose) with two crowders, checked against the reference's output
formats ("simulation engine/src/main.cpp:634-713,1378-1521",
Replica.cpp:1717-1780) and, record by record, against the oracle: every
bound-conformation record's poses are rebuilt from the printed quaternions
and centres and its printed energies recomputed in FP64."""
import os
import re
import subprocess

import numpy as np
import pytest

from facade_helpers import DATA, python_molecules, run_sim, split_chains, write_config
from helpers import say
from remcdockingsimulator_amd import forcefield as ffm
from remcdockingsimulator_amd import model

pytestmark = pytest.mark.gpu

PLACE = [
    ("t", (0.0, 0.0, 0.0), (1.0, 0.0, 0.0, -1.0), "A", False),
    ("t", (0.0, 0.0, 0.0), (1.0, 0.0, 0.0, -1.0), "B", False),
    ("p", (100.0, 100.0, 100.0), (0.0, 1.0, 0.0, 0.29), "A", True),
    ("p", (120.0, 60.0, 90.0), (1.0, 1.0, 0.0, 0.5), "A", True),
]
PARAMS = {"mcsteps": 2000, "resteps": 4, "replicas": 4, "samplefrequency": 100, "sampleafter": 0,
          "boundry": 150, "temperaturemax": 420, "temperaturemin": 290, "prefix": "g1"}


def parse_bound(path):
    recs = []
    head = re.compile(r"^(\d+); (-?[\d.]+|-?nan|-?inf) \((-?[\d.]+)\); ([\d.]+)$")
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    assert lines[0] == "# iteration; complex free energy (replica free energy); temperature;"
    assert lines[1].startswith("# molecule: rotation(w,x,y,z) position(x,y,z)")
    i = 2
    while i < len(lines):
        m = head.match(lines[i])
        assert m, lines[i]
        poses = []
        for a in range(len(PLACE)):
            t = lines[i + 1 + a].split()
            assert t[0] == ("0" if a < 2 else "1") and t[1] == "%d:" % a, lines[i + 1 + a]
            poses.append([float(v) for v in t[2:]])
        recs.append((int(m.group(1)), float(m.group(2)), float(m.group(3)), float(m.group(4)), np.array(poses)))
        i += 1 + len(PLACE)
    return recs


def test_remc_sim_short_run(tmp_path, oracle, aas):
    tmp = str(tmp_path)
    paths = split_chains(1, tmp)
    cfg = os.path.join(tmp, "cfg")
    write_config(cfg, PARAMS, PLACE)
    say("remc_sim short run")
    r = run_sim(["-f", cfg, "--seed", "5", "--outdir", "out", "--pid", "9", "--final-pdb", "final.pdb"], cwd=tmp)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    out = os.path.join(tmp, "out")
    pre = os.path.join(out, "g1_repulsive_9_")

    T = 290.0 * (420.0 / 290.0) ** (np.arange(4) / 3.0)
    fb = open(pre + "fractionBound").read().splitlines()
    assert fb[:2] == ["# Fraction Bound", "#Iteration InstantaniousAve CumulativeAve"]
    assert fb[2] == "Iteration:  " + "".join("%0.1fKi %0.1fKc " % (np.float32(t), np.float32(t)) for t in T) + "\t"
    rows = fb[3:]
    assert [int(x.split(":")[0]) for x in rows] == [500, 1000, 1500, 2000]
    for x in rows:
        vals = [float(v) for v in re.findall(r"\| (\S+) (\S+)", x) for v in v]
        assert len(vals) == 8 and all(0.0 <= v <= 1.0 for v in vals), x

    ar = open(pre + "acceptance_ratios").read().splitlines()
    assert ar[0] == "# Iteration AcceptanceRatio" and ar[1].startswith("Iteration:  ")
    assert [int(x.split(":")[0]) for x in ar[2:]] == [500, 1000, 1500, 2000]

    ex = open(pre + "exchange_freq").read().splitlines()
    assert ex[0] == "# Iteration continous_ratio, inst_ratio"
    assert len(ex) == 4
    for k, x in enumerate(ex[1:], start=1):
        assert re.match(r"^ {9}%d \d\.\d{4} \d\.\d{4}$" % k, x), x

    idx = open(pre + "fileindex").read()
    assert idx.endswith(" 2 seg_A.pdb Y\n 3 seg_A.pdb Y\n")

    # every bound record against the oracle
    recs = parse_bound(pre + "boundconformations")
    say("bound records", len(recs))
    assert recs, "the docked complex should be sampled bound at least once"
    mols = python_molecules(paths, aas, PLACE)
    sa0 = model.pack_molecules(mols, 150.0)
    ff = ffm.ForceField.mj1996(DATA)
    osys = oracle.OracleSystem(sa0, ff)
    body = np.repeat(np.arange(sa0.n_bodies), np.diff(sa0.body_start))
    # a record's rotation is relative to the input structure (before the
    # configuration's r(...) rotation), so poses are rebuilt from the input
    # molecules' relative coordinates
    rel0 = np.concatenate([model.read_ca_pdb(paths[seg], aas).relative for _, _, _, seg, _ in PLACE]).astype(np.float64)
    for step, complex_e, pot, temp, poses in recs:
        assert step % 100 == 0 and 0 <= step < 2000
        assert any(abs(temp - float("%0.1f" % np.float32(t))) < 1e-9 for t in T)
        pos = np.empty_like(rel0)
        rel = np.empty_like(rel0)
        for a in range(sa0.n_bodies):
            q = poses[a, :4] / np.linalg.norm(poses[a, :4])
            sel = body == a
            rel[sel] = model.quat_rotate(q, rel0[sel].astype(np.float32)).astype(np.float64)
            pos[sel] = rel[sel] + poses[a, 4:]
        e_full = osys.energy_d(pos)
        e_nc = osys.energy_noncrowder(pos.astype(np.float32))
        # the record prints rotations and centres to 6 decimals: a bead moves by up
        # to ~1e-6 A from the centre and ~2e-6 rad x |rel| from the rotation; the
        # energy error this allows is bounded with the oracle's bead gradient
        g, _, _ = osys.gradient(pos, rel)
        slack = 2.0 * np.sum(np.linalg.norm(g, axis=1) * (1e-6 + 2e-6 * np.linalg.norm(rel, axis=1)))
        assert abs(pot - e_full) <= 1e-4 + slack + 1e-5 * abs(e_full), (step, pot, e_full, slack)
        assert abs(complex_e - (e_full + e_nc)) <= 1e-4 + 2 * slack + 1e-5 * abs(e_full + e_nc), \
            (step, complex_e, e_full + e_nc, slack)
        assert complex_e < model_bound_threshold()

    pdb = open(os.path.join(tmp, "final.pdb")).read()
    assert pdb.count("\nATOM  ") + pdb.startswith("ATOM") == sa0.n_beads and pdb.endswith("END \n")


def model_bound_threshold():
    return -1.1844  # BOUND_ENERGY_VALUE (definitions.h:84)


def test_remc_sim_checkpoint_resume_is_exact(tmp_path, aas):
    """A run that checkpoints after exchange round 2 of 4, and a second process
    resumed from that checkpoint, end in the same state and write the same
    output lines for rounds 3 and 4 (main.cpp:89-122 intent, working load)."""
    tmp = str(tmp_path)
    split_chains(3, tmp)
    cfg = os.path.join(tmp, "cfg")
    write_config(cfg, PARAMS, PLACE)
    common = ["-bx", "1024", "--coop", "1", "--checkpoint-dir", "ck"]
    say("remc_sim run with a checkpoint")
    a = run_sim(["-f", cfg, "--seed", "11", "--outdir", "outA", "--pid", "21", "--checkpoint", "1000",
                 "--final-pdb", "finalA.pdb"] + common, cwd=tmp)
    assert a.returncode == 0, a.stdout[-3000:] + a.stderr[-3000:]
    ckf = os.path.join(tmp, "ck", "g1_repulsive_21")
    assert os.path.exists(ckf)
    say("remc_sim resumed from the checkpoint")
    b = run_sim(["-z", ckf, "--outdir", "outB", "--final-pdb", "finalB.pdb"] + common, cwd=tmp)
    assert b.returncode == 0, b.stdout[-3000:] + b.stderr[-3000:]
    assert "Resuming after exchange round 2" in b.stdout

    def body(path):
        return open(os.path.join(tmp, path)).read().splitlines()[1:]  # drop the REMARK with the file name

    assert body("finalA.pdb") == body("finalB.pdb")
    pa, pb = (os.path.join(tmp, d, "g1_repulsive_21_") for d in ("outA", "outB"))
    for name in ("fractionBound", "acceptance_ratios"):
        la = open(pa + name).read().splitlines()
        lb = open(pb + name).read().splitlines()
        assert lb == [x for x in la if re.match(r"^ +(1500|2000):", x)], name
    ea = open(pa + "exchange_freq").read().splitlines()
    eb = open(pb + "exchange_freq").read().splitlines()
    assert eb == ea[-1:]
    ba = open(pa + "boundconformations").read().splitlines()[2:]
    bb = open(pb + "boundconformations").read().splitlines()
    cut = next((i for i, x in enumerate(ba) if re.match(r"^\d+;", x) and int(x.split(";")[0]) >= 1000), len(ba))
    assert bb == ba[cut:]


def test_replica_device_api(tmp_path, oracle):
    """The reference's per-replica device calls through the facade
    (Replica::ReplicaDataToDevice, EonDevice, MCSearch, MoleculeDataToDevice,
    sample, FreeDevice; Replica.cpp:368-478,1308-1417,1530-1655,1717-1780):
    the plain kernel and the full energy agree; MCSearch takes the oracle's
    steps from the same counters (seed 1234 mixed with label 3) and brings
    back the state, potential and counters; a host move reaches the device;
    sample writes the reference's record."""
    from facade_api_harness import DATA as FDATA
    from facade_api_harness import build, read_dump

    tmp = str(tmp_path)
    exe = build(tmp)
    paths = split_chains(1, tmp)
    say("facade replica device api")
    r = subprocess.run([exe, "gpu", FDATA, paths["A"], paths["B"], tmp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.strip().split("\n")}
    e_plain, e_full = (float.fromhex(v) for v in out["energy"])
    assert abs(e_plain - e_full) <= 1e-5 * abs(e_full)
    pot, acc, acc_a, rej = float.fromhex(out["search"][0]), *(int(v) for v in out["search"][1:])
    assert acc + acc_a + rej == 600
    after = float.fromhex(out["after"][0])
    kt = ffm.kT(320.0)
    assert abs(pot - after) <= 1e-5 * abs(after) + 1e-4 * kt
    moved_plain, moved_full = (float.fromhex(v) for v in out["moved"])
    assert abs(moved_plain - moved_full) <= 1e-5 * abs(moved_full) and moved_full != after

    # the oracle from the same initial conformation and counters
    d0, d1 = read_dump(os.path.join(tmp, "initial.txt")), read_dump(os.path.join(tmp, "searched.txt"))
    n = len(d0["x"])
    body = np.repeat(np.arange(len(d0["crowder"])), np.diff(d0["body_start"]))
    sa = model.SystemArrays(
        body_start=np.array(d0["body_start"], np.int32), type=np.array(d0["type"], np.int32),
        charge=np.array(d0["charge"], np.float32), sigma=np.array(d0["sigma"], np.float32),
        crowder=np.array(d0["crowder"], np.uint8), box=d0["box"][0],
        pos=np.stack([d0["x"], d0["y"], d0["z"]], 1).astype(np.float32),
        rel=np.stack([d0["rx"], d0["ry"], d0["rz"]], 1).astype(np.float32),
        center=np.stack([d0["cx"], d0["cy"], d0["cz"]], 1).astype(np.float32),
        quat=np.array(d0["quat"]).reshape(-1, 4))
    assert body.shape[0] == n
    ff = ffm.ForceField.mj1996(DATA)
    o = oracle.OracleSystem(sa, ff)
    key = (1234 ^ (0x9E3779B97F4A7C15 * 4)) & 0xFFFFFFFFFFFFFFFF
    rep = oracle.OracleReplica(o, 320.0, 2.0, 0.2, 0)
    rec = rep.mc(seed=key, nsteps=600, mode=0)
    pos1 = np.stack([d1["x"], d1["y"], d1["z"]], 1).astype(np.float32)
    if np.sum(rec["accepted"] == 0) == rej and np.array_equal(pos1, rep.pos):
        assert np.array_equal(np.array(d1["quat"]).reshape(-1, 4), rep.quat)
    else:
        assert np.min(rec["margin"]) < 1e-4  # a near-threshold decision
    assert abs(pot - o.energy(pos1)) <= 1e-5 * abs(pot) + 1e-4 * kt

    with open(os.path.join(tmp, "bound.txt")) as f:
        lines = [ln.rstrip("\n") for ln in f]
    assert re.match(r"^77; -?[\d.]+ \(-?[\d.]+\); 320\.0$", lines[0]), lines[0]
    assert [ln.split()[:2] for ln in lines[1:4]] == [["0", "0:"], ["0", "1:"], ["1", "2:"]]
    assert len(lines) == 4  # the unbound sample wrote nothing
    assert out["sample"] == ["1", "0", "1", "2"] and out["freed"] == ["1"]
