"""The PDB files the drop-in writes, pinned to the reference's own format
strings rather than to this build's Python model: every line must parse with
the record's C format and print back byte for byte through it.

- single-conformation PDB, Replica::saveAsSinglePDB (Replica.cpp:495-530):
  remc_sim --save-pdb and every remc_pdbgen sample;
- multi-model trajectory, writeToTrajectory ("post processor/postprocess/
  main.cpp:252-291"): remc_pdbgen --trajectory;
- remc_pdbgen's sample file names (pdbgen.cpp:222).

Beyond the layout: ATOM serials count from 1 across the file (within a model
for the trajectory), chain letters run A, B, ... per molecule, each TER
repeats its molecule's last serial, residue name, chain and residue number,
and the REMARK centres and rotations are the ones the records hold.  CPU only
(the dry run and the replay tool never touch a GPU)."""
import os
import re
import subprocess

import pytest

from conftest import REPO
from facade_helpers import python_molecules, run_sim, split_chains, write_config
from test_pdbgen import PDBGEN, write_inputs

PARAMS = {"mcsteps": 2000, "resteps": 4, "replicas": 4, "samplefrequency": 100, "sampleafter": 0,
          "boundry": 120, "temperaturemax": 420, "temperaturemin": 290, "prefix": "t1"}

# the reference's records (C format strings as written in its sources)
ATOM = "ATOM  %5d %4s%C%3s %C%4d%C  %8.3f%8.3f%8.3f%6.2f%6.2f\n"      # Replica.cpp:519
TER = "TER   %5d      %3s %C%4d \n"                                    # Replica.cpp:524
END = "END \n"                                                          # Replica.cpp:528
R_NAME = "REMARK %s \n"                                                 # Replica.cpp:499
R_POT = "REMARK potential: %0.10f \n"                                   # Replica.cpp:500
R_TEMP = "REMARK temperature: %5.1f \n"                                 # Replica.cpp:501
R_MOL = "REMARK Molecule: %d\n"                                         # Replica.cpp:505
R_ROT = "REMARK Rotation relative to input Q(w,x,y,z): %f %f %f %f\n"   # Replica.cpp:506
R_CEN = "REMARK Centriod position P(x,y,z): %f %f %f\n"                 # Replica.cpp:507
T_TITLE = "TITLE     frame t=%7.3f\n"                                   # postprocess/main.cpp:258
T_MODEL = "MODEL%9d\n"                                                  # postprocess/main.cpp:260
T_REM = "REMARK potential,drms: %0.10f %0.10f \n"                       # postprocess/main.cpp:262
T_END = "ENDMDL\n"                                                      # postprocess/main.cpp:287

CONV = re.compile(r"%(\d*)(?:\.(\d+))?([dfsCc])")


def parse(fmt, line):
    """Values of `line` (without its newline) under the C format `fmt`, or
    None; a match must also print back through the format byte for byte."""
    pattern, kinds, pos = "", [], 0
    for m in CONV.finditer(fmt):
        pattern += re.escape(fmt[pos:m.start()])
        width, kind = m.group(1), m.group(3)
        if width and int(width) > 0:  # "%0.10f": a zero flag, no width
            pattern += "(.{%d})" % int(width)
        else:
            pattern += {"d": r"(-?\d+)", "f": r"(-?\d+\.\d+)", "s": r"(.*?)", "C": "(.)", "c": "(.)"}[kind]
        kinds.append(kind)
        pos = m.end()
    pattern += re.escape(fmt[pos:].rstrip("\n"))
    mm = re.fullmatch(pattern, line)
    if not mm:
        return None
    vals = []
    for g, k in zip(mm.groups(), kinds):
        try:
            vals.append(int(g) if k == "d" else float(g) if k == "f" else g.strip() if k == "s" else g)
        except ValueError:
            return None
    if (fmt.replace("%C", "%c") % tuple(vals)).rstrip("\n") != line:
        return None
    return vals


def expect(fmt, line, what):
    v = parse(fmt, line)
    assert v is not None, "%s: %r does not print back through %r" % (what, line, fmt)
    return v


def check_molecule_records(lines, i, n_mol, what):
    """ATOM/TER records of n_mol molecules from lines[i]; returns (next index,
    [(chain, [atom values])])."""
    serial, mols = 0, []
    for m in range(n_mol):
        chain = chr(ord("A") + m)
        atoms = []
        while lines[i].startswith("ATOM"):
            v = expect(ATOM, lines[i], what)
            serial += 1
            assert v[0] == serial and v[1] == "CA" and v[4] == chain, (what, lines[i])
            atoms.append(v)
            i += 1
        assert atoms, (what, "molecule", m, "has no ATOM records")
        t = expect(TER, lines[i], what)
        assert t == [serial, atoms[-1][3], chain, atoms[-1][5]], (what, lines[i])
        mols.append((chain, atoms))
        i += 1
    return i, mols


def check_single_pdb(path, name_remark, n_mol):
    lines = open(path).read().split("\n")
    assert lines[-1] == "", "the file ends with a newline"
    lines = lines[:-1]
    what = os.path.basename(path)
    assert expect(R_NAME, lines[0], what) == [name_remark]
    pot = expect(R_POT, lines[1], what)[0]
    temp = expect(R_TEMP, lines[2], what)[0]
    remarks = []
    i = 3
    for m in range(n_mol):
        assert expect(R_MOL, lines[i], what) == [m]
        q = expect(R_ROT, lines[i + 1], what)
        c = expect(R_CEN, lines[i + 2], what)
        remarks.append((q, c))
        i += 3
    i, mols = check_molecule_records(lines, i, n_mol, what)
    assert lines[i:] == [END.rstrip("\n")], (what, lines[i:])
    return pot, temp, remarks, mols


def test_dry_run_pdb_follows_the_reference_records(tmp_path, aas):
    paths = split_chains(1, str(tmp_path))
    cfg = os.path.join(str(tmp_path), "cfg")
    write_config(cfg, PARAMS)
    r = run_sim(["-f", cfg, "-q", "--save-pdb", "init.pdb", "--outdir", "o", "--pid", "3"], cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    pot, temp, remarks, mols = check_single_pdb(os.path.join(str(tmp_path), "init.pdb"), "init.pdb", 4)
    assert temp == 300.0
    # the centres in the REMARKs are the molecules' centres; the rotations are unit quaternions
    ref = python_molecules(paths, aas)
    for (q, c), (chain, atoms), mol in zip(remarks, mols, ref):
        assert abs(sum(x * x for x in q) - 1.0) < 1e-5
        assert [round(x, 6) for x in c] == [round(float(x), 6) for x in mol.center]
        assert len(atoms) == len(mol.types)


def test_pdbgen_samples_and_trajectory_follow_the_reference_records(tmp_path):
    tmp = str(tmp_path)
    write_inputs(tmp)
    r = subprocess.run([PDBGEN, "run_1_fileindex", "100", "0", "0", "--outdir", "res", "--trajectory", "traj.pdb"],
                       cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    outs = sorted(os.listdir(os.path.join(tmp, "res")))
    # pdbgen.cpp:222: "sample_%d_%0.1fK_%5.2f.pdb" of the sample number, temperature and energy
    assert outs == ["sample_%d_%0.1fK_%5.2f.pdb" % (1, 300.2, -5.12345)]
    pot, temp, remarks, mols = check_single_pdb(os.path.join(tmp, "res", outs[0]), "res/" + outs[0], 2)
    assert temp == 300.2 and abs(pot - (-5.12345)) < 1e-6
    assert [round(x, 6) for x in remarks[0][0]] == [0.9, 0.1, -0.3, 0.2] or abs(
        sum(x * x for x in remarks[0][0]) - 1.0) < 1e-5
    assert [c for c, _ in mols] == ["A", "B"]

    # the trajectory: one model per sample
    lines = open(os.path.join(tmp, "traj.pdb")).read().split("\n")[:-1]
    i, frames = 0, 0
    while i < len(lines):
        frames += 1
        assert expect(T_TITLE, lines[i], "trajectory") == [float(frames)]
        assert expect(T_MODEL, lines[i + 1], "trajectory") == [frames]
        expect(T_REM, lines[i + 2], "trajectory")
        i, mols_t = check_molecule_records(lines, i + 3, 2, "trajectory")
        assert [[a[6:9] for a in atoms] for _, atoms in mols_t] == [[a[6:9] for a in atoms] for _, atoms in mols]
        assert lines[i] == T_END.rstrip("\n")
        i += 1
    assert frames == 1


@pytest.mark.parametrize("fmt,line", [(ATOM, "ATOM      1  CA  ALA A   1      11.104   6.134  -6.504  1.00  1.00"),
                                      (TER, "TER       7      LYS A   7"),
                                      (R_POT, "REMARK potential: -12.5000000000"),
                                      (R_POT, "REMARK potential: -12.500000000 ")])
def test_checker_rejects_a_shifted_field(fmt, line):
    """The checker is not vacuous: the same records with one column moved
    (the standard PDB x column, a TER or potential REMARK without its trailing
    blank, nine decimals for ten) fail it."""
    assert parse(fmt, line) is None
