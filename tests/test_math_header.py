"""The device/host arithmetic header (remc_math.h) compiled for the host:
-ln(u) against numpy's log, and the Metropolis test in its -ln(u) form
against the exp form of the reference (Replica.cpp:438-459) away from the
decision boundary."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO

HARNESS = r"""
#include <cstdio>
#include <cstdlib>
#include "remc_math.h"
int main(int argc, char **argv)
{
    // stdin: lines "u delta T"; stdout: "-ln(u) metropolis exp_form"
    double u, d, T;
    while (std::scanf("%lf %lf %lf", &u, &d, &T) == 3) {
        int expf = d < 0.0 ? 1 : (u < remc_exp(-d * remc_beta(T)) ? 2 : 0);
        std::printf("%.17g %d %d\n", remc_neglog_u(u), remc_metropolis(d, T, u), expf);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("mathh")
    src = os.path.join(str(d), "h.cpp")
    exe = os.path.join(str(d), "h")
    with open(src, "w") as f:
        f.write(HARNESS)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    "-I", os.path.join(REPO, "remcdockingsimulator_amd", "csrc"), src, "-o", exe], check=True)
    return exe


def run(exe, rows):
    inp = "".join("%.17g %.17g %.17g\n" % r for r in rows)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    res = [ln.split() for ln in out if ln.strip()]
    return np.array([float(a) for a, _, _ in res]), np.array([int(b) for _, b, _ in res]), \
        np.array([int(c) for _, _, c in res])


def test_neglog_matches_numpy(harness):
    rng = np.random.default_rng(7)
    w = rng.integers(1, 2 ** 32, size=4000, dtype=np.uint64)
    u = np.concatenate([w.astype(np.float64) * 2.0 ** -32,
                        [2.0 ** -32, 0.5, 0.7071067811865476, 0.7071067811865475, 1.0 - 2.0 ** -32, 0.25]])
    got, _, _ = run(harness, [(x, 0.0, 300.0) for x in u])
    want = -np.log(u)
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-300)
    # |ln u| can be tiny near u = 1: compare absolutely there
    err = np.where(np.abs(want) > 1e-3, rel, np.abs(got - want))
    assert np.max(err) < 4e-16, np.max(err)


def test_zero_uniform_maps_to_exp_underflow_edge(harness):
    got, dec, _ = run(harness, [(0.0, 1.0, 300.0), (0.0, 1e6, 300.0)])
    assert got[0] == 745.1332191019412
    assert dec[0] == 2 and dec[1] == 0  # exp(-dE/kT) > 0 = u accepts; underflowed exp rejects


def test_metropolis_forms_agree_off_boundary(harness):
    rng = np.random.default_rng(11)
    n = 5000
    T = rng.uniform(250, 600, n)
    kT = 8.314472 * T / 4184.0
    u = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.float64) * 2.0 ** -32
    x = rng.uniform(-1, 8, n)  # dE in kT units
    d = x * kT
    _, dec, dec_exp = run(harness, list(zip(u, d, T)))
    margin = np.abs(np.log(np.maximum(u, 1e-300)) + x)
    far = margin > 1e-9
    assert np.array_equal(dec[far], dec_exp[far])
    assert set(np.unique(dec)) <= {0, 1, 2} and np.sum(dec == 2) > 100 and np.sum(dec == 0) > 100
