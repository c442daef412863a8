"""The device/host arithmetic header (remc_math.h) compiled for the host:
-ln(u) against numpy's log, and the Metropolis test in its -ln(u) form
against the exp form of the reference (Replica.cpp:438-459) away from the
decision boundary."""
import ctypes
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


FB_HARNESS = r"""
#include <cstdio>
#include "remc_math.h"
int main()
{
    // stdin: "s t" (sin and cos of t turns) or "d step rng_id k0 k1 n_bodies" (force-biased draws)
    char kind;
    while (std::scanf(" %c", &kind) == 1) {
        if (kind == 's') {
            double t, sn, cs;
            if (std::scanf("%lf", &t) != 1) return 1;
            remc_sincos_turns(t, &sn, &cs);
            std::printf("%.17g %.17g\n", sn, cs);
        } else {
            unsigned long long step;
            unsigned rng_id, k0, k1;
            int nb;
            if (std::scanf("%llu %u %u %u %d", &step, &rng_id, &k0, &k1, &nb) != 5) return 1;
            remc_fb_draws d = remc_fb_step_draws(step, rng_id, k0, k1, nb);
            std::printf("%d %d %.17g %.17g %.17g %.17g\n", d.body, d.rotate, d.n[0], d.n[1], d.n[2], d.nlu);
        }
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def fb_harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("mathfb")
    src = os.path.join(str(d), "fb.cpp")
    exe = os.path.join(str(d), "fb")
    with open(src, "w") as f:
        f.write(FB_HARNESS)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    "-I", os.path.join(REPO, "remcdockingsimulator_amd", "csrc"), src, "-o", exe], check=True)
    return exe


def test_sincos_turns_against_extended_precision(fb_harness):
    """sin and cos of 2 pi t against the x87 extended-precision libm (about
    1e-19), over [0, 1), the quadrant boundaries and some large t."""
    rng = np.random.default_rng(5)
    t = np.concatenate([rng.integers(0, 2 ** 32, 4000).astype(np.float64) * 2.0 ** -32,
                        [0.0, 0.125, 0.25, 0.375, 0.5, 0.75, 1.0 - 2.0 ** -32, 2.0 ** -32, 0.2499999999, 0.7500000001],
                        rng.uniform(-50, 50, 200)])
    out = subprocess.run([fb_harness], input="".join("s %.17g\n" % x for x in t), capture_output=True, text=True,
                         check=True).stdout.split()
    got = np.array([float(v) for v in out]).reshape(-1, 2)
    tl = t.astype(np.longdouble)
    two_pi = np.arctan(np.longdouble(1)) * 8
    frac = tl - np.rint(tl)  # exact in binary64 already; the reduction the header makes
    want_s = np.sin(two_pi * frac).astype(np.float64)
    want_c = np.cos(two_pi * frac).astype(np.float64)
    assert np.max(np.abs(got[:, 0] - want_s)) < 3e-16
    assert np.max(np.abs(got[:, 1] - want_c)) < 3e-16
    # exact values at the quadrant points
    q = got[4000:4010]
    assert abs(q[1, 0] - np.sqrt(0.5)) < 2e-16 and abs(q[1, 1] - np.sqrt(0.5)) < 2e-16
    assert q[0, 0] == 0.0 and q[0, 1] == 1.0 and q[2, 0] == 1.0 and q[2, 1] == 0.0
    assert q[4, 0] == 0.0 and q[4, 1] == -1.0 and q[5, 0] == -1.0 and q[5, 1] == 0.0


def test_fb_draws_follow_the_oracle_layout(fb_harness, oracle):
    """remc_fb_step_draws against the oracle's draws (orc_mc_fb_steps): Philox
    words of tags 0x46420000 and +1, body, move type, Box-Muller normals from
    libm and -ln u."""
    L = oracle.lib()
    rng = np.random.default_rng(9)
    cases = [(int(s), int(r), int(k0), int(k1), int(nb)) for s, r, k0, k1, nb in zip(
        rng.integers(0, 2 ** 40, 500), rng.integers(0, 512, 500), rng.integers(0, 2 ** 32, 500),
        rng.integers(0, 2 ** 32, 500), rng.integers(1, 3000, 500))]
    out = subprocess.run([fb_harness], input="".join("d %d %d %d %d %d\n" % c for c in cases), capture_output=True,
                         text=True, check=True).stdout.strip().split("\n")
    u32 = ctypes.c_uint32 * 4
    for line, (step, rid, k0, k1, nb) in zip(out, cases):
        v = line.split()
        words = []
        for tag in (0x46420000, 0x46420001):
            ctr = u32(step & 0xFFFFFFFF, step >> 32, rid, tag)
            key = (ctypes.c_uint32 * 2)(k0, k1)
            w = u32()
            L.orc_philox4x32_10(ctr, key, w)
            words.append([L.orc_u01(w[i]) for i in range(4)] + [w[0]])
        body = (words[0][4] * nb) >> 32
        u1, u2, u3, u4, u = words[0][2], words[0][3], words[1][0], words[1][1], words[1][2]
        rho1, rho2 = np.sqrt(-2 * np.log(1 - u1)), np.sqrt(-2 * np.log(1 - u3))
        n = [rho1 * np.cos(2 * np.pi * u2), rho1 * np.sin(2 * np.pi * u2), rho2 * np.cos(2 * np.pi * u4)]
        assert int(v[0]) == body and int(v[1]) == int(words[0][1] < 0.5)
        for k in range(3):
            assert abs(float(v[2 + k]) - n[k]) < 1e-14 * max(1.0, abs(n[k]))
        assert abs(float(v[5]) - (-np.log(u) if u > 0 else 745.1332191019412)) < 1e-14 * max(1.0, -np.log(max(u, 1e-300)))


FMOD_HARNESS = r"""
#include <cstdio>
#include "remc_math.h"
int main()
{
    // stdin: lines "x L" (float bit patterns); stdout: remc_fmodf(x, L) bit pattern
    unsigned int xb, lb;
    while (std::scanf("%u %u", &xb, &lb) == 2) {
        float x, l;
        std::memcpy(&x, &xb, 4);
        std::memcpy(&l, &lb, 4);
        const float r = remc_fmodf(x, l);
        unsigned int rb;
        std::memcpy(&rb, &r, 4);
        std::printf("%u\n", rb);
    }
    return 0;
}
"""


def test_header_fmodf_is_exact(tmp_path):
    """remc_fmodf of the header (centre wrapping of every translation, on
    the GPU and the host) equals numpy.fmod bit for bit: the branch-free
    wraps of [L, 4L) and the subtraction loop beyond, both signs."""
    src, exe = str(tmp_path / "f.cpp"), str(tmp_path / "f")
    with open(src, "w") as f:
        f.write("#include <cstring>\n" + FMOD_HARNESS)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    "-I", os.path.join(REPO, "remcdockingsimulator_amd", "csrc"), src, "-o", exe], check=True)
    rng = np.random.default_rng(11)
    xs, ls = [], []
    for box in (150.0, 263.37, 300.0, 600.0, 7.5):
        b = np.float32(box)
        for lo, hi in ((-4, 4), (0, 2), (1, 4), (3.999, 4.001), (-100, 100)):
            x = rng.uniform(lo * box, hi * box, 3000).astype(np.float32)
            xs.append(x)
            ls.append(np.full(x.shape, b, np.float32))
        edges = np.array([b, 2 * b, 3 * b, 4 * b, -b, -2 * b, -3 * b, -4 * b], np.float32)
        edges = np.concatenate([edges, np.nextafter(edges, np.float32(0)), np.nextafter(edges, np.float32(np.inf))])
        xs.append(edges.astype(np.float32))
        ls.append(np.full(edges.shape, b, np.float32))
    x = np.concatenate(xs)
    L = np.concatenate(ls)
    inp = "\n".join("%d %d" % (a, c) for a, c in zip(x.view(np.uint32), L.view(np.uint32))) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
    got = np.array([int(v) for v in out], dtype=np.uint32)
    want = np.fmod(x, L).view(np.uint32)
    assert np.array_equal(got, want), int(np.sum(got != want))
