// remc_math.h -- arithmetic shared by the HIP kernels and the library's host
// code (the host-side reference protocol of remc_mc_steps_plain).  Every
// function here must give bit-identical results on gfx950 and on the host, so
// contraction is switched off and transcendental functions are self-written
// from IEEE basic operations and fused multiply-adds.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define REMC_HD __host__ __device__ inline
#else
#include <math.h>
#include <string.h>
#define REMC_HD static inline
#endif

#define REMC_R_GAS 8.314472 /* Rgas, definitions.h:16 */
#define REMC_TAG_MC 0x4D430000u
#define REMC_TAG_EX 0x58000000u

#pragma clang fp contract(off)

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of the
// multiply-xor bijection with Weyl-sequence key schedule.
struct remc_u32x4 {
    uint32_t v[4];
};

REMC_HD remc_u32x4 remc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    remc_u32x4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

// uniform [0,1) with 32-bit resolution (gsl_rng_uniform on mt19937 resolution)
REMC_HD double remc_u01(uint32_t w) { return (double)w * (1.0 / 4294967296.0); }

// The draws of MC step `step` of replica `rng_id`, in the reference's draw
// order (Replica.cpp:386-451): molecule, move type, three vector components,
// Metropolis uniform.  The Metropolis uniform is drawn at a fixed counter and
// used only when dE >= 0, as the reference draws it only then (:451).
struct remc_draws {
    int body;
    int rotate;
    double a, b, c, u;
};

REMC_HD remc_draws remc_step_draws(uint64_t step, uint32_t rng_id, uint32_t k0, uint32_t k1, int n_bodies)
{
    remc_u32x4 w0 = remc_philox((uint32_t)step, (uint32_t)(step >> 32), rng_id, REMC_TAG_MC, k0, k1);
    remc_u32x4 w1 = remc_philox((uint32_t)step, (uint32_t)(step >> 32), rng_id, REMC_TAG_MC + 1u, k0, k1);
    remc_draws d;
    d.body = (int)(((uint64_t)w0.v[0] * (uint64_t)n_bodies) >> 32);
    d.rotate = remc_u01(w0.v[1]) < 0.5;
    d.a = remc_u01(w0.v[2]) - 0.5;
    d.b = 0.5 - remc_u01(w0.v[3]);
    d.c = remc_u01(w1.v[0]) - 0.5;
    d.u = remc_u01(w1.v[1]);
    return d;
}

REMC_HD double remc_exchange_u(uint64_t round, uint32_t pair, uint32_t system_id, uint32_t k0, uint32_t k1)
{
    remc_u32x4 w = remc_philox((uint32_t)round, (uint32_t)(round >> 32), pair, REMC_TAG_EX + (system_id & 0xFFFFFFu), k0, k1);
    return remc_u01(w.v[0]);
}

// exp(x) in FP64: Cody-Waite reduction x = k ln2 + r (fdlibm split of ln2),
// Taylor series of exp(r) to degree 13 in fused Horner form, exact 2^k scale.
REMC_HD double remc_exp(double x)
{
    if (x != x) return x;
    if (x > 709.0) return __builtin_huge_val();
    if (x < -708.0) return 0.0;
    const double log2e = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double k = rint(x * log2e);
    double r = fma(-k, ln2_hi, x);
    r = fma(-k, ln2_lo, r);
    double p = 1.0 / 6227020800.0;
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    uint64_t bits = (uint64_t)((int64_t)((int)k + 1023)) << 52;
#if defined(__HIP_DEVICE_COMPILE__)
    double two_k = __longlong_as_double((long long)bits);
#else
    double two_k;
    memcpy(&two_k, &bits, sizeof two_k);
#endif
    return p * two_k;
}

// fmodf for finite x, L > 0, computed exactly by subtracting the largest
// power-of-two multiple of L (each step exact by Sterbenz's lemma).
REMC_HD float remc_fmodf(float x, float L)
{
    float ax = fabsf(x);
    if (!(ax >= L)) return x;           // also NaN
    if (!(ax <= 3.4028234663852886e38f)) return x; // inf: no finite remainder, leave it
    if (ax < 4.0f * L) {
        // the wraps that occur (L <= |x| < 4L): |x| - L, |x| - 2L and, from
        // [L, 2L), that minus L are exact (Sterbenz), the same numbers the
        // subtraction loop below forms, without its branches
        const float r2 = ax - 2.0f * L;
        const float r = ax < 2.0f * L ? ax - L : (r2 < L ? r2 : r2 - L);
        return x < 0.0f ? -r : r;
    }
    float r = ax;
    while (r >= L) {
        float d = L;
        while (2.0f * d <= r) d = 2.0f * d;
        r = r - d;
    }
    return x < 0.0f ? -r : r;
}

// Quaternion product a*b (Quaternion.cpp:35-47), (w,x,y,z).
REMC_HD void remc_qmul(const double a[4], const double b[4], double o[4])
{
    double w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    double x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    double y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    double z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    o[0] = w; o[1] = x; o[2] = y; o[3] = z;
}

// Rotation of a float vector by a unit quaternion in FP64, rounded to float
// (Quaternion::rotateVector, Quaternion.cpp:84-101).  The coefficients are
// passed precomputed: m = {w2+x2-y2-z2, xy-wz, wy+xz, wz+xy, yz-wx,
// w2-x2+y2-z2, xz-wy, yz+wx, w2-x2-y2+z2}.
REMC_HD void remc_qcoeffs(const double q[4], double m[9])
{
    double w = q[0], x = q[1], y = q[2], z = q[3];
    double w2 = w * w, x2 = x * x, y2 = y * y, z2 = z * z;
    double wx = w * x, wy = w * y, wz = w * z, xy = x * y, xz = x * z, yz = y * z;
    m[0] = w2 + x2 - y2 - z2;
    m[1] = xy - wz;
    m[2] = wy + xz;
    m[3] = wz + xy;
    m[4] = yz - wx;
    m[5] = w2 - x2 + y2 - z2;
    m[6] = xz - wy;
    m[7] = yz + wx;
    m[8] = w2 - x2 - y2 + z2;
}

REMC_HD void remc_qrotate(const double m[9], float vx_, float vy_, float vz_, float *ox, float *oy, float *oz)
{
    double vx = vx_, vy = vy_, vz = vz_;
    *ox = (float)(vx * m[0] + (vy * m[1] + vz * m[2]) * 2.0);
    *oy = (float)((vx * m[3] + vz * m[4]) * 2.0 + vy * m[5]);
    *oz = (float)((vx * m[6] + vy * m[7]) * 2.0 + vz * m[8]);
}

// Rotation quaternion of a rotate move: axis normalize(a,b,c) in FP64
// (Vector3double::normalizeInPlace), q = (cos, sin*axis) normalised
// (Molecule::rotateQ, Molecule.cpp:290-296).
REMC_HD void remc_move_quat(double a, double b, double c, double ch, double sh, double q[4])
{
    double mag = sqrt(a * a + b * b + c * c);
    double ax = a / mag, ay = b / mag, az = c / mag;
    q[0] = ch; q[1] = sh * ax; q[2] = sh * ay; q[3] = sh * az;
    double qm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    q[0] = q[0] / qm; q[1] = q[1] / qm; q[2] = q[2] / qm; q[3] = q[3] / qm;
}

// Translation vector of a translate move (Replica::translate,
// Replica.cpp:333-337): t = step * normalize((float)a,(float)b,(float)c) in
// FP32.  It depends on the draws only, so it can be formed ahead of the step.
REMC_HD void remc_move_translation(double a, double b, double c, float step, float *tx, float *ty, float *tz)
{
    float vx = (float)a, vy = (float)b, vz = (float)c;
    float mag = sqrtf(vx * vx + vy * vy + vz * vz);
    if (mag != 0.0f) { vx = vx / mag; vy = vy / mag; vz = vz / mag; }
    *tx = vx * step; *ty = vy * step; *tz = vz * step;
}

// New centre: c + t wrapped with fmod(. + L, L) per axis (Replica.cpp:358-362).
REMC_HD void remc_wrap_centre(float cx, float cy, float cz, float tx, float ty, float tz, float box,
                              float *px, float *py, float *pz)
{
    float x = cx + tx, y = cy + ty, z = cz + tz;
    if (box > 0.0f) {
        x = remc_fmodf(x + box, box);
        y = remc_fmodf(y + box, box);
        z = remc_fmodf(z + box, box);
    }
    *px = x; *py = y; *pz = z;
}

REMC_HD void remc_move_centre(double a, double b, double c, float step, float box, float cx, float cy, float cz,
                              float *px, float *py, float *pz)
{
    float tx, ty, tz;
    remc_move_translation(a, b, c, step, &tx, &ty, &tz);
    remc_wrap_centre(cx, cy, cz, tx, ty, tz, box, px, py, pz);
}

// -ln(u) in FP64 for a uniform u in [0, 1): u = m 2^e with m in
// [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716,
// the odd series of atanh to s^23 (truncation below 1e-19 relative), and
// e ln2 with the fdlibm split of ln2.  u = 0 maps to the largest x with
// exp(-x) > 0 in FP64, so that the Metropolis test below accepts exactly the
// moves u < exp(-x) accepts when u = 0.
REMC_HD double remc_neglog_u(double u)
{
    if (!(u > 0.0)) return 745.1332191019412;
    int64_t bits;
#if defined(__HIP_DEVICE_COMPILE__)
    bits = __double_as_longlong(u);
#else
    memcpy(&bits, &u, sizeof bits);
#endif
    int e = (int)((bits >> 52) & 0x7ff) - 1023; // u >= 2^-32 here: never subnormal
    int64_t mb = (bits & 0x000FFFFFFFFFFFFFLL) | 0x3FF0000000000000LL;
    double m;
#if defined(__HIP_DEVICE_COMPILE__)
    m = __longlong_as_double(mb);
#else
    memcpy(&m, &mb, sizeof m);
#endif
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, s2, 1.0 / 21.0);
    p = fma(p, s2, 1.0 / 19.0);
    p = fma(p, s2, 1.0 / 17.0);
    p = fma(p, s2, 1.0 / 15.0);
    p = fma(p, s2, 1.0 / 13.0);
    p = fma(p, s2, 1.0 / 11.0);
    p = fma(p, s2, 1.0 / 9.0);
    p = fma(p, s2, 1.0 / 7.0);
    p = fma(p, s2, 1.0 / 5.0);
    p = fma(p, s2, 1.0 / 3.0);
    const double lm = fma(2.0 * s * s2, p, 2.0 * s);
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return -(fma((double)e, ln2_hi, fma((double)e, ln2_lo, lm)));
}

// Metropolis test in FP64 (Replica.cpp:438-459) with the uniform given as
// L = -ln(u): 1 = accept (dE < 0), 2 = Boltzmann accept (u < exp(-dE beta),
// i.e. dE beta < L), 0 = reject.  beta = 4184 / (R T) in mol/kcal.
REMC_HD int remc_metropolis_l(double delta, double beta, double L)
{
    if (delta < 0.0) return 1;
    return delta * beta < L ? 2 : 0;
}

REMC_HD double remc_beta(double temperature) { return 4184.0 / (REMC_R_GAS * temperature); }

REMC_HD int remc_metropolis(double delta, double temperature, double u)
{
    return remc_metropolis_l(delta, remc_beta(temperature), remc_neglog_u(u));
}

// sin and cos of 2 pi t in FP64 for finite t: t reduced to f = t - rint(t),
// then to r = f - k/4 with |r| <= 1/8 (both exact), x = 2 pi r, |x| <= pi/4,
// and the Taylor series of sin to x^15 and of cos to x^16 (truncation below
// 1e-16), rotated by the quadrant k.
REMC_HD void remc_sincos_turns(double t, double *sn, double *cs)
{
    const double f = t - rint(t);
    const double k = rint(4.0 * f);
    const double x = (f - 0.25 * k) * 6.283185307179586476925;
    const double x2 = x * x;
    double ps = -1.0 / 1307674368000.0;
    ps = fma(ps, x2, 1.0 / 6227020800.0);
    ps = fma(ps, x2, -1.0 / 39916800.0);
    ps = fma(ps, x2, 1.0 / 362880.0);
    ps = fma(ps, x2, -1.0 / 5040.0);
    ps = fma(ps, x2, 1.0 / 120.0);
    ps = fma(ps, x2, -1.0 / 6.0);
    const double sx = fma(x * x2, ps, x);
    double pc = 1.0 / 20922789888000.0;
    pc = fma(pc, x2, -1.0 / 87178291200.0);
    pc = fma(pc, x2, 1.0 / 479001600.0);
    pc = fma(pc, x2, -1.0 / 3628800.0);
    pc = fma(pc, x2, 1.0 / 40320.0);
    pc = fma(pc, x2, -1.0 / 720.0);
    pc = fma(pc, x2, 1.0 / 24.0);
    pc = fma(pc, x2, -0.5);
    const double cx = fma(x2, pc, 1.0);
    switch (((int)k) & 3) {
    case 0: *sn = sx; *cs = cx; break;
    case 1: *sn = cx; *cs = -sx; break;
    case 2: *sn = -sx; *cs = -cx; break;
    default: *sn = -cx; *cs = sx; break;
    }
}

// Draws of a force-biased MC step: body, move type, three standard normals
// (Box-Muller on (1 - u1, u2) and (1 - u3, u4)) and -ln u of the acceptance
// uniform, from the step's two Philox blocks under the tag 0x46420000.
#define REMC_TAG_FB 0x46420000u
struct remc_fb_draws {
    int body;
    int rotate;
    double n[3];
    double nlu;
};

REMC_HD remc_fb_draws remc_fb_step_draws(uint64_t step, uint32_t rng_id, uint32_t k0, uint32_t k1, int n_bodies)
{
    remc_u32x4 w0 = remc_philox((uint32_t)step, (uint32_t)(step >> 32), rng_id, REMC_TAG_FB, k0, k1);
    remc_u32x4 w1 = remc_philox((uint32_t)step, (uint32_t)(step >> 32), rng_id, REMC_TAG_FB + 1u, k0, k1);
    remc_fb_draws d;
    d.body = (int)(((uint64_t)w0.v[0] * (uint64_t)n_bodies) >> 32);
    d.rotate = remc_u01(w0.v[1]) < 0.5;
    const double rho1 = sqrt(2.0 * remc_neglog_u(1.0 - remc_u01(w0.v[2])));
    const double rho2 = sqrt(2.0 * remc_neglog_u(1.0 - remc_u01(w1.v[0])));
    double s2, c2, s4, c4;
    remc_sincos_turns(remc_u01(w0.v[3]), &s2, &c2);
    remc_sincos_turns(remc_u01(w1.v[1]), &s4, &c4);
    d.n[0] = rho1 * c2;
    d.n[1] = rho1 * s2;
    d.n[2] = rho2 * c4;
    d.nlu = remc_neglog_u(remc_u01(w1.v[2]));
    return d;
}

// Exchange test of ladder rungs (Ti, Tj) holding energies (Ei, Ej)
// (main.cpp:1440-1441).
REMC_HD int remc_exchange_accept(double Ti, double Tj, double Ei, double Ej, double u)
{
    double delta = (1.0 / Ti - 1.0 / Tj) * (Ei - Ej) * (4184.0 / REMC_R_GAS);
    return delta >= 0.0 || u < remc_exp(delta);
}

#pragma clang fp contract(fast)
