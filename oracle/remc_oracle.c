/*
 * remc_oracle.c -- CPU FP64 restatement of the reference REMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 *
 * Built with -ffp-contract=off so that every float/double operation is a single
 * IEEE-rounded step, as in the reference's scalar C++ (the move arithmetic must
 * match the GPU implementation bit for bit).
 *
 * Parity status of each function is recorded in the test files that pin it;
 * no function here is left without a check.
 */
#include "remc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_R_GAS 8.314472 /* Rgas (definitions.h:16) */

double orc_r_gas(void) { return ORC_R_GAS; }

/* Reference constants.  They are float expressions in definitions.h:11-24 and
 * are promoted to double inside Replica::E() (Replica.cpp:613,739); evaluate
 * them in float exactly as written, then widen. */
void orc_default_constants(orc_forcefield *ff)
{
    volatile float rgas = 8.314472f, t294 = 294.0f, k4184 = 4184.0f, one = 1.0f;
    volatile float btu = 0.948f, e = 1.602176487f, lj = 0.3507221006079f;
    float kbt = one / (t294 * rgas / k4184);       /* definitions.h:18 */
    float dh = btu * e * e;                        /* Replica.cpp:613  */
    ff->lambda = (double)0.159f;                   /* definitions.h:35 */
    ff->e0 = (double)-2.27f;                       /* definitions.h:34 */
    ff->xi = 10.0;                                 /* definitions.h:22 */
    ff->lj_cf = (double)lj;                        /* definitions.h:23 */
    ff->dh_cf = (double)dh;
    ff->kbt_cf = (double)kbt;
    ff->crowder_sigma = 6.0;                       /* Replica.cpp:586  */
    ff->crowder_cutoff = 15.0;                     /* definitions.h:89 */
    ff->eps_r = (double)1e-38f;                    /* definitions.h:61 */
}

/* ---------------------------------------------------------------- RNG ---- */
/* Philox4x32-10: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11.  Multipliers
 * 0xD2511F53/0xCD9E8D57, Weyl key increments 0x9E3779B9/0xBB67AE85. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 32-bit resolution uniform in [0,1), as gsl_rng_uniform on mt19937. */
double orc_u01(uint32_t w) { return (double)w * (1.0 / 4294967296.0); }

#define ORC_TAG_MC 0x4D430000u
#define ORC_TAG_EX 0x58000000u

void orc_step_words(uint64_t step, uint32_t rng_id, const uint32_t key[2], uint32_t w[6])
{
    uint32_t ctr[4] = {(uint32_t)step, (uint32_t)(step >> 32), rng_id, ORC_TAG_MC};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    w[0] = o[0]; w[1] = o[1]; w[2] = o[2]; w[3] = o[3];
    ctr[3] = ORC_TAG_MC + 1u;
    orc_philox4x32_10(ctr, key, o);
    w[4] = o[0]; w[5] = o[1];
}

double orc_exchange_uniform(uint64_t round, uint32_t pair, uint32_t system_id, const uint32_t key[2])
{
    uint32_t ctr[4] = {(uint32_t)round, (uint32_t)(round >> 32), pair, ORC_TAG_EX + (system_id & 0xFFFFFFu)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return orc_u01(o[0]);
}

/* --------------------------------------------------------------- math ---- */
/* exp(x): W. J. Cody and W. Waite, "Software Manual for the Elementary
 * Functions" (1980) argument reduction x = k ln2 + r with ln2 split in two
 * parts (the fdlibm split), exp(r) by its Taylor series to degree 13 in Horner
 * form with fused multiply-adds, then exact scaling by 2^k.  |r| <= 0.347 makes
 * the truncation error < 1e-17 relative.  Results below 2^-1021 flush to 0. */
double orc_exp(double x)
{
    if (x != x) return x;
    if (x > 709.0) return HUGE_VAL;
    if (x < -708.0) return 0.0;
    const double log2e = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double k = nearbyint(x * log2e);
    double r = fma(-k, ln2_hi, x);
    r = fma(-k, ln2_lo, r);
    double p = 1.0 / 6227020800.0;          /* 1/13! */
    p = fma(p, r, 1.0 / 479001600.0);       /* 1/12! */
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
    int64_t bits = (int64_t)((int)k + 1023) << 52;
    double two_k;
    memcpy(&two_k, &bits, sizeof two_k);
    return p * two_k;
}

/* Exact fmodf for finite x and L > 0 (result has the sign of x).  Repeatedly
 * subtracts the largest L*2^k not above the remainder; each subtraction is
 * exact by Sterbenz's lemma, so the result equals C fmodf. */
float orc_fmodf(float x, float L)
{
    float ax = fabsf(x);
    if (!(ax >= L)) return x;
    if (!(ax <= 3.4028234663852886e38f)) return x;
    float r = ax;
    while (r >= L) {
        float d = L;
        while (2.0f * d <= r) d = 2.0f * d;
        r = r - d;
    }
    return x < 0.0f ? -r : r;
}

/* ------------------------------------------------------------- energy ---- */
/* distance() with minimum image (Replica.cpp:589-600).  The reference rounds
 * with round() on the CPU and rintf() on the GPU (CudaFunctions.cu:436); this
 * restatement uses round-half-even, the GPU's choice; the two differ only for
 * a separation of exactly L/2 along an axis, where both images are equidistant. */
static double orc_min_image(double d, double L)
{
    return L > 0.0 ? d - L * nearbyint(d / L) : d;
}

double orc_pair_energy(const orc_forcefield *ff, double L,
                       double xi_, double yi_, double zi_, int ti, double qi, double si, int ci,
                       double xj_, double yj_, double zj_, int tj, double qj, double sj, int cj,
                       double *lj_raw, double *dh_raw)
{
    double dx = orc_min_image(xi_ - xj_, L);
    double dy = orc_min_image(yi_ - yj_, L);
    double dz = orc_min_image(zi_ - zj_, L);
    double r = sqrt(dx * dx + dy * dy + dz * dz) + ff->eps_r;
    double lj = 0.0, dh = 0.0;
    if (ci || cj) {
        /* repulsive crowder (Replica.cpp:640-690; CudaFunctions.cu:566-577) */
        if (r < ff->crowder_cutoff) {
            double t = ff->crowder_sigma / r;
            double t2 = t * t;
            lj = t2 * t2 * t2;
        }
    } else {
        /* Debye-Hueckel (Replica.cpp:701-706) */
        dh = qi * qj * exp(-r / ff->xi) / r;
        /* contact term (Replica.cpp:710-724) */
        double eij = ff->lambda * (ff->table[ti * ORC_NTYPES + tj] - ff->e0);
        double sij = 0.5 * (si + sj);
        double t = sij / r;
        double t2 = t * t;
        double x6 = t2 * t2 * t2;
        lj = -4.0 * eij * x6 * (x6 - 1.0);
        if (eij > 0.0 && r < sij * 1.122462048309372981433533049679)
            lj = -lj + 2.0 * eij;
    }
    if (lj_raw) *lj_raw = lj;
    if (dh_raw) *dh_raw = dh;
    return (lj * ff->lj_cf + dh * ff->dh_cf) * ff->kbt_cf;
}

static double orc_pair_ij(const orc_system *s, const orc_forcefield *ff,
                          double xi_, double yi_, double zi_, int i, int ci,
                          double xj_, double yj_, double zj_, int j, int cj,
                          double *lj, double *dh)
{
    return orc_pair_energy(ff, s->box, xi_, yi_, zi_, s->type[i], s->charge[i], s->sigma[i], ci,
                           xj_, yj_, zj_, s->type[j], s->charge[j], s->sigma[j], cj, lj, dh);
}

/* Replica::E() (Replica.cpp:603-745): molecule pairs mI < mJ, residue pairs,
 * LJ and DH accumulated separately in FP64 and converted at the end (:739). */
double orc_energy_full(const orc_system *s, const orc_forcefield *ff,
                       const float *x, const float *y, const float *z,
                       double *lj_sum, double *dh_sum)
{
    double ljacc = 0.0, dhacc = 0.0;
    for (int a = 0; a < s->n_bodies; ++a) {
        int ca = s->crowder[a];
        for (int b = a + 1; b < s->n_bodies; ++b) {
            int cb = s->crowder[b];
            for (int i = s->body_start[a]; i < s->body_start[a + 1]; ++i) {
                for (int j = s->body_start[b]; j < s->body_start[b + 1]; ++j) {
                    double lj, dh;
                    orc_pair_ij(s, ff, x[i], y[i], z[i], i, ca, x[j], y[j], z[j], j, cb, &lj, &dh);
                    ljacc += lj;
                    dhacc += dh;
                }
            }
        }
    }
    if (lj_sum) *lj_sum = ljacc;
    if (dh_sum) *dh_sum = dhacc;
    return (ljacc * ff->lj_cf + dhacc * ff->dh_cf) * ff->kbt_cf;
}

double orc_energy_bodies(const orc_system *s, const orc_forcefield *ff,
                         const float *x, const float *y, const float *z, int a, int b)
{
    double ljacc = 0.0, dhacc = 0.0;
    int ca = s->crowder[a], cb = s->crowder[b];
    for (int i = s->body_start[a]; i < s->body_start[a + 1]; ++i)
        for (int j = s->body_start[b]; j < s->body_start[b + 1]; ++j) {
            double lj, dh;
            orc_pair_ij(s, ff, x[i], y[i], z[i], i, ca, x[j], y[j], z[j], j, cb, &lj, &dh);
            ljacc += lj;
            dhacc += dh;
        }
    return (ljacc * ff->lj_cf + dhacc * ff->dh_cf) * ff->kbt_cf;
}

double orc_energy_noncrowder(const orc_system *s, const orc_forcefield *ff,
                             const float *x, const float *y, const float *z)
{
    double e = 0.0;
    for (int a = 0; a < s->n_bodies; ++a) {
        if (s->crowder[a]) continue;
        for (int b = a + 1; b < s->n_bodies; ++b) {
            if (s->crowder[b]) continue;
            e += orc_energy_bodies(s, ff, x, y, z, a, b);
        }
    }
    return e;
}

void orc_pair_table(const orc_system *s, const orc_forcefield *ff,
                    const float *x, const float *y, const float *z, int a, int b, double *out)
{
    int ca = s->crowder[a], cb = s->crowder[b];
    int nb = s->body_start[b + 1] - s->body_start[b];
    for (int i = s->body_start[a]; i < s->body_start[a + 1]; ++i)
        for (int j = s->body_start[b]; j < s->body_start[b + 1]; ++j) {
            double lj, dh;
            double e = orc_pair_ij(s, ff, x[i], y[i], z[i], i, ca, x[j], y[j], z[j], j, cb, &lj, &dh);
            size_t k = (size_t)(i - s->body_start[a]) * nb + (j - s->body_start[b]);
            out[3 * k + 0] = e;
            out[3 * k + 1] = dh * ff->dh_cf * ff->kbt_cf;
            out[3 * k + 2] = lj * ff->lj_cf * ff->kbt_cf;
        }
}

double orc_delta_energy(const orc_system *s, const orc_forcefield *ff,
                        const float *x, const float *y, const float *z, int m,
                        const float *nx, const float *ny, const float *nz)
{
    double acc = 0.0;
    int m0 = s->body_start[m], m1 = s->body_start[m + 1];
    int cm = s->crowder[m];
    for (int b = 0; b < s->n_bodies; ++b) {
        if (b == m) continue;
        int cb = s->crowder[b];
        for (int i = m0; i < m1; ++i) {
            int k = i - m0;
            for (int j = s->body_start[b]; j < s->body_start[b + 1]; ++j) {
                double e_new = orc_pair_ij(s, ff, nx[k], ny[k], nz[k], i, cm, x[j], y[j], z[j], j, cb, NULL, NULL);
                double e_old = orc_pair_ij(s, ff, x[i], y[i], z[i], i, cm, x[j], y[j], z[j], j, cb, NULL, NULL);
                acc += e_new - e_old;
            }
        }
    }
    return acc;
}

/* ------------------------------------------------------------ adjoint ---- */
double orc_energy_full_d(const orc_system *s, const orc_forcefield *ff,
                         const double *x, const double *y, const double *z)
{
    double e = 0.0;
    for (int a = 0; a < s->n_bodies; ++a)
        for (int b = a + 1; b < s->n_bodies; ++b)
            for (int i = s->body_start[a]; i < s->body_start[a + 1]; ++i)
                for (int j = s->body_start[b]; j < s->body_start[b + 1]; ++j)
                    e += orc_pair_ij(s, ff, x[i], y[i], z[i], i, s->crowder[a], x[j], y[j], z[j], j,
                                     s->crowder[b], NULL, NULL);
    return e;
}

/* Pair energy derivatives at distance r (the same branches as
 * orc_pair_energy): de/dr, de/dE_ij and de/dxi, all in kcal/mol units. */
static void orc_pair_derivs(const orc_forcefield *ff, double r, int ti, double qi, double si, int ci,
                            int tj, double qj, double sj, int cj, double *dedr, double *de_deij, double *de_dxi)
{
    const double c_lj = ff->lj_cf * ff->kbt_cf, c_dh = ff->dh_cf * ff->kbt_cf;
    double dlj = 0.0, ddh = 0.0, dlj_de = 0.0, ddh_dxi = 0.0;
    if (ci || cj) {
        if (r < ff->crowder_cutoff) {
            double t = ff->crowder_sigma / r, t2 = t * t, t6 = t2 * t2 * t2;
            dlj = -6.0 * t6 / r;
        }
    } else {
        double dh = qi * qj * exp(-r / ff->xi) / r;
        ddh = -dh * (1.0 / ff->xi + 1.0 / r);
        ddh_dxi = dh * r / (ff->xi * ff->xi);
        double eij = ff->lambda * (ff->table[ti * ORC_NTYPES + tj] - ff->e0);
        double sij = 0.5 * (si + sj);
        double t = sij / r, t2 = t * t, x6 = t2 * t2 * t2;
        double base = 4.0 * x6 * (x6 - 1.0);
        double dbase = 4.0 * (2.0 * x6 - 1.0) * (-6.0 * x6 / r);
        if (eij > 0.0 && r < sij * 1.122462048309372981433533049679) {
            dlj = eij * dbase;          /* LJ = E (base + 2) */
            dlj_de = base + 2.0;
        } else {
            dlj = -eij * dbase;         /* LJ = -E base */
            dlj_de = -base;
        }
    }
    *dedr = dlj * c_lj + ddh * c_dh;
    *de_deij = dlj_de * c_lj;
    *de_dxi = ddh_dxi * c_dh;
}

void orc_gradient(const orc_system *s, const orc_forcefield *ff,
                  const double *x, const double *y, const double *z,
                  const double *rx, const double *ry, const double *rz,
                  double *bead_grad, double *force_torque, double *dparams)
{
    const int n = s->n_beads;
    double *g = (double *)calloc((size_t)3 * n, sizeof(double));
    double dl = 0.0, de0 = 0.0, dxi = 0.0;
    for (int a = 0; a < s->n_bodies; ++a)
        for (int b = a + 1; b < s->n_bodies; ++b)
            for (int i = s->body_start[a]; i < s->body_start[a + 1]; ++i)
                for (int j = s->body_start[b]; j < s->body_start[b + 1]; ++j) {
                    double dx = orc_min_image(x[i] - x[j], s->box);
                    double dy = orc_min_image(y[i] - y[j], s->box);
                    double dz = orc_min_image(z[i] - z[j], s->box);
                    double d = sqrt(dx * dx + dy * dy + dz * dz);
                    double r = d + ff->eps_r;
                    double dedr, de_deij, de_dxi;
                    orc_pair_derivs(ff, r, s->type[i], s->charge[i], s->sigma[i], s->crowder[a], s->type[j],
                                    s->charge[j], s->sigma[j], s->crowder[b], &dedr, &de_deij, &de_dxi);
                    double k = d > 0.0 ? dedr / d : 0.0;
                    g[3 * i + 0] += k * dx; g[3 * i + 1] += k * dy; g[3 * i + 2] += k * dz;
                    g[3 * j + 0] -= k * dx; g[3 * j + 1] -= k * dy; g[3 * j + 2] -= k * dz;
                    if (!(s->crowder[a] || s->crowder[b])) {
                        dl += de_deij * (ff->table[s->type[i] * ORC_NTYPES + s->type[j]] - ff->e0);
                        de0 += de_deij * (-ff->lambda);
                        dxi += de_dxi;
                    }
                }
    if (bead_grad) memcpy(bead_grad, g, sizeof(double) * 3 * (size_t)n);
    if (force_torque)
        for (int m = 0; m < s->n_bodies; ++m) {
            double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
            for (int i = s->body_start[m]; i < s->body_start[m + 1]; ++i) {
                double f[3] = {-g[3 * i], -g[3 * i + 1], -g[3 * i + 2]};
                F[0] += f[0]; F[1] += f[1]; F[2] += f[2];
                T[0] += ry[i] * f[2] - rz[i] * f[1];
                T[1] += rz[i] * f[0] - rx[i] * f[2];
                T[2] += rx[i] * f[1] - ry[i] * f[0];
            }
            for (int k = 0; k < 3; ++k) {
                force_torque[6 * m + k] = F[k];
                force_torque[6 * m + 3 + k] = T[k];
            }
        }
    if (dparams) { dparams[0] = dl; dparams[1] = de0; dparams[2] = dxi; }
    free(g);
}

/* -------------------------------------------------------------- moves ---- */
/* Quaternion::multiply (Quaternion.cpp:35-47): out = a * b, (w,x,y,z). */
void orc_quat_mul(const double a[4], const double b[4], double out[4])
{
    double w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    double x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    double y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    double z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    out[0] = w; out[1] = x; out[2] = y; out[3] = z;
}

/* Quaternion::rotateVector (Quaternion.cpp:84-101): FP64 arithmetic on a
 * float vector, result rounded to float. */
void orc_quat_rotate(const double q[4], const float v[3], float out[3])
{
    double w = q[0], x = q[1], y = q[2], z = q[3];
    double w2 = w * w, x2 = x * x, y2 = y * y, z2 = z * z;
    double wx = w * x, wy = w * y, wz = w * z, xy = x * y, xz = x * z, yz = y * z;
    double vx = v[0], vy = v[1], vz = v[2];
    out[0] = (float)(vx * (w2 + x2 - y2 - z2) + (vy * (xy - wz) + vz * (wy + xz)) * 2.0);
    out[1] = (float)((vx * (wz + xy) + vz * (yz - wx)) * 2.0 + vy * (w2 - x2 + y2 - z2));
    out[2] = (float)((vx * (xz - wy) + vy * (yz + wx)) * 2.0 + vz * (w2 - x2 - y2 + z2));
}

void orc_propose(const orc_system *s, const orc_replica *r, int m, int rotate,
                 double a, double b, double c,
                 float *nx, float *ny, float *nz, float *nrx, float *nry, float *nrz,
                 float cnew[3], double qnew[4])
{
    int m0 = s->body_start[m], n = s->body_start[m + 1] - m0;
    float cx = r->cx[m], cy = r->cy[m], cz = r->cz[m];
    double qold[4] = {r->qw[m], r->qx[m], r->qy[m], r->qz[m]};
    if (rotate) {
        /* Replica::createNormalisedRandomVectord + Vector3double::normalizeInPlace
         * (Replica.cpp:312-318, vector3f.cpp:36-43) */
        double mag = sqrt(a * a + b * b + c * c);
        double ax = a / mag, ay = b / mag, az = c / mag;
        /* Molecule::rotateQ (Molecule.cpp:288-333) */
        double q[4] = {r->rot_cos_half, r->rot_sin_half * ax, r->rot_sin_half * ay, r->rot_sin_half * az};
        double qm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        q[0] = q[0] / qm; q[1] = q[1] / qm; q[2] = q[2] / qm; q[3] = q[3] / qm;
        orc_quat_mul(q, qold, qnew);
        for (int k = 0; k < n; ++k) {
            float v[3] = {r->rx[m0 + k], r->ry[m0 + k], r->rz[m0 + k]}, o[3];
            orc_quat_rotate(q, v, o);
            nrx[k] = o[0]; nry[k] = o[1]; nrz[k] = o[2];
            nx[k] = o[0] + cx; ny[k] = o[1] + cy; nz[k] = o[2] + cz;
        }
        cnew[0] = cx; cnew[1] = cy; cnew[2] = cz;
    } else {
        /* Replica::createNormalisedRandomVector + Vector3f::normalize
         * (Replica.cpp:303-310, vector3f.cpp:58-62) in float */
        float vx = (float)a, vy = (float)b, vz = (float)c;
        float mag = sqrtf(vx * vx + vy * vy + vz * vz);
        if (mag != 0.0f) { vx = vx / mag; vy = vy / mag; vz = vz / mag; }
        float st = r->translate_step;
        float tx = vx * st, ty = vy * st, tz = vz * st;
        /* Replica::translate periodic branch (Replica.cpp:358-362) */
        float L = (float)s->box;
        float px = cx + tx, py = cy + ty, pz = cz + tz;
        if (L > 0.0f) {
            px = orc_fmodf(px + L, L);
            py = orc_fmodf(py + L, L);
            pz = orc_fmodf(pz + L, L);
        }
        cnew[0] = px; cnew[1] = py; cnew[2] = pz;
        /* Molecule::setPosition (Molecule.cpp:247-260) */
        for (int k = 0; k < n; ++k) {
            nrx[k] = r->rx[m0 + k]; nry[k] = r->ry[m0 + k]; nrz[k] = r->rz[m0 + k];
            nx[k] = nrx[k] + px; ny[k] = nry[k] + py; nz[k] = nrz[k] + pz;
        }
        memcpy(qnew, qold, sizeof qold);
    }
}

/* ----------------------------------------------------------------- MC ---- */
int orc_metropolis(double delta, double temperature, double u, double *margin)
{
    double beta = 4184.0 / (ORC_R_GAS * temperature);
    if (delta < 0.0) {
        if (margin) *margin = -delta * beta;
        return 1;
    }
    double p = orc_exp(-delta * beta);
    /* two thresholds: dE = 0 (accept kind 1 or 2) and -ln u (accept or reject) */
    if (margin) *margin = fmin(delta * beta, fabs(log(u) + delta * beta));
    return u < p ? 2 : 0;
}

void orc_mc_steps(const orc_system *s, const orc_forcefield *ff, orc_replica *r,
                  const uint32_t key[2], int nsteps, int mode, orc_step_record *rec)
{
    int maxn = 0;
    for (int b = 0; b < s->n_bodies; ++b) {
        int n = s->body_start[b + 1] - s->body_start[b];
        if (n > maxn) maxn = n;
    }
    float *buf = (float *)malloc(sizeof(float) * 6 * (size_t)(maxn > 0 ? maxn : 1));
    float *nx = buf, *ny = buf + maxn, *nz = buf + 2 * maxn;
    float *nrx = buf + 3 * maxn, *nry = buf + 4 * maxn, *nrz = buf + 5 * maxn;
    for (int st = 0; st < nsteps; ++st) {
        uint32_t w[6];
        orc_step_words(r->step, r->rng_id, key, w);
        int m = (int)(((uint64_t)w[0] * (uint64_t)s->n_bodies) >> 32);
        int rotate = orc_u01(w[1]) < 0.5;
        double a = orc_u01(w[2]) - 0.5, b = 0.5 - orc_u01(w[3]), c = orc_u01(w[4]) - 0.5;
        double u = orc_u01(w[5]);
        float cnew[3];
        double qnew[4];
        orc_propose(s, r, m, rotate, a, b, c, nx, ny, nz, nrx, nry, nrz, cnew, qnew);
        int m0 = s->body_start[m], n = s->body_start[m + 1] - m0;
        double delta, e_new = 0.0;
        if (mode == 0) {
            delta = orc_delta_energy(s, ff, r->x, r->y, r->z, m, nx, ny, nz);
        } else {
            /* full recomputation with the moved body swapped in (Replica.cpp:435) */
            float *ox = (float *)malloc(sizeof(float) * 3 * (size_t)n);
            memcpy(ox, r->x + m0, sizeof(float) * n);
            memcpy(ox + n, r->y + m0, sizeof(float) * n);
            memcpy(ox + 2 * n, r->z + m0, sizeof(float) * n);
            memcpy(r->x + m0, nx, sizeof(float) * n);
            memcpy(r->y + m0, ny, sizeof(float) * n);
            memcpy(r->z + m0, nz, sizeof(float) * n);
            e_new = orc_energy_full(s, ff, r->x, r->y, r->z, NULL, NULL);
            memcpy(r->x + m0, ox, sizeof(float) * n);
            memcpy(r->y + m0, ox + n, sizeof(float) * n);
            memcpy(r->z + m0, ox + 2 * n, sizeof(float) * n);
            free(ox);
            delta = e_new - r->energy;
        }
        double margin;
        int acc = orc_metropolis(delta, r->temperature, u, &margin);
        if (acc) {
            memcpy(r->x + m0, nx, sizeof(float) * n);
            memcpy(r->y + m0, ny, sizeof(float) * n);
            memcpy(r->z + m0, nz, sizeof(float) * n);
            memcpy(r->rx + m0, nrx, sizeof(float) * n);
            memcpy(r->ry + m0, nry, sizeof(float) * n);
            memcpy(r->rz + m0, nrz, sizeof(float) * n);
            r->cx[m] = cnew[0]; r->cy[m] = cnew[1]; r->cz[m] = cnew[2];
            r->qw[m] = qnew[0]; r->qx[m] = qnew[1]; r->qy[m] = qnew[2]; r->qz[m] = qnew[3];
            r->energy = mode == 0 ? r->energy + delta : e_new;
            if (acc == 1) r->n_accept++; else r->n_accept_boltz++;
        } else {
            r->n_reject++;
        }
        if (rec) {
            rec[st].body = m;
            rec[st].move = rotate;
            rec[st].accepted = acc;
            rec[st].pad = 0;
            rec[st].delta = delta;
            rec[st].margin = margin;
        }
        r->step++;
    }
    free(buf);
}

int orc_exchange_round(int n_rep, const double *ladder, int *slot_replica,
                       const double *energy, int offset, uint64_t round, uint32_t system_id,
                       const uint32_t key[2], int *swapped)
{
    int n = 0;
    if (swapped)
        for (int k = 0; k < n_rep; ++k) swapped[k] = 0;
    for (int k = offset; k < n_rep - 1; k += 2) {
        int ri = slot_replica[k], rj = slot_replica[k + 1];
        double delta = (1.0 / ladder[k] - 1.0 / ladder[k + 1]) * (energy[ri] - energy[rj]) * (4184.0 / ORC_R_GAS);
        double u = orc_exchange_uniform(round, (uint32_t)k, system_id, key);
        int acc = delta >= 0.0 || u < orc_exp(delta);
        if (acc) {
            slot_replica[k] = rj;
            slot_replica[k + 1] = ri;
            if (swapped) swapped[k] = 1;
            ++n;
        }
    }
    return n;
}

/* ------------------------------------------------ threaded CPU baseline ---- */
typedef struct orc_thread_arg {
    const orc_system *s;
    const orc_forcefield *ff;
    orc_replica *reps;
    int first, count, nsteps, mode;
    const uint32_t *key;
    orc_step_record *rec;
} orc_thread_arg;

static void *orc_thread_main(void *p)
{
    orc_thread_arg *a = (orc_thread_arg *)p;
    for (int i = 0; i < a->count; ++i)
        orc_mc_steps(a->s, a->ff, &a->reps[a->first + i], a->key, a->nsteps, a->mode,
                     a->rec ? a->rec + (size_t)(a->first + i) * a->nsteps : NULL);
    return NULL;
}

double orc_mc_threaded(const orc_system *s, const orc_forcefield *ff, orc_replica *reps,
                       int n_rep, const uint32_t key[2], int nsteps, int mode, int threads,
                       orc_step_record *rec)
{
    if (threads < 1) threads = 1;
    if (threads > n_rep) threads = n_rep;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    orc_thread_arg *args = (orc_thread_arg *)malloc(sizeof(orc_thread_arg) * threads);
    /* contiguous replica blocks per thread (main.cpp:879-889) */
    int per = (n_rep + threads - 1) / threads;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int launched = 0;
    for (int t = 0; t < threads; ++t) {
        int first = t * per;
        int count = n_rep - first < per ? n_rep - first : per;
        if (count <= 0) break;
        args[t] = (orc_thread_arg){s, ff, reps, first, count, nsteps, mode, key, rec};
        pthread_create(&th[t], NULL, orc_thread_main, &args[t]);
        ++launched;
    }
    for (int t = 0; t < launched; ++t) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(th);
    free(args);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------- force-biased MC ---- */
#define ORC_TAG_FB 0x46420000u

/* Force (3) and torque about the centre (3) on body m whose beads sit at
 * (bx,by,bz) with relative positions (brx,bry,brz), from every bead of the
 * other bodies at (x,y,z).  Pair derivatives as orc_gradient. */
void orc_body_force(const orc_system *s, const orc_forcefield *ff, const float *x, const float *y, const float *z,
                    int m, const float *bx, const float *by, const float *bz, const float *brx, const float *bry,
                    const float *brz, double ft[6])
{
    int m0 = s->body_start[m], m1 = s->body_start[m + 1];
    for (int k = 0; k < 6; ++k) ft[k] = 0.0;
    for (int i = m0; i < m1; ++i) {
        int k = i - m0;
        double f[3] = {0.0, 0.0, 0.0};
        for (int b = 0; b < s->n_bodies; ++b) {
            if (b == m) continue;
            for (int j = s->body_start[b]; j < s->body_start[b + 1]; ++j) {
                double dx = orc_min_image((double)bx[k] - (double)x[j], s->box);
                double dy = orc_min_image((double)by[k] - (double)y[j], s->box);
                double dz = orc_min_image((double)bz[k] - (double)z[j], s->box);
                double d = sqrt(dx * dx + dy * dy + dz * dz);
                double r = d + ff->eps_r;
                double dedr, de_deij, de_dxi;
                orc_pair_derivs(ff, r, s->type[i], s->charge[i], s->sigma[i], s->crowder[m], s->type[j],
                                s->charge[j], s->sigma[j], s->crowder[b], &dedr, &de_deij, &de_dxi);
                double g = d > 0.0 ? dedr / d : 0.0;
                f[0] -= g * dx; f[1] -= g * dy; f[2] -= g * dz;
            }
        }
        ft[0] += f[0]; ft[1] += f[1]; ft[2] += f[2];
        double ax = brx[k], ay = bry[k], az = brz[k];
        ft[3] += ay * f[2] - az * f[1];
        ft[4] += az * f[0] - ax * f[2];
        ft[5] += ax * f[1] - ay * f[0];
    }
}

void orc_mc_fb_steps(const orc_system *s, const orc_forcefield *ff, orc_replica *r, const uint32_t key[2],
                     int nsteps, double a_t, double a_r, int hastings, orc_step_record *rec)
{
    int maxn = 1;
    for (int b = 0; b < s->n_bodies; ++b) {
        int n = s->body_start[b + 1] - s->body_start[b];
        if (n > maxn) maxn = n;
    }
    float *buf = (float *)malloc(sizeof(float) * 6 * (size_t)maxn);
    float *nx = buf, *ny = buf + maxn, *nz = buf + 2 * maxn;
    float *nrx = buf + 3 * maxn, *nry = buf + 4 * maxn, *nrz = buf + 5 * maxn;
    const double two_pi = 6.283185307179586476925286766559;
    for (int st = 0; st < nsteps; ++st) {
        uint32_t ctr[4] = {(uint32_t)r->step, (uint32_t)(r->step >> 32), r->rng_id, ORC_TAG_FB};
        uint32_t w0[4], w1[4];
        orc_philox4x32_10(ctr, key, w0);
        ctr[3] = ORC_TAG_FB + 1u;
        orc_philox4x32_10(ctr, key, w1);
        int m = (int)(((uint64_t)w0[0] * (uint64_t)s->n_bodies) >> 32);
        int rotate = orc_u01(w0[1]) < 0.5;
        double u1 = orc_u01(w0[2]), u2 = orc_u01(w0[3]), u3 = orc_u01(w1[0]), u4 = orc_u01(w1[1]);
        double u = orc_u01(w1[2]);
        /* Box-Muller on (1 - u1, u2) and (1 - u3, u4): three standard normals */
        double rho1 = sqrt(-2.0 * log(1.0 - u1)), rho2 = sqrt(-2.0 * log(1.0 - u3));
        double nrm[3] = {rho1 * cos(two_pi * u2), rho1 * sin(two_pi * u2), rho2 * cos(two_pi * u4)};
        const double beta = 4184.0 / (ORC_R_GAS * r->temperature);
        int m0 = s->body_start[m], n = s->body_start[m + 1] - m0;
        double ft_old[6], ft_new[6];
        orc_body_force(s, ff, r->x, r->y, r->z, m, r->x + m0, r->y + m0, r->z + m0, r->rx + m0, r->ry + m0,
                       r->rz + m0, ft_old);
        float cnew[3] = {r->cx[m], r->cy[m], r->cz[m]};
        double qold[4] = {r->qw[m], r->qx[m], r->qy[m], r->qz[m]}, qnew[4];
        double disp[3];
        const double A = rotate ? a_r : a_t;
        const double *fo = rotate ? ft_old + 3 : ft_old;
        for (int k = 0; k < 3; ++k) disp[k] = beta * A * fo[k] + sqrt(2.0 * A) * nrm[k];
        if (rotate) {
            double ang = sqrt(disp[0] * disp[0] + disp[1] * disp[1] + disp[2] * disp[2]);
            double q[4] = {1.0, 0.0, 0.0, 0.0};
            if (ang > 0.0) {
                double sh = sin(0.5 * ang), ch = cos(0.5 * ang);
                q[0] = ch; q[1] = sh * disp[0] / ang; q[2] = sh * disp[1] / ang; q[3] = sh * disp[2] / ang;
                double qm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
                q[0] = q[0] / qm; q[1] = q[1] / qm; q[2] = q[2] / qm; q[3] = q[3] / qm;
            }
            orc_quat_mul(q, qold, qnew);
            for (int k = 0; k < n; ++k) {
                float v[3] = {r->rx[m0 + k], r->ry[m0 + k], r->rz[m0 + k]}, o[3];
                orc_quat_rotate(q, v, o);
                nrx[k] = o[0]; nry[k] = o[1]; nrz[k] = o[2];
                nx[k] = o[0] + cnew[0]; ny[k] = o[1] + cnew[1]; nz[k] = o[2] + cnew[2];
            }
        } else {
            float t[3] = {(float)disp[0], (float)disp[1], (float)disp[2]};
            for (int k = 0; k < 3; ++k) disp[k] = (double)t[k]; /* the displacement actually made */
            float L = (float)s->box;
            float px = cnew[0] + t[0], py = cnew[1] + t[1], pz = cnew[2] + t[2];
            if (L > 0.0f) {
                px = orc_fmodf(px + L, L);
                py = orc_fmodf(py + L, L);
                pz = orc_fmodf(pz + L, L);
            }
            cnew[0] = px; cnew[1] = py; cnew[2] = pz;
            for (int k = 0; k < n; ++k) {
                nrx[k] = r->rx[m0 + k]; nry[k] = r->ry[m0 + k]; nrz[k] = r->rz[m0 + k];
                nx[k] = nrx[k] + px; ny[k] = nry[k] + py; nz[k] = nrz[k] + pz;
            }
            memcpy(qnew, qold, sizeof qold);
        }
        double delta = orc_delta_energy(s, ff, r->x, r->y, r->z, m, nx, ny, nz);
        orc_body_force(s, ff, r->x, r->y, r->z, m, nx, ny, nz, nrx, nry, nrz, ft_new);
        const double *fn = rotate ? ft_new + 3 : ft_new;
        /* log q(new->old) - log q(old->new): reverse displacement -disp */
        double fwd = 0.0, rev = 0.0;
        for (int k = 0; k < 3; ++k) {
            double a = disp[k] - beta * A * fo[k], b = -disp[k] - beta * A * fn[k];
            fwd += a * a;
            rev += b * b;
        }
        /* accept if x < -ln u; hastings = 0 drops the proposal ratio (a planted
         * mistake for the tests: the chain then leaves exp(-beta E)) */
        double x = beta * delta + (hastings ? (rev - fwd) / (4.0 * A) : 0.0);
        double Lu = u > 0.0 ? -log(u) : 745.1332191019412;
        int acc = x < 0.0 ? 1 : (x < Lu ? 2 : 0);
        if (acc) {
            memcpy(r->x + m0, nx, sizeof(float) * n);
            memcpy(r->y + m0, ny, sizeof(float) * n);
            memcpy(r->z + m0, nz, sizeof(float) * n);
            memcpy(r->rx + m0, nrx, sizeof(float) * n);
            memcpy(r->ry + m0, nry, sizeof(float) * n);
            memcpy(r->rz + m0, nrz, sizeof(float) * n);
            r->cx[m] = cnew[0]; r->cy[m] = cnew[1]; r->cz[m] = cnew[2];
            r->qw[m] = qnew[0]; r->qx[m] = qnew[1]; r->qy[m] = qnew[2]; r->qz[m] = qnew[3];
            r->energy += delta;
            if (acc == 1) r->n_accept++; else r->n_accept_boltz++;
        } else {
            r->n_reject++;
        }
        if (rec) {
            rec[st].body = m;
            rec[st].move = rotate;
            rec[st].accepted = acc;
            rec[st].pad = 0;
            rec[st].delta = delta;
            rec[st].margin = x < 0.0 ? -x : fmin(x, fabs(x - Lu));
        }
        r->step++;
    }
    free(buf);
}
