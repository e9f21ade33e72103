/*
 * oracle.c -- CPU restatement of the reference's Hermite-4 force path.
 *
 * TEST INFRASTRUCTURE ONLY.  The parity tests compare against it: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it.  The product
 * (n-body-code_amd/) never links, loads or falls back to it.
 *
 * Every function restates one piece of jlalmerol/N-Body-Code and cites the
 * file:line it follows.  Built with gcc -O3 -ffp-contract=off (no FMA
 * contraction, no -march) so that its arithmetic matches the reference's
 * x86-64 build operation for operation; bit-exactness against the compiled
 * reference (oracle/_ref) is checked in tests/test_oracle.py.
 *
 * Pinning (how each function is checked beyond retyping):
 *   or_invsqrt2d        bitwise vs _ref InvSqrt2D, and <=5e-16 vs 1/sqrt
 *   or_force_f64        bitwise vs _ref force_calculation (FP64 build) on
 *                       golden fixtures; closed-form 2-body a and jerk;
 *                       momentum (sum m a = 0); long-double brute force
 *   or_force_f32        bitwise vs _ref force_calculation (_F32 build)
 *   or_predict /
 *   or_snap_correct     bitwise vs _ref NBody::timestepping trajectories;
 *                       constant-jerk motion integrated exactly
 *   or_updatedt /
 *   or_initialdt        bitwise vs _ref adaptive-dt trajectories
 *   or_kinetic /
 *   or_potential_compat bitwise vs _ref compute_kinetic/compute_potential
 *   or_potential        closed form on 2 and 3 bodies
 *   or_reference_ic     bitwise vs _ref NBody::initialize (mt19937(42))
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Fast inverse square root, src/utils.cpp:155-165.                    */
/* Published magic-constant method with four modified Newton steps     */
/* (Walczyk, Moroz, Cieslinski, 2021); constants must equal the        */
/* reference's for bit-exactness.  Bit casts go through memcpy.        */
/* ------------------------------------------------------------------ */
double or_invsqrt2d(double x)
{
    double half = 0.5 * x;
    int64_t bits;
    memcpy(&bits, &x, sizeof bits);
    bits = 0x5FE6ED2102DCBFDALL - (bits >> 1);
    double y;
    memcpy(&y, &bits, sizeof y);
    y *= 1.50087895511633457 - half * y * y;
    y *= 1.50000057967625766 - half * y * y;
    y *= 1.5000000000002520 - half * y * y;
    y *= 1.5000000000000000 - half * y * y;
    return y;
}

/* ------------------------------------------------------------------ */
/* FP64 pair sum, src/kernel.cpp:1164-1224.                           */
/* Newton-3rd-law triangle j<i, i cyclic over threads (:1185), each    */
/* thread accumulating into private arrays which are merged at the end */
/* (the OpenMP array reduction of :1180).  With nthreads == 1 this is  */
/* the reference's single-thread order exactly.  The merge here is in  */
/* thread order (deterministic); libgomp's is not ordered, so only     */
/* nthreads == 1 is compared bitwise.                                  */
/* ------------------------------------------------------------------ */
static void f64_rows(int n, int first, int stride,
                     const double *x, const double *y, const double *z,
                     const double *w, const double *vx, const double *vy,
                     const double *vz, double eps2,
                     double *ax, double *ay, double *az,
                     double *jx, double *jy, double *jz)
{
    for (int i = first; i < n; i += stride) {
        for (int j = 0; j < i; j++) {
            double dx = x[j] - x[i];
            double dy = y[j] - y[i];
            double dz = z[j] - z[i];
            double ux = vx[j] - vx[i];
            double uy = vy[j] - vy[i];
            double uz = vz[j] - vz[i];
            double s = dx * dx + dy * dy + dz * dz + eps2;
            double vr = dx * ux + dy * uy + dz * uz;
            double inv = or_invsqrt2d(s);
            double inv2 = inv * inv;
            double q = inv * inv2;
            double qq = -3.0 * vr * inv2;
            double ti = w[i] * q;
            double tj = w[j] * q;
            ax[i] += dx * tj;
            ay[i] += dy * tj;
            az[i] += dz * tj;
            ax[j] -= dx * ti;
            ay[j] -= dy * ti;
            az[j] -= dz * ti;
            jx[i] += tj * (ux + qq * dx);
            jy[i] += tj * (uy + qq * dy);
            jz[i] += tj * (uz + qq * dz);
            jx[j] -= ti * (ux + qq * dx);
            jy[j] -= ti * (uy + qq * dy);
            jz[j] -= ti * (uz + qq * dz);
        }
    }
}

/* ------------------------------------------------------------------ */
/* _F32 pair sum, src/kernel.cpp:1101-1163.  Rounding points follow    */
/* C/C++ promotion exactly: differences in double rounded to float;    */
/* s summed in float then "+ eps2" in double and rounded; the inverse  */
/* sqrt in double rounded to float; "-3.0 * vr * inv2" in double then  */
/* rounded; "w * q" in double then rounded; accumulation of float     */
/* products into the double outputs.                                   */
/* ------------------------------------------------------------------ */
static void f32_rows(int n, int first, int stride,
                     const double *x, const double *y, const double *z,
                     const double *w, const double *vx, const double *vy,
                     const double *vz, double eps2,
                     double *ax, double *ay, double *az,
                     double *jx, double *jy, double *jz)
{
    for (int i = first; i < n; i += stride) {
        for (int j = 0; j < i; j++) {
            float dx = (float)(x[j] - x[i]);
            float dy = (float)(y[j] - y[i]);
            float dz = (float)(z[j] - z[i]);
            float ux = (float)(vx[j] - vx[i]);
            float uy = (float)(vy[j] - vy[i]);
            float uz = (float)(vz[j] - vz[i]);
            float s = (float)((double)(dx * dx + dy * dy + dz * dz) + eps2);
            float vr = dx * ux + dy * uy + dz * uz;
            float inv = (float)or_invsqrt2d((double)s);
            float inv2 = inv * inv;
            float q = inv2 * inv;
            float qq = (float)(-3.0 * (double)vr * (double)inv2);
            float ti = (float)(w[i] * (double)q);
            float tj = (float)(w[j] * (double)q);
            ax[i] += dx * tj;
            ay[i] += dy * tj;
            az[i] += dz * tj;
            ax[j] -= dx * ti;
            ay[j] -= dy * ti;
            az[j] -= dz * ti;
            jx[i] += tj * (ux + qq * dx);
            jy[i] += tj * (uy + qq * dy);
            jz[i] += tj * (uz + qq * dz);
            jx[j] -= ti * (ux + qq * dx);
            jy[j] -= ti * (uy + qq * dy);
            jz[j] -= ti * (uz + qq * dz);
        }
    }
}

typedef void (*rows_fn)(int, int, int, const double *, const double *,
                        const double *, const double *, const double *,
                        const double *, const double *, double, double *,
                        double *, double *, double *, double *, double *);

static int run_force(rows_fn rows, int n, const double *x, const double *y,
                     const double *z, const double *w, const double *vx,
                     const double *vy, const double *vz, double eps2,
                     double *ax, double *ay, double *az, double *jx,
                     double *jy, double *jz, int nthreads)
{
    /* outputs zeroed first, src/kernel.cpp:1174-1178 */
    for (int i = 0; i < n; i++) {
        ax[i] = 0.0; ay[i] = 0.0; az[i] = 0.0;
        jx[i] = 0.0; jy[i] = 0.0; jz[i] = 0.0;
    }
    if (nthreads <= 1) {
        /* One private copy starting from zero; merging it into the zeroed
         * outputs is an exact "0 + p", so accumulate in place. */
        rows(n, 0, 1, x, y, z, w, vx, vy, vz, eps2, ax, ay, az, jx, jy, jz);
        return 0;
    }
    size_t len = (size_t)n * 6;
    double *priv = (double *)calloc(len * (size_t)nthreads, sizeof(double));
    if (!priv)
        return -1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
        int nt = omp_get_num_threads();
#else
        int t = 0, nt = 1;
#endif
        if (nt == nthreads) {
            double *p = priv + len * (size_t)t;
            rows(n, t, nthreads, x, y, z, w, vx, vy, vz, eps2, p, p + n,
                 p + 2 * (size_t)n, p + 3 * (size_t)n, p + 4 * (size_t)n,
                 p + 5 * (size_t)n);
        }
    }
    double *outs[6] = {ax, ay, az, jx, jy, jz};
    for (int t = 0; t < nthreads; t++) {
        const double *p = priv + len * (size_t)t;
        for (int c = 0; c < 6; c++)
            for (int i = 0; i < n; i++)
                outs[c][i] += p[(size_t)c * n + i];
    }
    free(priv);
    return 0;
}

int or_force_f64(int n, const double *x, const double *y, const double *z,
                 const double *w, const double *vx, const double *vy,
                 const double *vz, double eps2, double *ax, double *ay,
                 double *az, double *jx, double *jy, double *jz, int nthreads)
{
    return run_force(f64_rows, n, x, y, z, w, vx, vy, vz, eps2, ax, ay, az,
                     jx, jy, jz, nthreads);
}

int or_force_f32(int n, const double *x, const double *y, const double *z,
                 const double *w, const double *vx, const double *vy,
                 const double *vz, double eps2, double *ax, double *ay,
                 double *az, double *jx, double *jy, double *jz, int nthreads)
{
    return run_force(f32_rows, n, x, y, z, w, vx, vy, vz, eps2, ax, ay, az,
                     jx, jy, jz, nthreads);
}

/* ------------------------------------------------------------------ */
/* Directed FP64 sum for a subset of target bodies: acc/jerk of body   */
/* idx[k] from all j != idx[k], same per-pair arithmetic as            */
/* src/kernel.cpp:1188-1205 with 1/sqrt in long double.  Used to check */
/* big-N GPU results on a sample without the O(N^2) triangle.          */
/* ------------------------------------------------------------------ */
int or_force_sample_ld(int n, const double *x, const double *y,
                       const double *z, const double *w, const double *vx,
                       const double *vy, const double *vz, double eps2,
                       int nidx, const int *idx, double *out6, int nthreads)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int k = 0; k < nidx; k++) {
        int i = idx[k];
        long double a[3] = {0, 0, 0}, jk[3] = {0, 0, 0};
        for (int j = 0; j < n; j++) {
            if (j == i)
                continue;
            long double dx = (long double)x[j] - x[i];
            long double dy = (long double)y[j] - y[i];
            long double dz = (long double)z[j] - z[i];
            long double ux = (long double)vx[j] - vx[i];
            long double uy = (long double)vy[j] - vy[i];
            long double uz = (long double)vz[j] - vz[i];
            long double s = dx * dx + dy * dy + dz * dz + eps2;
            long double vr = dx * ux + dy * uy + dz * uz;
            long double inv = 1.0L / sqrtl(s);
            long double inv2 = inv * inv;
            long double tj = w[j] * inv * inv2;
            long double qq = -3.0L * vr * inv2;
            a[0] += dx * tj; a[1] += dy * tj; a[2] += dz * tj;
            jk[0] += tj * (ux + qq * dx);
            jk[1] += tj * (uy + qq * dy);
            jk[2] += tj * (uz + qq * dz);
        }
        for (int c = 0; c < 3; c++) {
            out6[(size_t)k * 6 + c] = (double)a[c];
            out6[(size_t)k * 6 + 3 + c] = (double)jk[c];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Predictor, src/nbody.cpp:152-163 (same expression order; the macro  */
/* ONESIXTH expands unparenthesised to 1.0/6.0, which binds first).    */
/* ------------------------------------------------------------------ */
void or_predict(int n, double dt, double *x, double *y, double *z,
                double *vx, double *vy, double *vz, const double *ax,
                const double *ay, const double *az, const double *jx,
                const double *jy, const double *jz)
{
    for (int i = 0; i < n; i++) {
        x[i] += vx[i] * dt + 0.5 * ax[i] * dt * dt + 1.0 / 6.0 * jx[i] * dt * dt * dt;
        y[i] += vy[i] * dt + 0.5 * ay[i] * dt * dt + 1.0 / 6.0 * jy[i] * dt * dt * dt;
        z[i] += vz[i] * dt + 0.5 * az[i] * dt * dt + 1.0 / 6.0 * jz[i] * dt * dt * dt;
        vx[i] += ax[i] * dt + 0.5 * jx[i] * dt * dt;
        vy[i] += ay[i] * dt + 0.5 * jy[i] * dt * dt;
        vz[i] += az[i] * dt + 0.5 * jz[i] * dt * dt;
    }
}

/* ------------------------------------------------------------------ */
/* Snap/crackle + corrector, src/nbody.cpp:165-196.  a0,j0 = forces at */
/* the start of the step; a1,j1 = forces at the predicted state.       */
/* Outputs a2[3n], a3[3n] (x block, y block, z block).                 */
/* ------------------------------------------------------------------ */
void or_snap_correct(int n, double dt, double *x, double *y, double *z,
                     double *vx, double *vy, double *vz,
                     const double *a0, const double *j0,
                     const double *a1, const double *j1,
                     double *a2, double *a3)
{
    /* a0/j0/a1/j1/a2/a3 are [3][n] blocks */
    for (int i = 0; i < n; i++) {
        double invdt2 = 1.0 / (dt * dt);
        double invdt3 = 1.0 / (dt * dt * dt);
        for (int c = 0; c < 3; c++) {
            size_t k = (size_t)c * n + i;
            a2[k] = 2.0 * invdt2 * (-3.0 * (a0[k] - a1[k]) - (2.0 * j0[k] + j1[k]) * dt);
            a3[k] = 6.0 * invdt3 * (2.0 * (a0[k] - a1[k]) + (j0[k] + j1[k]) * dt);
        }
    }
    double *pos[3] = {x, y, z};
    double *vel[3] = {vx, vy, vz};
    for (int i = 0; i < n; i++) {
        double dt3 = dt * dt * dt;
        double dt4 = dt3 * dt;
        double dt5 = dt4 * dt;
        for (int c = 0; c < 3; c++) {
            size_t k = (size_t)c * n + i;
            pos[c][i] += 1.0 / 24.0 * a2[k] * dt4 + 1.0 / 120.0 * a3[k] * dt5;
        }
        for (int c = 0; c < 3; c++) {
            size_t k = (size_t)c * n + i;
            vel[c][i] += 1.0 / 6.0 * a2[k] * dt3 + 1.0 / 24.0 * a3[k] * dt4;
        }
    }
}

/* Aarseth step criterion, src/nbody.cpp:198-222 (ETA = 1e-2): uses the */
/* start-of-step a0, j0 with the a2, a3 just computed.                   */
double or_updatedt(int n, double eta, const double *a0, const double *j0,
                   const double *a2, const double *a3)
{
    double best = DBL_MAX;
    for (int i = 0; i < n; i++) {
        const double *A = a0, *J = j0, *S = a2, *C = a3;
        size_t X = i, Y = (size_t)n + i, Z = 2 * (size_t)n + i;
        double na2 = A[X] * A[X] + A[Y] * A[Y] + A[Z] * A[Z];
        double na = sqrt(na2);
        double nj2 = J[X] * J[X] + J[Y] * J[Y] + J[Z] * J[Z];
        double nj = sqrt(nj2);
        double ns2 = S[X] * S[X] + S[Y] * S[Y] + S[Z] * S[Z];
        double ns = sqrt(ns2);
        double nc2 = C[X] * C[X] + C[Y] * C[Y] + C[Z] * C[Z];
        double nc = sqrt(nc2);
        double num = na * ns + nj2;
        double den = nj * nc + ns2;
        double cur = sqrt(eta * num / den);
        if (cur < best)
            best = cur;
    }
    return best;
}

/* Initial step, src/nbody.cpp:131-150 (ETA0 = 1e-4). */
double or_initialdt(int n, double eta0, const double *a, const double *j)
{
    double best = DBL_MAX;
    for (int i = 0; i < n; i++) {
        double a2x = a[i] * a[i];
        double a2y = a[n + i] * a[n + i];
        double a2z = a[2 * (size_t)n + i] * a[2 * (size_t)n + i];
        double an = sqrt(a2x + a2y + a2z);
        double j2x = j[i] * j[i];
        double j2y = j[n + i] * j[n + i];
        double j2z = j[2 * (size_t)n + i] * j[2 * (size_t)n + i];
        double jn = sqrt(j2x + j2y + j2z);
        double cur = eta0 * an / jn;
        if (cur < best)
            best = cur;
    }
    return best;
}

/* Kinetic energy, src/utils.cpp:125-135 (serial order). */
double or_kinetic(int n, const double *w, const double *vx, const double *vy,
                  const double *vz)
{
    double t = 0.0;
    for (int i = 0; i < n; i++) {
        double a = vx[i] * vx[i];
        double b = vy[i] * vy[i];
        double c = vz[i] * vz[i];
        t += 0.5 * w[i] * (a + b + c);
    }
    return t;
}

/* Potential as the reference computes it, src/utils.cpp:137-153:      */
/* -sum_{i<j} w_j*w_j / sqrt(r^2 + eps2) (the m_j^2 weight is the       */
/* reference's, kept here for compatibility), serial inner order.       */
double or_potential_compat(int n, const double *x, const double *y,
                           const double *z, const double *w, double eps2)
{
    double u = 0.0;
    for (int i = 0; i < n; i++) {
        double t = 0.0;
        for (int j = i + 1; j < n; j++) {
            double rx2 = (x[i] - x[j]) * (x[i] - x[j]);
            double ry2 = (y[i] - y[j]) * (y[i] - y[j]);
            double rz2 = (z[i] - z[j]) * (z[i] - z[j]);
            double r2 = rx2 + ry2 + rz2 + eps2;
            double inv = 1. / sqrt(r2);
            t -= w[j] * w[j] * inv;
        }
        u += t;
    }
    return u;
}

/* Physical softened potential -sum_{i<j} m_i m_j / sqrt(r^2 + eps2),  */
/* accumulated in long double (the energy oracle for drift checks).     */
double or_potential(int n, const double *x, const double *y, const double *z,
                    const double *w, double eps2, int nthreads)
{
    long double u = 0.0L;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : u) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int i = 0; i < n; i++) {
        long double t = 0.0L;
        for (int j = i + 1; j < n; j++) {
            long double dx = (long double)x[i] - x[j];
            long double dy = (long double)y[i] - y[j];
            long double dz = (long double)z[i] - z[j];
            t += (long double)w[j] / sqrtl(dx * dx + dy * dy + dz * dz + eps2);
        }
        u -= (long double)w[i] * t;
    }
    return (double)u;
}

/* ------------------------------------------------------------------ */
/* Reference initial conditions, src/nbody.cpp:69-114: mt19937(42),    */
/* uniform_real_distribution<double>(-1,1) as libstdc++ draws it        */
/* (generate_canonical with two 32-bit draws), rejection in a sphere,   */
/* power-law masses with the u in (-1,1) draw, and the float virial     */
/* rescale factor of :104.                                              */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t mt[624];
    int idx;
} mt32;

static void mt32_seed(mt32 *g, uint32_t s)
{
    g->mt[0] = s;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

static uint32_t mt32_next(mt32 *g)
{
    if (g->idx >= 624) {
        for (int k = 0; k < 624; k++) {
            uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(k + 397) % 624] ^ (y >> 1);
            if (y & 1u)
                v ^= 0x9908b0dfu;
            g->mt[k] = v;
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* a + (b - a) * canonical, with canonical built from two 32-bit draws */
static double mt32_uniform(mt32 *g, double a, double b)
{
    double sum = (double)mt32_next(g);
    sum += (double)mt32_next(g) * 4294967296.0;
    double u = sum / 18446744073709551616.0;
    if (u >= 1.0)
        u = nextafter(1.0, 0.0);
    return u * (b - a) + a;
}

int or_reference_ic(int n, double vmax, double rmax, double mmin, double mmax,
                    double *x, double *y, double *z, double *w, double *vx,
                    double *vy, double *vz)
{
    mt32 *g = (mt32 *)malloc(sizeof(mt32));
    if (!g)
        return -1;
    mt32_seed(g, 42u);
    const double alpha = -2.35 + 1.0;
    int i = 0;
    while (i < n) {
        double xt = mt32_uniform(g, -1.0, 1.0) * rmax;
        double yt = mt32_uniform(g, -1.0, 1.0) * rmax;
        double zt = mt32_uniform(g, -1.0, 1.0) * rmax;
        double r2 = xt * xt + yt * yt + zt * zt;
        if (r2 < rmax * rmax) {
            x[i] = xt; y[i] = yt; z[i] = zt;
            i++;
        }
    }
    double m1 = pow(mmin, alpha);
    double m2 = pow(mmax, alpha);
    double ia = 1.0 / alpha;
    for (i = 0; i < n; i++) {
        vx[i] = mt32_uniform(g, -1.0, 1.0) * vmax;
        vy[i] = mt32_uniform(g, -1.0, 1.0) * vmax;
        vz[i] = mt32_uniform(g, -1.0, 1.0) * vmax;
        double t = (m2 - m1) * mt32_uniform(g, -1.0, 1.0) + m1;
        w[i] = pow(t, ia);
    }
    free(g);
    double T = or_kinetic(n, w, vx, vy, vz);
    double U = or_potential_compat(n, x, y, z, w, 1.0e-7);
    double Q = -2.0 * T / U;
    float f = (float)(1.0 / sqrt(Q));
    for (i = 0; i < n; i++) {
        vx[i] = vx[i] * f;
        vy[i] = vy[i] * f;
        vz[i] = vz[i] * f;
    }
    return 0;
}
