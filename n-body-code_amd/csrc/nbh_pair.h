// nbh_pair.h -- the per-pair arithmetic shared by every force kernel
// (single system: nbh_kernels.hip; ensemble: nbh_ensemble.hip).
//
// Per directed pair i <- j:
//   d = x_j - x_i, u = v_j - v_i, s = |d|^2 + eps2, rv = d.u, r = s^-1/2,
//   a_i += m_j r^3 d,  jerk_i += m_j r^3 (u - 3 rv r^2 d)
// the reference's per-pair formula (src/kernel.cpp:1188-1205).
#pragma once

#include "nbh_internal.h"

namespace nbh {

// FP64 source body: 64 bytes (pad is never read).  Held in LDS tiles by the
// ensemble kernel (three 16-byte and one 8-byte broadcast read) and packed in
// HBM for the single-system directed kernel, which reads it with scalar loads.
struct __align__(16) JB64 {
    double x, y, z, m, vx, vy, vz, pad;
};

// Mixed-precision source body: FP32 velocity and mass first, FP64 position
// after, 48 bytes.  Held in LDS tiles (ensemble, symmetric kernels: two
// 16-byte and one 8-byte read) or packed in HBM like JB64.
struct __align__(16) JBMX {
    float vx, vy, vz, m;
    double x, y, z, pad;
};

// s^-1/2 in FP64 for every FP64 kernel: the v_rsq_f64 seed (measured
// relative error <= 5.2e-8 on gfx950) refined by one third-order step,
// y' = y + y e (1/2 + 3/8 e) with e = 1 - s y^2.  The seed's relative error
// d becomes O(d^3) ~ 1e-22 plus the rounding of e and of the last fma, i.e.
// about one ulp -- closer to 1/sqrt(s) than the reference's InvSqrt2D
// (src/utils.cpp:155-165, up to 4.3e-16), which a plain Newton step
// (d -> 1.5 d^2 ~ 4e-15) is not.
__device__ __forceinline__ double rsqrt_f64(double s)
{
    const double y = __builtin_amdgcn_rsq(s);
    const double e = fma(-s, y * y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

__device__ __forceinline__ JB64 make_jb64(double x, double y, double z, double m, double vx, double vy, double vz)
{
    JB64 b;
    b.x = x;
    b.y = y;
    b.z = z;
    b.m = m;
    b.vx = vx;
    b.vy = vy;
    b.vz = vz;
    b.pad = 0.0;
    return b;
}

// One source body gathered from the SoA state (coalesced when lanes take
// consecutive j).
__device__ __forceinline__ JB64 load_jb64(const State& s, int j)
{
    return make_jb64(s.x[j], s.y[j], s.z[j], s.m[j], s.vx[j], s.vy[j], s.vz[j]);
}

// Reference velocity of a mixed-precision evaluation: every velocity is
// stored in FP32 as (float)(v - vref), with vref one FP64 velocity per system
// (its centre-of-mass velocity, k_vref_*).  Velocity differences, the only
// way velocities enter the jerk, are unchanged by the shift, but their FP32
// rounding error then scales with the velocity dispersion about vref rather
// than with the system's bulk velocity: a cluster moving at 1e3 times its
// internal speeds keeps the jerk accuracy it has at rest.
struct VRef {
    double x, y, z;
};

__device__ __forceinline__ VRef load_vref(const double* __restrict__ p)
{
    return VRef{p[0], p[1], p[2]};
}

__device__ __forceinline__ JBMX make_jbmx(double x, double y, double z, double m, double vx, double vy, double vz,
                                          const VRef& vr)
{
    JBMX b;
    b.vx = (float)(vx - vr.x);
    b.vy = (float)(vy - vr.y);
    b.vz = (float)(vz - vr.z);
    b.m = (float)m;
    b.x = x;
    b.y = y;
    b.z = z;
    b.pad = 0.0;
    return b;
}

__device__ __forceinline__ JBMX load_jbmx(const State& s, int j, const VRef& vr)
{
    return make_jbmx(s.x[j], s.y[j], s.z[j], s.m[j], s.vx[j], s.vy[j], s.vz[j], vr);
}

// ---------------------------------------------------------------- FP64 --
template <bool MASK>
__device__ __forceinline__ void pair_f64(const JB64& b, int j, int i,
                                         double xi, double yi, double zi,
                                         double vxi, double vyi, double vzi,
                                         double eps2, double (&acc)[6])
{
    const double dx = b.x - xi;
    const double dy = b.y - yi;
    const double dz = b.z - zi;
    const double ux = b.vx - vxi;
    const double uy = b.vy - vyi;
    const double uz = b.vz - vzi;
    double s = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2)));
    const double vr = fma(dx, ux, fma(dy, uy, dz * uz));
    double mj = b.m;
    if (MASK) {  // eps2 too small for a finite self pair (self_mask): drop it
        const bool self = (j == i);
        s = self ? 1.0 : s;
        mj = self ? 0.0 : mj;
    }
    const double r = rsqrt_f64(s);
    const double r2 = r * r;
    const double mr3 = mj * r * r2;
    const double al = -3.0 * vr * r2;
    acc[0] = fma(mr3, dx, acc[0]);
    acc[1] = fma(mr3, dy, acc[1]);
    acc[2] = fma(mr3, dz, acc[2]);
    acc[3] = fma(mr3, fma(al, dx, ux), acc[3]);
    acc[4] = fma(mr3, fma(al, dy, uy), acc[4]);
    acc[5] = fma(mr3, fma(al, dz, uz), acc[5]);
}

// --------------------------------------------------------------- mixed --
// Position differences are formed in FP64 and rounded; everything after is
// FP32.  Terms are summed in FP32 over runs of kFold sources, and each run
// total is folded into an FP32 (hi, lo) pair with Knuth's two-sum, so the
// rounding error of the long sum is carried in lo.
constexpr int kFold = 64;

__device__ __forceinline__ void two_sum_fold(float& hi, float& lo, float v)
{
    const float s = hi + v;
    const float bv = s - hi;
    const float err = (hi - (s - bv)) + (v - bv);
    hi = s;
    lo += err;
}

template <bool MASK>
__device__ __forceinline__ void pair_mx(const JBMX& b, int j, int i,
                                        double xi, double yi, double zi,
                                        float vxi, float vyi, float vzi,
                                        float eps2, float (&acc)[6])
{
    const float dx = (float)(b.x - xi);
    const float dy = (float)(b.y - yi);
    const float dz = (float)(b.z - zi);
    const float ux = b.vx - vxi;
    const float uy = b.vy - vyi;
    const float uz = b.vz - vzi;
    float s = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
    const float vr = fmaf(dx, ux, fmaf(dy, uy, dz * uz));
    float mj = b.m;
    if (MASK) {
        const bool self = (j == i);
        s = self ? 1.0f : s;
        mj = self ? 0.0f : mj;
    }
    const float r = __builtin_amdgcn_rsqf(s);
    const float r2 = r * r;
    const float mr3 = mj * r * r2;
    const float al = -3.0f * vr * r2;
    acc[0] = fmaf(mr3, dx, acc[0]);
    acc[1] = fmaf(mr3, dy, acc[1]);
    acc[2] = fmaf(mr3, dz, acc[2]);
    acc[3] = fmaf(mr3, fmaf(al, dx, ux), acc[3]);
    acc[4] = fmaf(mr3, fmaf(al, dy, uy), acc[4]);
    acc[5] = fmaf(mr3, fmaf(al, dz, uz), acc[5]);
}

}  // namespace nbh
