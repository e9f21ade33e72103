// nbh_adjoint.hip -- reverse-mode derivatives of the force evaluation and of
// the Hermite trajectory (FP64).
//
// Force VJP.  The forward map is, per directed pair (target i, source j),
//   a_i    += m_j q d,                      q = s^-3/2, s = |d|^2 + eps2
//   jerk_i += m_j (q u - 3 (d.u) p d),      p = s^-5/2, d = x_j - x_i, u = v_j - v_i
// (src/kernel.cpp:1188-1205).  Given cotangents (alpha_i, beta_i) of (a_i,
// jerk_i), a pair contributes to body k both as a target and as a source.
// With G(A, B; d, u) = q A - 3p (A.d + B.u) d - 3p ((B.d) u + (d.u) B)
//                      + 15 (d.u)(B.d) s^-7/2 d
// and H(B; d) = q B - 3p (B.d) d, both even in (d, u) and linear in (A, B),
// the two roles combine per pair (k, o) into
//   xbar_k += G(m_k alpha_o - m_o alpha_k, m_k beta_o - m_o beta_k)
//   vbar_k += H(m_k beta_o - m_o beta_k)
//   mbar_k += -q (alpha_o.d + beta_o.u) + 3 (d.u) p (beta_o.d)
// with d = x_o - x_k, u = v_o - v_k: a gather over o, the same streaming
// pattern as the force (one target per lane, sources through LDS, j-split
// with an ordered second pass), no atomics.
//
// Step VJP.  One step (src/nbody.cpp:152-196, fixed dt) is, with
// (a1, j1) = F(x_p, v_p, m),
//   x_p = x + v dt + a dt^2/2 + j dt^3/6,   v_p = v + a dt + j dt^2/2
//   x'  = x_p - 0.15 dt^2 (a - a1) - 7/60 dt^3 j - 1/30 dt^3 j1
//   v'  = v_p - 0.5 dt (a - a1) - 5/12 dt^2 j - 1/12 dt^2 j1
// and carries (x', v', a1, j1); the corrector's snap/crackle algebra
// reduces to those coefficients.  Its adjoint is two O(n) halves around one
// force VJP; the second half of step t and the first of step t-1 run as one
// kernel (k_adj_pre_post).  Trajectories keep checkpoints every `every`
// steps and recompute each segment forward during the backward sweep,
// keeping every step's prediction, the VJP's input; the recompute of segment
// k-1 runs on a second stream beside segment k's step adjoints.  A forward
// step is the force at the prediction and one kernel (k_adj_advance) that
// adds the split sums, corrects, and predicts the next step.
#include "nbh_ctx.h"
#include "nbh_hermite.h"
#include "nbh_pair.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace nbh {
namespace {

constexpr int kVB = 64;       // targets per workgroup of the VJP kernel
constexpr int kOB = 256;      // threads per workgroup of the O(n) kernels

// Source record in LDS: position, mass, velocity, and the cotangents of the
// source's acceleration and jerk.
struct __align__(16) VJB {
    double x, y, z, m;
    double vx, vy, vz, p0;
    double ax, ay, az, p1;
    double jx, jy, jz, p2;
};

struct VjpArgs {
    State s;
    const double* cot;  // [6][n]: alpha (3 blocks), beta (3 blocks)
    double* out;        // [split][7][n]: xbar(3), mbar, vbar(3)
    int n;
    int jchunk;
    double eps2;
};

template <bool MASK>
__global__ __launch_bounds__(kVB) void k_force_vjp(VjpArgs a)
{
    __shared__ VJB tile[kVB];
    const int t = threadIdx.x;
    const int k = blockIdx.x * kVB + t;
    const int n = a.n;
    const int kc = k < n ? k : n - 1;
    const double xk = a.s.x[kc], yk = a.s.y[kc], zk = a.s.z[kc], mk = a.s.m[kc];
    const double vxk = a.s.vx[kc], vyk = a.s.vy[kc], vzk = a.s.vz[kc];
    const double axk = a.cot[kc], ayk = a.cot[n + kc], azk = a.cot[2 * (size_t)n + kc];
    const double bxk = a.cot[3 * (size_t)n + kc], byk = a.cot[4 * (size_t)n + kc], bzk = a.cot[5 * (size_t)n + kc];
    double gx = 0, gy = 0, gz = 0, hx = 0, hy = 0, hz = 0, gm = 0;
    const int jbeg = blockIdx.y * a.jchunk;
    const int jend = min(n, jbeg + a.jchunk);
    for (int jt = jbeg; jt < jend; jt += kVB) {
        const int cnt = min(kVB, jend - jt);
        __syncthreads();
        if (t < cnt) {
            const int o = jt + t;
            VJB b;
            b.x = a.s.x[o]; b.y = a.s.y[o]; b.z = a.s.z[o]; b.m = a.s.m[o];
            b.vx = a.s.vx[o]; b.vy = a.s.vy[o]; b.vz = a.s.vz[o];
            b.ax = a.cot[o]; b.ay = a.cot[n + o]; b.az = a.cot[2 * (size_t)n + o];
            b.jx = a.cot[3 * (size_t)n + o]; b.jy = a.cot[4 * (size_t)n + o]; b.jz = a.cot[5 * (size_t)n + o];
            b.p0 = b.p1 = b.p2 = 0.0;
            tile[t] = b;
        }
        __syncthreads();
        for (int q0 = 0; q0 < cnt; ++q0) {
            const VJB& b = tile[q0];
            const double dx = b.x - xk, dy = b.y - yk, dz = b.z - zk;
            const double ux = b.vx - vxk, uy = b.vy - vyk, uz = b.vz - vzk;
            double s = fma(dx, dx, fma(dy, dy, fma(dz, dz, a.eps2)));
            double mo = b.m, mkk = mk;
            if (MASK) {
                const bool self = (jt + q0) == k;
                s = self ? 1.0 : s;
                mo = self ? 0.0 : mo;
                mkk = self ? 0.0 : mkk;
            }
            const double r = rsqrt_f64(s);
            const double r2 = r * r;
            const double q = r * r2;
            const double p = q * r2;
            const double ps = p * r2;
            const double du = fma(dx, ux, fma(dy, uy, dz * uz));
            // combined cotangents of the pair
            const double Ax = fma(mkk, b.ax, -mo * axk), Ay = fma(mkk, b.ay, -mo * ayk), Az = fma(mkk, b.az, -mo * azk);
            const double Bx = fma(mkk, b.jx, -mo * bxk), By = fma(mkk, b.jy, -mo * byk), Bz = fma(mkk, b.jz, -mo * bzk);
            const double Ad = fma(Ax, dx, fma(Ay, dy, Az * dz));
            const double Bu = fma(Bx, ux, fma(By, uy, Bz * uz));
            const double Bd = fma(Bx, dx, fma(By, dy, Bz * dz));
            const double p3 = -3.0 * p;
            const double c1 = fma(15.0 * du * Bd, ps, p3 * (Ad + Bu));
            // -3p ((B.d) u + (d.u) B) as e u + f B
            const double e = p3 * Bd;
            const double f = p3 * du;
            gx = fma(q, Ax, fma(c1, dx, fma(e, ux, fma(f, Bx, gx))));
            gy = fma(q, Ay, fma(c1, dy, fma(e, uy, fma(f, By, gy))));
            gz = fma(q, Az, fma(c1, dz, fma(e, uz, fma(f, Bz, gz))));
            hx = fma(q, Bx, fma(e, dx, hx));
            hy = fma(q, By, fma(e, dy, hy));
            hz = fma(q, Bz, fma(e, dz, hz));
            // mass: the source role only (force is linear in the source mass)
            const double ad = fma(b.ax, dx, fma(b.ay, dy, b.az * dz));
            const double bu = fma(b.jx, ux, fma(b.jy, uy, b.jz * uz));
            const double bd = fma(b.jx, dx, fma(b.jy, dy, b.jz * dz));
            const double gmo = fma(-f, bd, -q * (ad + bu));   // 3 (d.u) p (beta_o.d) - q (...)
            gm += MASK ? ((jt + q0) == k ? 0.0 : gmo) : gmo;
        }
    }
    if (k < n) {
        double* o = a.out + (size_t)blockIdx.y * 7 * n;
        o[k] = gx;
        o[n + k] = gy;
        o[2 * (size_t)n + k] = gz;
        o[3 * (size_t)n + k] = gm;
        o[4 * (size_t)n + k] = hx;
        o[5 * (size_t)n + k] = hy;
        o[6 * (size_t)n + k] = hz;
    }
}

// grad[c][k] (+)= sum over splits, in split order.
__global__ __launch_bounds__(kOB) void k_vjp_reduce(const double* __restrict__ part, int n, int splits,
                                                    double* grad, int accumulate)
{
    const int k = blockIdx.x * kOB + threadIdx.x;
    const int c = blockIdx.y;
    if (k >= n) return;
    double v = 0.0;
    // unrolled so that eight loads are in flight at once; the adds keep split order
#pragma unroll 8
    for (int s = 0; s < splits; ++s) v += part[((size_t)s * 7 + c) * n + k];
    const size_t o = (size_t)c * n + k;
    grad[o] = accumulate ? grad[o] + v : v;
}

// Step adjoint, first half: from the cotangents of the step's outputs
// (xbar', vbar', abar', jbar' in `cot` = [12][n]: x(3) v(3) a(3) j(3)),
// writes the cotangents of (a1, j1) into cf = [6][n], keeps xbar_p = xbar',
// vbar_p = vbar' in cot[0..5], and puts the direct part of (abar, jbar)
// into cot[6..11].
__device__ __forceinline__ void adj_post_body(int i, int n, double dt, double* cot, double* cf)
{
    const double ca = -0.15 * dt * dt, cj = -7.0 / 60.0 * dt * dt * dt, cj1 = -1.0 / 30.0 * dt * dt * dt;
    const double da = -0.5 * dt, dj = -5.0 / 12.0 * dt * dt, dj1 = -1.0 / 12.0 * dt * dt;
    for (int c = 0; c < 3; ++c) {
        const size_t X = (size_t)c * n + i, V = (size_t)(3 + c) * n + i;
        const size_t A = (size_t)(6 + c) * n + i, J = (size_t)(9 + c) * n + i;
        const double xb = cot[X], vb = cot[V];
        cf[(size_t)c * n + i] = cot[A] - ca * xb - da * vb;        // abar of a1
        cf[(size_t)(3 + c) * n + i] = cot[J] + cj1 * xb + dj1 * vb; // jbar of j1
        cot[A] = ca * xb + da * vb;                                 // abar of a (direct)
        cot[J] = cj * xb + dj * vb;                                 // jbar of j (direct)
    }
}

// Step adjoint, second half: adds the force VJP g = (x(3) m v(3)) of body i
// to (xbar_p, vbar_p), accumulates mbar, and pulls the cotangents back
// through the predictor onto the step's inputs (x, v, a, j).
__device__ __forceinline__ void adj_pre_body(int i, int n, double dt, double* cot, const double (&g)[7],
                                             double* mbar)
{
    mbar[i] += g[3];
    for (int c = 0; c < 3; ++c) {
        const size_t X = (size_t)c * n + i, V = (size_t)(3 + c) * n + i;
        const size_t A = (size_t)(6 + c) * n + i, J = (size_t)(9 + c) * n + i;
        const double xp = cot[X] + g[c];
        const double vp = cot[V] + g[4 + c];
        cot[X] = xp;
        cot[V] = fma(xp, dt, vp);
        cot[A] = cot[A] + xp * (0.5 * dt * dt) + vp * dt;
        cot[J] = cot[J] + xp * (dt * dt * dt / 6.0) + vp * (0.5 * dt * dt);
    }
}

__global__ __launch_bounds__(kOB) void k_adj_post(int n, double dt, double* cot, double* cf)
{
    const int i = blockIdx.x * kOB + threadIdx.x;
    if (i >= n) return;
    adj_post_body(i, n, dt, cot, cf);
}

// Bodies per workgroup of the two fused step kernels below.  Their split
// sums take one thread per body and component (a body's six or seven sums
// done by one thread would run one after another, ~3x slower at n = 4096),
// combined through LDS for the body's own arithmetic, done by the first wave.
constexpr int kAB = 64;

// The second half of step t's adjoint, then (post != 0) the first half of
// step t-1's, for kAB bodies.  g is the force VJP of step t, summed here
// over its split planes (split s, component c at (s*7 + c)*n) in split
// order, as k_vjp_reduce does.
__global__ __launch_bounds__(7 * kAB) void k_adj_pre_post(int n, double dt, double* cot,
                                                          const double* __restrict__ part, int splits, double* mbar,
                                                          double* cf, int post)
{
    __shared__ double gs[7][kAB];
    const int b = threadIdx.x % kAB, c = threadIdx.x / kAB;
    const int i = blockIdx.x * kAB + b;
    if (i < n) {
        double v = 0.0;
#pragma unroll 8
        for (int s = 0; s < splits; ++s) v += part[((size_t)s * 7 + c) * n + i];
        gs[c][b] = v;
    }
    __syncthreads();
    if (c != 0 || i >= n) return;
    double g[7];
    for (int k = 0; k < 7; ++k) g[k] = gs[k][b];
    adj_pre_body(i, n, dt, cot, g, mbar);
    if (post) adj_post_body(i, n, dt, cot, cf);
}

// The end of one forward step for kAB bodies, fixed dt (*dtp): the new
// forces (a1, j1) summed over the pair kernel's split planes (component c,
// split k at (c*splits + k)*n) in split order, as k_reduce_splits does (one
// split: its plane as is, as the unsplit kernel writes it); the corrector on
// the step's prediction `pin` (x, m, v: 7n) with the start-of-step forces of
// `in` (state and forces: 13n), as k_correct; the corrected state and
// (a1, j1) written to `out` (13n); and, when pnext is set, the next step's
// prediction from them, as k_predict, written to pnext (7n).  Every input of
// a body is read before any of its outputs is written, so in == out and
// pin == pnext are allowed.
__global__ __launch_bounds__(6 * kAB) void k_adj_advance(int n, const double* dtp, const double* __restrict__ part,
                                                         int splits, const double* in, const double* pin,
                                                         double* out, double* pnext)
{
#pragma clang fp contract(off)
    __shared__ double fs[6][kAB];
    const int b = threadIdx.x % kAB, c = threadIdx.x / kAB;
    const int i = blockIdx.x * kAB + b;
    if (i < n) {
        const double* p = part + (size_t)c * splits * n + i;
        double v;
        if (splits == 1) {
            v = p[0];
        } else {
            v = 0.0;
#pragma unroll 8
            for (int s = 0; s < splits; ++s) v += p[(size_t)s * n];
        }
        fs[c][b] = v;
    }
    __syncthreads();
    if (c != 0 || i >= n) return;
    const double dt = *dtp;
    const double invdt2 = 1.0 / (dt * dt);
    const double invdt3 = 1.0 / (dt * dt * dt);
    double f1[6], a0[3], j0[3], x[3], u[3];
    for (int k = 0; k < 6; ++k) f1[k] = fs[k][b];
    for (int k = 0; k < 3; ++k) {
        a0[k] = in[(size_t)(7 + k) * n + i];
        j0[k] = in[(size_t)(10 + k) * n + i];
        x[k] = pin[(size_t)k * n + i];
        u[k] = pin[(size_t)(4 + k) * n + i];
    }
    const double m = in[(size_t)3 * n + i];
    for (int k = 0; k < 3; ++k) {
        double a2, a3;
        hermite_correct_axis(dt, invdt2, invdt3, a0[k], j0[k], f1[k], f1[3 + k], x[k], u[k], a2, a3);
    }
    for (int k = 0; k < 3; ++k) {
        out[(size_t)k * n + i] = x[k];
        out[(size_t)(4 + k) * n + i] = u[k];
        out[(size_t)(7 + k) * n + i] = f1[k];
        out[(size_t)(10 + k) * n + i] = f1[3 + k];
    }
    out[(size_t)3 * n + i] = m;
    if (pnext) {
        for (int k = 0; k < 3; ++k) {
            hermite_predict_axis(dt, f1[k], f1[3 + k], x[k], u[k]);
            pnext[(size_t)k * n + i] = x[k];
            pnext[(size_t)(4 + k) * n + i] = u[k];
        }
        pnext[(size_t)3 * n + i] = m;
    }
}

inline int oblk(int n) { return (n + kOB - 1) / kOB; }

}  // namespace

// Sources per split and split planes the VJP kernel writes for `splits`
// requested splits.
static int vjp_chunk(int n, int splits)
{
    const int tiles = (n + kVB - 1) / kVB;
    return (tiles + splits - 1) / splits * kVB;
}
static int vjp_planes(int n, int splits) { return (n + vjp_chunk(n, splits) - 1) / vjp_chunk(n, splits); }

// Force VJP on device arrays: grad = [7][n] (x(3), m, v(3)) of
// sum_i alpha_i.a_i + beta_i.jerk_i, cot = [6][n].  accumulate adds to grad.
// With grad null the split planes are left in `partial` (vjp_planes of
// them, split s and component c at (s*7 + c)*n) for the caller to add.
static hipError_t launch_force_vjp(const State& s, int n, double eps2, const double* cot, double* grad,
                                   int accumulate, double* partial, int splits, hipStream_t st)
{
    VjpArgs a;
    a.s = s;
    a.cot = cot;
    a.out = partial;
    a.n = n;
    a.jchunk = vjp_chunk(n, splits);
    const int S = vjp_planes(n, splits);
    const int tiles = (n + kVB - 1) / kVB;
    a.eps2 = eps2;
    const dim3 grid(tiles, S);
    // the self pair is dropped explicitly when eps2 is small enough for its
    // r^7 = eps2^-7/2 term to overflow FP64 (eps2 < ~1e-88); above, it adds
    // finite multiples of d = u = 0
    if (eps2 < 1e-80)
        k_force_vjp<true><<<grid, kVB, 0, st>>>(a);
    else
        k_force_vjp<false><<<grid, kVB, 0, st>>>(a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !grad) return e;
    k_vjp_reduce<<<dim3(oblk(n), 7), kOB, 0, st>>>(partial, n, S, grad, accumulate);
    return hipGetLastError();
}

// j-split of the FP64 force evaluations inside the trajectory (64-lane
// workgroups, about 8 per CU; 16 per CU measured slower at n = 4096).
static int adj_force_splits(int n, int cus)
{
    const int tiles = (n + 63) / 64;
    return std::max(1, std::min(tiles, (8 * (cus > 0 ? cus : 256) + tiles - 1) / tiles));
}

// j-split of the force VJP: one-wave workgroups, about 8 per CU (at n = 4096
// 6% faster than 4 per CU; 16 per CU no faster).
static int vjp_splits(int n, int cus)
{
    const int tiles = (n + kVB - 1) / kVB;
    const int want = 8 * (cus > 0 ? cus : 256);
    int s = (want + tiles - 1) / tiles;
    return std::max(1, std::min(s, tiles));
}

}  // namespace nbh

using nbh::DevBuf;
using nbh::fail;

struct nbh_adjoint {
    nbh_ctx* ctx = nullptr;
    int n = 0;
    int every = 1;
    int nsteps = 0;
    double dt = 0.0;
    bool have_forward = false;
    DevBuf ckpt;      // [ncp][13][n]: state(7) forces(6)
    DevBuf cur;       // [13][n]: the recompute's working state
    DevBuf pred;      // [2][every][7][n]: the predictions of two segments' steps
    DevBuf cot;       // [12][n]
    DevBuf cf;        // [6][n]
    DevBuf g;         // [7][n]
    DevBuf mbar;      // [n]
    DevBuf partial;   // [splits][7][n]: the VJP's split planes (and the forward run's force planes)
    DevBuf rpartial;  // [6][splits][n]: the recompute's force split planes
    DevBuf dtbuf;     // [1]
    // The backward sweep recomputes segment k-1 on its own stream while the
    // context's stream runs segment k's step adjoints (default priority).
    hipStream_t rstream = nullptr;
    hipEvent_t ev_start = nullptr;    // context stream: the sweep's inputs are in place
    hipEvent_t ev_ready[2] = {};      // recompute stream: pred[b] holds its segment
    hipEvent_t ev_free[2] = {};       // context stream: pred[b] has been used
};

namespace {

int splits_of(const nbh_adjoint* a) { return nbh::vjp_splits(a->n, a->ctx->cus); }

// Launch shape of the trajectory's FP64 force evaluations: LDS tiles of 64,
// one target per lane, j-split about 8 workgroups per CU.
nbh::Shape adj_shape(const nbh_adjoint* a)
{
    const int n = a->n;
    nbh::Shape sh{};
    sh.block = 64;
    sh.ipl = 1;
    const int tiles = (n + 63) / 64;
    const int split = nbh::adj_force_splits(n, a->ctx->cus);
    const int per = (tiles + split - 1) / split;
    sh.jchunk = per * 64;
    sh.jsplit = (n + sh.jchunk - 1) / sh.jchunk;
    return sh;
}

// One FP64 force evaluation of the 7n state at `st` into `f` (6n).
int adj_force(nbh_adjoint* a, double* st, double* f)
{
    nbh_ctx* c = a->ctx;
    const int n = a->n;
    hipError_t e = nbh::launch_force(NBH_FP64, nbh::state_of(st, n), n, c->cfg.eps2, adj_shape(a),
                                     nbh::forces_of(f, n), a->partial.d(), nullptr, nullptr, c->stream, nullptr,
                                     nullptr);  // LDS tiles
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "adjoint force launch failed: %s", hipGetErrorString(e));
    return NBH_OK;
}

// The prediction of the step that starts at `in` (state and forces, 13n),
// into p (7n), on stream st: a copy, then the predictor in place.
int adj_predict(nbh_adjoint* a, const double* in, double* p, hipStream_t st)
{
    const int n = a->n;
    NBH_HIP(hipMemcpyAsync(p, in, sizeof(double) * 7 * (size_t)n, hipMemcpyDeviceToDevice, st));
    const nbh::Forces f0 = nbh::forces_of(const_cast<double*>(in) + 7 * (size_t)n, n);
    NBH_HIP(nbh::launch_predict(n, a->dtbuf.d(), p, p + n, p + 2 * (size_t)n, p + 4 * (size_t)n, p + 5 * (size_t)n,
                                p + 6 * (size_t)n, f0, st));
    return NBH_OK;
}

// One forward Hermite step from `in` (13n) with its prediction `pin` (7n),
// on stream st with the split planes in `part`: the force at the prediction
// into the planes, then k_adj_advance, which sums them, corrects into `out`
// (13n) and, with pnext set, predicts the next step into pnext (7n).
// in == out and pin == pnext are allowed.  The states and forces are
// bitwise those of the step sequence k_predict, force, k_reduce_splits,
// k_correct.
int adj_step(nbh_adjoint* a, const double* in, double* pin, double* out, double* pnext, hipStream_t st,
             double* part)
{
    nbh_ctx* c = a->ctx;
    const int n = a->n;
    const nbh::Shape sh = adj_shape(a);
    // one split: the kernel writes the six planes directly, at c*n
    hipError_t e = nbh::launch_force(NBH_FP64, nbh::state_of(pin, n), n, c->cfg.eps2, sh, nbh::forces_of(part, n),
                                     part, nullptr, nullptr, st, nullptr, nullptr, 0, false);
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "adjoint force launch failed: %s", hipGetErrorString(e));
    nbh::k_adj_advance<<<(n + nbh::kAB - 1) / nbh::kAB, 6 * nbh::kAB, 0, st>>>(n, a->dtbuf.d(), part, sh.jsplit, in,
                                                                               pin, out, pnext);
    NBH_HIP(hipGetLastError());
    return NBH_OK;
}

// Segment sgi (steps s0 .. s1-1) recomputed from its checkpoint on the
// recompute stream, keeping every step's prediction: step t's in slot
// t - s0 of `pb` ([every][7][n]).
int adj_recompute(nbh_adjoint* a, int sgi, double* pb)
{
    const int n = a->n;
    const int s0 = sgi * a->every;
    const int s1 = std::min(a->nsteps, s0 + a->every);
    hipStream_t st = a->rstream;
    double* cur = a->cur.d();
    NBH_HIP(hipMemcpyAsync(cur, a->ckpt.d() + (size_t)sgi * 13 * n, sizeof(double) * 13 * (size_t)n,
                           hipMemcpyDeviceToDevice, st));
    int rc = adj_predict(a, cur, pb, st);
    if (rc) return rc;
    for (int t = s0; t + 1 < s1; ++t) {
        double* p = pb + (size_t)(t - s0) * 7 * n;
        if ((rc = adj_step(a, cur, p, cur, p + 7 * (size_t)n, st, a->rpartial.d()))) return rc;
    }
    return NBH_OK;
}

size_t sz(int n, int k) { return sizeof(double) * (size_t)k * (size_t)n; }

}  // namespace

extern "C" {

int nbh_force_vjp(nbh_ctx* c, int n, const double* state, const double* cot, double* grad)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (n < 1) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    if (!state || !cot || !grad) return fail(NBH_ERR_ARG, "state/cot/grad array is NULL");
    NBH_HIP(hipSetDevice(c->cfg.device));
    const int splits = nbh::vjp_splits(n, c->cus);
    DevBuf ds, dc, dg, dp;
    if (ds.reserve(sz(n, 7)) != hipSuccess || dc.reserve(sz(n, 6)) != hipSuccess ||
        dg.reserve(sz(n, 7)) != hipSuccess || dp.reserve(sz(n, 7 * splits)) != hipSuccess) {
        ds.release(); dc.release(); dg.release(); dp.release();
        return fail(NBH_ERR_ALLOC, "device allocation for the force VJP failed");
    }
    int rc = NBH_OK;
    hipError_t e = hipMemcpyAsync(ds.p, state, sz(n, 7), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dc.p, cot, sz(n, 6), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = nbh::launch_force_vjp(nbh::state_of(ds.d(), n), n, c->cfg.eps2, dc.d(), dg.d(), 0, dp.d(), splits,
                                  c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(grad, dg.p, sz(n, 7), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(NBH_ERR_HIP, "force VJP failed: %s", hipGetErrorString(e));
    ds.release(); dc.release(); dg.release(); dp.release();
    return rc;
}

int nbh_adj_create(nbh_ctx* c, int n, int every, nbh_adjoint** out)
{
    if (!out) return fail(NBH_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (n < 1) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    if (every < 1) return fail(NBH_ERR_ARG, "checkpoint interval must be >= 1, got %d", every);
    if (c->cfg.precision != NBH_FP64) return fail(NBH_ERR_ARG, "the adjoint needs an FP64 context");
    NBH_HIP(hipSetDevice(c->cfg.device));
    auto* a = new (std::nothrow) nbh_adjoint();
    if (!a) return fail(NBH_ERR_ALLOC, "adjoint allocation failed");
    a->ctx = c;
    a->n = n;
    a->every = every;
    // partial sums of both the force j-split (6 planes per split) and the VJP
    // j-split (7 planes per split)
    const int planes = std::max(6 * nbh::adj_force_splits(n, c->cus), 7 * splits_of(a));
    if (a->cur.reserve(sz(n, 13)) != hipSuccess || a->pred.reserve(sz(n, 14 * every)) != hipSuccess ||
        a->rpartial.reserve(sz(n, 6 * nbh::adj_force_splits(n, c->cus))) != hipSuccess ||
        a->cot.reserve(sz(n, 12)) != hipSuccess || a->cf.reserve(sz(n, 6)) != hipSuccess ||
        a->g.reserve(sz(n, 7)) != hipSuccess || a->mbar.reserve(sz(n, 1)) != hipSuccess ||
        a->partial.reserve(sz(n, planes)) != hipSuccess || a->dtbuf.reserve(sizeof(double)) != hipSuccess) {
        nbh_adj_destroy(a);
        return fail(NBH_ERR_ALLOC, "adjoint device allocation failed (n=%d, every=%d)", n, every);
    }
    hipError_t he = hipStreamCreateWithFlags(&a->rstream, hipStreamNonBlocking);
    for (hipEvent_t* ev : {&a->ev_start, &a->ev_ready[0], &a->ev_ready[1], &a->ev_free[0], &a->ev_free[1]})
        if (he == hipSuccess) he = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (he != hipSuccess) {
        nbh_adj_destroy(a);
        return fail(NBH_ERR_HIP, "adjoint stream/event creation failed: %s", hipGetErrorString(he));
    }
    *out = a;
    return NBH_OK;
}

int nbh_adj_destroy(nbh_adjoint* a)
{
    if (!a) return NBH_OK;
    (void)hipSetDevice(a->ctx->cfg.device);
    (void)hipStreamSynchronize(a->ctx->stream);
    if (a->rstream) (void)hipStreamSynchronize(a->rstream);
    for (hipEvent_t ev : {a->ev_start, a->ev_ready[0], a->ev_ready[1], a->ev_free[0], a->ev_free[1]})
        if (ev) (void)hipEventDestroy(ev);
    if (a->rstream) (void)hipStreamDestroy(a->rstream);
    for (DevBuf* b : {&a->ckpt, &a->cur, &a->pred, &a->cot, &a->cf, &a->g, &a->mbar, &a->partial, &a->rpartial,
                      &a->dtbuf})
        b->release();
    delete a;
    return NBH_OK;
}

int nbh_adj_forward(nbh_adjoint* a, const double* state0, double dt, int nsteps, double* state_final)
{
    if (!a) return fail(NBH_ERR_ARG, "adjoint is NULL");
    if (!state0) return fail(NBH_ERR_ARG, "state0 is NULL");
    if (nsteps < 0) return fail(NBH_ERR_ARG, "nsteps must be >= 0");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(NBH_ERR_ARG, "dt must be finite and > 0");
    nbh_ctx* c = a->ctx;
    if (c->cfg.precision != NBH_FP64) return fail(NBH_ERR_ARG, "the adjoint needs an FP64 context");
    NBH_HIP(hipSetDevice(c->cfg.device));
    // a backward sweep that failed part-way may have left recompute work on
    // its stream, which reads the checkpoints and writes the working state
    NBH_HIP(hipStreamSynchronize(a->rstream));
    const int n = a->n;
    const int ncp = nsteps / a->every + 1;
    if (a->ckpt.reserve(sz(n, 13 * ncp)) != hipSuccess) return fail(NBH_ERR_ALLOC, "checkpoint allocation failed");
    a->have_forward = false;
    a->nsteps = nsteps;
    a->dt = dt;
    NBH_HIP(hipMemcpyAsync(a->dtbuf.p, &a->dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    double* cur = a->cur.d();
    NBH_HIP(hipMemcpyAsync(cur, state0, sz(n, 7), hipMemcpyHostToDevice, c->stream));
    int rc = adj_force(a, cur, cur + 7 * (size_t)n);
    if (rc) return rc;
    double* p = a->pred.d();
    if (nsteps > 0 && (rc = adj_predict(a, cur, p, c->stream))) return rc;
    for (int t = 0; t <= nsteps; ++t) {
        if (t % a->every == 0)
            NBH_HIP(hipMemcpyAsync(a->ckpt.d() + (size_t)(t / a->every) * 13 * n, cur, sz(n, 13),
                                   hipMemcpyDeviceToDevice, c->stream));
        if (t == nsteps) break;
        if ((rc = adj_step(a, cur, p, cur, t + 1 < nsteps ? p : nullptr, c->stream, a->partial.d()))) return rc;
    }
    if (state_final) NBH_HIP(hipMemcpyAsync(state_final, cur, sz(n, 7), hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    a->have_forward = true;
    return NBH_OK;
}

int nbh_adj_backward(nbh_adjoint* a, const double* cot_final, double* grad0)
{
    if (!a) return fail(NBH_ERR_ARG, "adjoint is NULL");
    if (!cot_final || !grad0) return fail(NBH_ERR_ARG, "cot_final/grad0 is NULL");
    if (!a->have_forward) return fail(NBH_ERR_STATE, "no forward run: call nbh_adj_forward first");
    nbh_ctx* c = a->ctx;
    NBH_HIP(hipSetDevice(c->cfg.device));
    const int n = a->n;
    const int splits = splits_of(a);
    const double dt = a->dt;
    double* cot = a->cot.d();
    // cotangent of the final carried state: x, v from the caller, a = j = 0
    NBH_HIP(hipMemsetAsync(cot, 0, sz(n, 12), c->stream));
    NBH_HIP(hipMemcpyAsync(cot, cot_final, sz(n, 3), hipMemcpyHostToDevice, c->stream));
    NBH_HIP(hipMemcpyAsync(cot + 3 * (size_t)n, cot_final + 4 * (size_t)n, sz(n, 3), hipMemcpyHostToDevice,
                           c->stream));
    NBH_HIP(hipMemcpyAsync(a->mbar.p, cot_final + 3 * (size_t)n, sz(n, 1), hipMemcpyHostToDevice, c->stream));
    const int nseg = (a->nsteps + a->every - 1) / a->every;
    const int planes = nbh::vjp_planes(n, splits);
    auto pbuf = [&](int sgi) { return a->pred.d() + (size_t)(sgi & 1) * a->every * 7 * n; };
    // The recompute stream produces segment k-1's predictions while this
    // stream runs segment k's step adjoints; segments alternate between the
    // two halves of pred.  ev_ready[b]: half b holds its segment; ev_free[b]:
    // this stream is done with half b, which the recompute may overwrite.
    NBH_HIP(hipEventRecord(a->ev_start, c->stream));
    NBH_HIP(hipStreamWaitEvent(a->rstream, a->ev_start, 0));
    // on a failure, nothing of the sweep is left running on the recompute stream
    auto failed = [&](int rc) {
        (void)hipStreamSynchronize(a->rstream);
        return rc;
    };
    if (nseg > 0) {
        if (int rc = adj_recompute(a, nseg - 1, pbuf(nseg - 1))) return failed(rc);
        NBH_HIP(hipEventRecord(a->ev_ready[(nseg - 1) & 1], a->rstream));
    }
    for (int sgi = nseg - 1; sgi >= 0; --sgi) {
        const int b = sgi & 1;
        if (sgi > 0) {
            // half (sgi-1)&1 last held segment sgi+1, if that exists
            if (sgi + 1 < nseg) NBH_HIP(hipStreamWaitEvent(a->rstream, a->ev_free[(sgi - 1) & 1], 0));
            if (int rc = adj_recompute(a, sgi - 1, pbuf(sgi - 1))) return failed(rc);
            NBH_HIP(hipEventRecord(a->ev_ready[(sgi - 1) & 1], a->rstream));
        }
        NBH_HIP(hipStreamWaitEvent(c->stream, a->ev_ready[b], 0));
        const int s0 = sgi * a->every;
        const int s1 = std::min(a->nsteps, s0 + a->every);
        for (int t = s1 - 1; t >= s0; --t) {
            // the first half of step t's adjoint: the last step's here, the
            // others' fused into the second half of step t + 1's
            if (t == a->nsteps - 1) {
                nbh::k_adj_post<<<nbh::oblk(n), nbh::kOB, 0, c->stream>>>(n, dt, cot, a->cf.d());
                NBH_HIP(hipGetLastError());
            }
            const double* p = pbuf(sgi) + (size_t)(t - s0) * 7 * n;
            NBH_HIP(nbh::launch_force_vjp(nbh::state_of(const_cast<double*>(p), n), n, c->cfg.eps2, a->cf.d(),
                                          nullptr, 0, a->partial.d(), splits, c->stream));
            nbh::k_adj_pre_post<<<(n + nbh::kAB - 1) / nbh::kAB, 7 * nbh::kAB, 0, c->stream>>>(
                n, dt, cot, a->partial.d(), planes, a->mbar.d(), a->cf.d(), t > 0);
            NBH_HIP(hipGetLastError());
        }
        NBH_HIP(hipEventRecord(a->ev_free[b], c->stream));
    }
    // the initial forces (a0, j0) = F(x0, v0, m)
    double* x0 = a->ckpt.d();
    NBH_HIP(nbh::launch_force_vjp(nbh::state_of(x0, n), n, c->cfg.eps2, cot + 6 * (size_t)n, a->g.d(), 0,
                                  a->partial.d(), splits, c->stream));
    // grad0 = (xbar + g_x, mbar + g_m, vbar + g_v)
    std::vector<double> h((size_t)7 * n), hc((size_t)6 * n), hm(n);
    NBH_HIP(hipMemcpyAsync(h.data(), a->g.p, sz(n, 7), hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipMemcpyAsync(hc.data(), cot, sz(n, 6), hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipMemcpyAsync(hm.data(), a->mbar.p, sz(n, 1), hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            grad0[(size_t)k * n + i] = hc[(size_t)k * n + i] + h[(size_t)k * n + i];
            grad0[(size_t)(4 + k) * n + i] = hc[(size_t)(3 + k) * n + i] + h[(size_t)(4 + k) * n + i];
        }
        grad0[3 * (size_t)n + i] = hm[i] + h[3 * (size_t)n + i];
    }
    return NBH_OK;
}

}  // extern "C"
