// nbh_kernels.hip -- gfx950 kernels for the Hermite-4 force path.
//
// Pair kernels (the hot path, replacing the reference's force loops
// src/kernel.cpp:1101-1224 and the Tensix nb_compute pipeline
// kernels/nb_compute.cpp:180-582):
//   * one or two target (i) bodies per lane, BLOCK lanes per workgroup;
//   * every lane of a wave takes the same source (j) body at the same time,
//     so the source record (FP64: 64 bytes; mixed: 48, FP64 position with
//     FP32 velocity and mass) is wave-uniform.  Two ways to bring it in:
//       - LDS tiles (k_force_*_lds): each lane loads one body of a tile of
//         BLOCK with coalesced SoA loads into LDS (two tiles alternate so
//         the loads of tile t+1 overlap the arithmetic on tile t), read back
//         as broadcasts.  Best for small n: 64 lanes fetch 64 sources at
//         once, which hides the latency with few waves resident;
//       - scalar loads (k_force_*): sources packed once per evaluation into
//         records in HBM (k_pack_*) and read at wave-uniform indices into
//         SGPRs, which the VALU instructions take as operands.  No LDS, no
//         barriers, and a record's registers are scalar, so more waves stay
//         resident; FP32 reaches its rate only by pairing instructions of
//         two waves.  Best for large n;
//     both sum each target's sources in the same order (bitwise the same);
//   * grid.y splits the j-range when n is too small to fill the CUs; the
//     partial sums are added by a second kernel in split order, so results
//     are deterministic run to run;
//   * no MFMA: the softened r^-3 interaction is rsqrt-bound VALU work.
//
// Per pair (directed i <- j), both precisions:
//   d = x_j - x_i, u = v_j - v_i, s = |d|^2 + eps2, rv = d.u,
//   r = s^-1/2, a_i += m_j r^3 d, jerk_i += m_j r^3 (u - 3 rv r^2 d)
// which is the reference's per-pair formula (src/kernel.cpp:1188-1205)
// summed over all j != i instead of over the j < i triangle.
#include "nbh_internal.h"
#include "nbh_pair.h"
#include "nbh_hermite.h"

#include <algorithm>
#include <cmath>

namespace nbh {
namespace {

constexpr int kOBlock = 256;  // threads per workgroup of the O(n) kernels

struct PairArgs {
    State s;
    int n;                // targets: bodies 0..n-1
    int nj;               // sources: bodies 0..nj-1 (n except in the autotuner's trial launches)
    int jchunk;           // sources per j-split
    double eps2;
    Forces out;
    size_t split_stride;  // element offset between j-split partial planes
    int diag;             // > 0: sources are the target's own block of diag bodies
    const double* vref;   // mixed: reference velocity (3 doubles), subtracted before rounding
};

// Source range of this workgroup: its j-split, or in the block-diagonal pass
// the block that holds its targets.
template <int BLOCK, int IPL>
__device__ __forceinline__ void source_range(const PairArgs& a, int& jbeg, int& jend)
{
    if (a.diag > 0) {
        jbeg = (int)(blockIdx.x * BLOCK * IPL) / a.diag * a.diag;
        jend = min(a.n, jbeg + a.diag);
    } else {
        jbeg = blockIdx.y * a.jchunk;
        jend = min(a.nj, jbeg + a.jchunk);
    }
}

// LDS-tile kernels.  BLOCK lanes per workgroup, IPL targets per lane (lane t
// holds targets base + t + q*BLOCK, q < IPL), so each broadcast source
// record read from LDS feeds IPL pair evaluations.
template <int BLOCK, int IPL, bool MASK>
__global__ __launch_bounds__(BLOCK) void k_force_f64_lds(PairArgs a)
{
    constexpr int UNROLL = IPL == 1 ? 4 : 2;
    __shared__ JB64 tile[2][BLOCK];
    const int t = threadIdx.x;
    const int base = blockIdx.x * BLOCK * IPL + t;
    const double eps2 = a.eps2;
    double xi[IPL], yi[IPL], zi[IPL], vxi[IPL], vyi[IPL], vzi[IPL];
    double acc[IPL][6];
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = base + q * BLOCK;
        const int ic = i < a.n ? i : a.n - 1;
        xi[q] = a.s.x[ic]; yi[q] = a.s.y[ic]; zi[q] = a.s.z[ic];
        vxi[q] = a.s.vx[ic]; vyi[q] = a.s.vy[ic]; vzi[q] = a.s.vz[ic];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[q][c] = 0.0;
    }

    int jbeg, jend;
    source_range<BLOCK, IPL>(a, jbeg, jend);
    JB64 st{};
    if (jbeg + t < jend) st = load_jb64(a.s, jbeg + t);
    tile[0][t] = st;
    __syncthreads();
    int buf = 0;
    for (int jt = jbeg; jt < jend; jt += BLOCK) {
        const int cnt = min(BLOCK, jend - jt);
        const bool more = jt + BLOCK < jend;
        if (more && jt + BLOCK + t < jend) st = load_jb64(a.s, jt + BLOCK + t);
        const JB64* tb = tile[buf];
        if (cnt == BLOCK) {
#pragma unroll UNROLL
            for (int k = 0; k < BLOCK; ++k) {
                const JB64 b = tb[k];
#pragma unroll
                for (int q = 0; q < IPL; ++q)
                    pair_f64<MASK>(b, jt + k, base + q * BLOCK, xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q],
                                   eps2, acc[q]);
            }
        } else {
            for (int k = 0; k < cnt; ++k) {
                const JB64 b = tb[k];
#pragma unroll
                for (int q = 0; q < IPL; ++q)
                    pair_f64<MASK>(b, jt + k, base + q * BLOCK, xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q],
                                   eps2, acc[q]);
            }
        }
        if (more) tile[buf ^ 1][t] = st;
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = base + q * BLOCK;
        if (i < a.n) {
            const size_t o = (size_t)blockIdx.y * a.split_stride + i;
#pragma unroll
            for (int c = 0; c < 6; ++c) a.out.c[c][o] = acc[q][c];
        }
    }
}

// Terms summed over runs of kFold sources from the split's first source (a
// tile holds whole runs), each run folded into the (hi, lo) totals.
template <int BLOCK, int IPL, bool MASK>
__global__ __launch_bounds__(BLOCK) void k_force_mixed_lds(PairArgs a)
{
    static_assert(BLOCK % kFold == 0, "tile must hold whole fold runs");
    constexpr int UNROLL = IPL == 1 ? 4 : 2;
    __shared__ JBMX tile[2][BLOCK];
    const int t = threadIdx.x;
    const int base = blockIdx.x * BLOCK * IPL + t;
    const float eps2 = (float)a.eps2;
    const VRef vr = load_vref(a.vref);
    double xi[IPL], yi[IPL], zi[IPL];
    float vxi[IPL], vyi[IPL], vzi[IPL];
    float hi[IPL][6], lo[IPL][6];
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = base + q * BLOCK;
        const int ic = i < a.n ? i : a.n - 1;
        xi[q] = a.s.x[ic]; yi[q] = a.s.y[ic]; zi[q] = a.s.z[ic];
        vxi[q] = (float)(a.s.vx[ic] - vr.x);
        vyi[q] = (float)(a.s.vy[ic] - vr.y);
        vzi[q] = (float)(a.s.vz[ic] - vr.z);
#pragma unroll
        for (int c = 0; c < 6; ++c) hi[q][c] = lo[q][c] = 0.f;
    }

    int jbeg, jend;
    source_range<BLOCK, IPL>(a, jbeg, jend);
    JBMX st{};
    if (jbeg + t < jend) st = load_jbmx(a.s, jbeg + t, vr);
    tile[0][t] = st;
    __syncthreads();
    int buf = 0;
    for (int jt = jbeg; jt < jend; jt += BLOCK) {
        const int cnt = min(BLOCK, jend - jt);
        const bool more = jt + BLOCK < jend;
        if (more && jt + BLOCK + t < jend) st = load_jbmx(a.s, jt + BLOCK + t, vr);
        const JBMX* tb = tile[buf];
        for (int k0 = 0; k0 < cnt; k0 += kFold) {
            float acc[IPL][6];
#pragma unroll
            for (int q = 0; q < IPL; ++q)
#pragma unroll
                for (int c = 0; c < 6; ++c) acc[q][c] = 0.f;
            if (k0 + kFold <= cnt) {
#pragma unroll UNROLL
                for (int k = 0; k < kFold; ++k) {
                    const JBMX b = tb[k0 + k];
#pragma unroll
                    for (int q = 0; q < IPL; ++q)
                        pair_mx<MASK>(b, jt + k0 + k, base + q * BLOCK, xi[q], yi[q], zi[q], vxi[q], vyi[q],
                                      vzi[q], eps2, acc[q]);
                }
            } else {
                for (int k = k0; k < cnt; ++k) {
                    const JBMX b = tb[k];
#pragma unroll
                    for (int q = 0; q < IPL; ++q)
                        pair_mx<MASK>(b, jt + k, base + q * BLOCK, xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q],
                                      eps2, acc[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < IPL; ++q)
#pragma unroll
                for (int c = 0; c < 6; ++c) two_sum_fold(hi[q][c], lo[q][c], acc[q][c]);
        }
        if (more) tile[buf ^ 1][t] = st;
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = base + q * BLOCK;
        if (i < a.n) {
            const size_t o = (size_t)blockIdx.y * a.split_stride + i;
#pragma unroll
            for (int c = 0; c < 6; ++c) a.out.c[c][o] = (double)hi[q][c] + (double)lo[q][c];
        }
    }
}

// Source records of both precisions, packed from the SoA state before each
// evaluation.
__global__ __launch_bounds__(kOBlock) void k_pack_f64(State s, int n, JB64* __restrict__ rec)
{
    const int j = blockIdx.x * kOBlock + threadIdx.x;
    if (j < n) rec[j] = load_jb64(s, j);
}

__global__ __launch_bounds__(kOBlock) void k_pack_mixed(State s, int n, const double* __restrict__ vref,
                                                        JBMX* __restrict__ rec)
{
    const int j = blockIdx.x * kOBlock + threadIdx.x;
    if (j < n) rec[j] = load_jbmx(s, j, load_vref(vref));
}

// Centre-of-mass velocity, first pass: workgroup g sums m, m vx, m vy, m vz
// over its contiguous share of the bodies (tree order in LDS) into
// work[4g .. 4g+3].
__global__ __launch_bounds__(kOBlock) void k_vref_part(State s, int n, double* __restrict__ work)
{
    __shared__ double red[4][kOBlock];
    const int t = threadIdx.x;
    const int chunk = (n + gridDim.x - 1) / gridDim.x;
    const int j0 = blockIdx.x * chunk, j1 = min(n, j0 + chunk);
    double m = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    for (int j = j0 + t; j < j1; j += kOBlock) {
        const double mj = s.m[j];
        m += mj;
        px = fma(mj, s.vx[j], px);
        py = fma(mj, s.vy[j], py);
        pz = fma(mj, s.vz[j], pz);
    }
    red[0][t] = m; red[1][t] = px; red[2][t] = py; red[3][t] = pz;
    __syncthreads();
    for (int h = kOBlock / 2; h > 0; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int c = 0; c < 4; ++c) red[c][t] += red[c][t + h];
        __syncthreads();
    }
    if (t < 4) work[4 * blockIdx.x + t] = red[t][0];
}

// Second pass: the parts added in tree order; vref = sum m v / sum m, or 0
// when the bodies carry no mass.
__global__ __launch_bounds__(kVrefParts) void k_vref_final(double* __restrict__ work, int parts)
{
    __shared__ double red[4][kVrefParts];
    const int t = threadIdx.x;
#pragma unroll
    for (int c = 0; c < 4; ++c) red[c][t] = t < parts ? work[4 * t + c] : 0.0;
    __syncthreads();
    for (int h = kVrefParts / 2; h > 0; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int c = 0; c < 4; ++c) red[c][t] += red[c][t + h];
        __syncthreads();
    }
    if (t < 3) {
        const double m = red[0][0];
        work[vref_offset() + t] = m > 0.0 ? red[t + 1][0] / m : 0.0;
    }
}

// Scalar-load kernels.  BLOCK lanes per workgroup, IPL targets per lane
// (lane t holds targets base + t + q*BLOCK, q < IPL), so each source record
// (from `rec`, at a wave-uniform index: scalar loads) feeds IPL pair
// evaluations.  Each target
// sums its sources in order, whatever BLOCK and IPL.
template <int BLOCK, int IPL, bool MASK>
__global__ __launch_bounds__(BLOCK) void k_force_f64(PairArgs a, const JB64* __restrict__ rec)
{
    constexpr int UNROLL = 4;
    const int t = threadIdx.x;
    const int base = blockIdx.x * BLOCK * IPL + t;
    const double eps2 = a.eps2;
    double xi[IPL], yi[IPL], zi[IPL], vxi[IPL], vyi[IPL], vzi[IPL];
    double acc[IPL][6];
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = base + q * BLOCK;
        const int ic = i < a.n ? i : a.n - 1;
        xi[q] = a.s.x[ic]; yi[q] = a.s.y[ic]; zi[q] = a.s.z[ic];
        vxi[q] = a.s.vx[ic]; vyi[q] = a.s.vy[ic]; vzi[q] = a.s.vz[ic];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[q][c] = 0.0;
    }

    int jbeg, jend;
    source_range<BLOCK, IPL>(a, jbeg, jend);
#pragma unroll UNROLL
    for (int j = jbeg; j < jend; ++j) {
        const JB64 b = rec[j];
#pragma unroll
        for (int q = 0; q < IPL; ++q)
            pair_f64<MASK>(b, j, base + q * BLOCK, xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q], eps2, acc[q]);
    }
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = base + q * BLOCK;
        if (i < a.n) {
            const size_t o = (size_t)blockIdx.y * a.split_stride + i;
#pragma unroll
            for (int c = 0; c < 6; ++c) a.out.c[c][o] = acc[q][c];
        }
    }
}

// Sources from `rec` (k_pack_mixed) at wave-uniform indices: scalar loads.
// Terms are summed over runs of kFold sources starting at the split's first
// source, whatever BLOCK and IPL, then folded into the (hi, lo) totals.
template <int BLOCK, int IPL, bool MASK>
__global__ __launch_bounds__(BLOCK) void k_force_mixed(PairArgs a, const JBMX* __restrict__ rec)
{
    // one target per lane: 16 sources per unrolled body (54 VGPRs, as with
    // 8, and 0.7% faster at n = 2^20)
    constexpr int UNROLL = IPL == 1 ? 16 : 4;
    const int t = threadIdx.x;
    const int base = blockIdx.x * BLOCK * IPL + t;
    const float eps2 = (float)a.eps2;
    const VRef vr = load_vref(a.vref);
    double xi[IPL], yi[IPL], zi[IPL];
    float vxi[IPL], vyi[IPL], vzi[IPL];
    float hi[IPL][6], lo[IPL][6];
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = base + q * BLOCK;
        const int ic = i < a.n ? i : a.n - 1;
        xi[q] = a.s.x[ic]; yi[q] = a.s.y[ic]; zi[q] = a.s.z[ic];
        vxi[q] = (float)(a.s.vx[ic] - vr.x);
        vyi[q] = (float)(a.s.vy[ic] - vr.y);
        vzi[q] = (float)(a.s.vz[ic] - vr.z);
#pragma unroll
        for (int c = 0; c < 6; ++c) hi[q][c] = lo[q][c] = 0.f;
    }

    int jbeg, jend;
    source_range<BLOCK, IPL>(a, jbeg, jend);
    for (int j0 = jbeg; j0 < jend; j0 += kFold) {
        float acc[IPL][6];
#pragma unroll
        for (int q = 0; q < IPL; ++q)
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[q][c] = 0.f;
        if (j0 + kFold <= jend) {
#pragma unroll UNROLL
            for (int j = j0; j < j0 + kFold; ++j) {
                const JBMX b = rec[j];
#pragma unroll
                for (int q = 0; q < IPL; ++q)
                    pair_mx<MASK>(b, j, base + q * BLOCK, xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q], eps2,
                                  acc[q]);
            }
        } else {
            for (int j = j0; j < jend; ++j) {
                const JBMX b = rec[j];
#pragma unroll
                for (int q = 0; q < IPL; ++q)
                    pair_mx<MASK>(b, j, base + q * BLOCK, xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q], eps2,
                                  acc[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < IPL; ++q)
#pragma unroll
            for (int c = 0; c < 6; ++c) two_sum_fold(hi[q][c], lo[q][c], acc[q][c]);
    }
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = base + q * BLOCK;
        if (i < a.n) {
            const size_t o = (size_t)blockIdx.y * a.split_stride + i;
#pragma unroll
            for (int c = 0; c < 6; ++c) a.out.c[c][o] = (double)hi[q][c] + (double)lo[q][c];
        }
    }
}

// Second pass of a j-split: out[c][i] = sum over splits in order 0..S-1.
__global__ __launch_bounds__(kOBlock) void k_reduce_splits(const double* __restrict__ part,
                                                           int n, int splits, Forces out)
{
    const int i = blockIdx.x * kOBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    const double* p = part + (size_t)c * splits * n + i;
    double v = 0.0;
    // unrolled so that eight loads are in flight at once; the adds keep split order
#pragma unroll 8
    for (int s = 0; s < splits; ++s) v += p[(size_t)s * n];
    out.c[c][i] = v;
}

// ------------------------------------------------------- integrator -----
// Predictor, src/nbody.cpp:152-163, same operations in the same order and
// no contraction, so results are bitwise those of the reference's x86-64
// build (which has no FMA).
__global__ __launch_bounds__(kOBlock) void k_predict(int n, const double* __restrict__ dtp,
                                                     double* __restrict__ x, double* __restrict__ y,
                                                     double* __restrict__ z, double* __restrict__ vx,
                                                     double* __restrict__ vy, double* __restrict__ vz,
                                                     Forces f)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * kOBlock + threadIdx.x;
    if (i >= n) return;
    const double dt = *dtp;
    double px = x[i], py = y[i], pz = z[i], ux = vx[i], uy = vy[i], uz = vz[i];
    hermite_predict_axis(dt, f.c[0][i], f.c[3][i], px, ux);
    hermite_predict_axis(dt, f.c[1][i], f.c[4][i], py, uy);
    hermite_predict_axis(dt, f.c[2][i], f.c[5][i], pz, uz);
    x[i] = px; y[i] = py; z[i] = pz;
    vx[i] = ux; vy[i] = uy; vz[i] = uz;
}

__device__ __forceinline__ void block_min_u64(unsigned long long v, unsigned long long* out)
{
    __shared__ unsigned long long red[kOBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_down(v, off, 64);
        v = o < v ? o : v;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = red[0];
        for (int k = 1; k < kOBlock / 64; ++k) m = red[k] < m ? red[k] : m;
        atomicMin(out, m);
    }
}

__device__ __forceinline__ unsigned long long dt_bits(double d)
{
    // Non-negative doubles order like their bit patterns; NaN sorts last.
    return (d >= 0.0) ? (unsigned long long)__double_as_longlong(d) : ~0ull;
}

// Snap/crackle and corrector, src/nbody.cpp:165-196 (no contraction, same
// order), then the pointer swap of :263-288 done as a copy f0 <- f1 so the
// captured step sequence keeps fixed addresses.  With dt_next set, the
// Aarseth criterion of NBody::updatedt (src/nbody.cpp:198-222), which uses
// the start-of-step a0, j0 and the new a2, a3, is min-reduced.
__global__ __launch_bounds__(kOBlock) void k_correct(int n, const double* __restrict__ dtp,
                                                     double* __restrict__ x, double* __restrict__ y,
                                                     double* __restrict__ z, double* __restrict__ vx,
                                                     double* __restrict__ vy, double* __restrict__ vz,
                                                     Forces f0, Forces f1,
                                                     unsigned long long* dt_next, double eta)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * kOBlock + threadIdx.x;
    unsigned long long cand = ~0ull;
    if (i < n) {
        const double dt = *dtp;
        const double invdt2 = 1.0 / (dt * dt);
        const double invdt3 = 1.0 / (dt * dt * dt);
        double a2[3], a3[3], a0[3], j0[3];
        double* const pos[3] = {x, y, z};
        double* const vel[3] = {vx, vy, vz};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a0[c] = f0.c[c][i];
            j0[c] = f0.c[3 + c][i];
            double p = pos[c][i], u = vel[c][i];
            hermite_correct_axis(dt, invdt2, invdt3, a0[c], j0[c], f1.c[c][i], f1.c[3 + c][i], p, u, a2[c], a3[c]);
            pos[c][i] = p;
            vel[c][i] = u;
        }
        if (dt_next) {
            const double na2 = a0[0] * a0[0] + a0[1] * a0[1] + a0[2] * a0[2];
            const double na = sqrt(na2);
            const double nj2 = j0[0] * j0[0] + j0[1] * j0[1] + j0[2] * j0[2];
            const double nj = sqrt(nj2);
            const double ns2 = a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2];
            const double ns = sqrt(ns2);
            const double nc2 = a3[0] * a3[0] + a3[1] * a3[1] + a3[2] * a3[2];
            const double nc = sqrt(nc2);
            const double num = na * ns + nj2;
            const double den = nj * nc + ns2;
            cand = dt_bits(sqrt(eta * num / den));
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) f0.c[c][i] = f1.c[c][i];
    }
    if (dt_next) block_min_u64(cand, dt_next);
}

__global__ void k_dt_advance(double* dt, unsigned long long* dt_next)
{
    if (threadIdx.x == 0) {
        *dt = __longlong_as_double((long long)*dt_next);
        *dt_next = (unsigned long long)__double_as_longlong(INFINITY);
    }
}

__global__ void k_fill_inf(unsigned long long* p)
{
    if (threadIdx.x == 0) *p = (unsigned long long)__double_as_longlong(INFINITY);
}

// NBody::initialdt, src/nbody.cpp:131-150.
__global__ __launch_bounds__(kOBlock) void k_initial_dt(int n, Forces f, double eta0,
                                                        unsigned long long* out)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * kOBlock + threadIdx.x;
    unsigned long long cand = ~0ull;
    if (i < n) {
        const double a2x = f.c[0][i] * f.c[0][i];
        const double a2y = f.c[1][i] * f.c[1][i];
        const double a2z = f.c[2][i] * f.c[2][i];
        const double an = sqrt(a2x + a2y + a2z);
        const double j2x = f.c[3][i] * f.c[3][i];
        const double j2y = f.c[4][i] * f.c[4][i];
        const double j2z = f.c[5][i] * f.c[5][i];
        const double jn = sqrt(j2x + j2y + j2z);
        cand = dt_bits(eta0 * an / jn);
    }
    block_min_u64(cand, out);
}

// ----------------------------------------------------------- energy -----
// Kinetic and potential energy in FP64 (src/utils.cpp:125-153), each
// unordered pair evaluated once, as the reference's j > i loop does.
// compat = 0: U = -sum_{i<j} m_i m_j / sqrt(r^2 + eps2).  compat = 1: the
// reference's own definition, U = -sum_i sum_{j>i} m_j^2 / sqrt(r^2 + eps2)
// (src/utils.cpp:148), i.e. each pair weighted by the squared mass of its
// higher-indexed body.
//
// Workgroup b holds targets b*kEBlock .. and visits the source tiles
// b + k (mod nb), k = 0 .. (nb-1)/2, plus k = nb/2 when nb is even and
// b < nb/2: tile pairs are taken once each and every workgroup does about
// nb/2 tiles.  In its own tile (k = 0) a target takes the sources j > i.
// Block sums are written per block and added in block order by
// k_energy_final: deterministic.
constexpr int kEBlock = 256;

template <bool COMPAT>
__global__ __launch_bounds__(kEBlock) void k_energy(State s, int n, double eps2, double* work)
{
    __shared__ double4 tile[kEBlock];
    __shared__ double red[2][kEBlock / 64];
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int nb = gridDim.x;
    const int i = b * kEBlock + t;
    const int ic = i < n ? i : n - 1;
    const double xi = s.x[ic], yi = s.y[ic], zi = s.z[ic];
    const int kmax = (nb - 1) / 2 + ((nb % 2 == 0 && b < nb / 2) ? 1 : 0);
    double phi = 0.0;  // sum over sources of w_j / r: w_j = m_j, or m_j^2 (compat, j > i)
    double rin = 0.0;  // compat: sum of 1 / r over sources j < i (weighted by m_i^2 at the end)
    for (int k = 0; k <= kmax; ++k) {
        const int q = b + k < nb ? b + k : b + k - nb;
        const int jt = q * kEBlock;
        const int j = jt + t;
        double4 src = make_double4(0.0, 0.0, 0.0, 0.0);
        if (j < n) src = make_double4(s.x[j], s.y[j], s.z[j], COMPAT ? s.m[j] * s.m[j] : s.m[j]);
        __syncthreads();
        tile[t] = src;
        __syncthreads();
        const int cnt = min(kEBlock, n - jt);
        if (k == 0) {
            for (int kk = 0; kk < cnt; ++kk) {
                const double4 p = tile[kk];
                const double dx = p.x - xi, dy = p.y - yi, dz = p.z - zi;
                double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2)));
                const bool keep = jt + kk > i;
                r2 = keep ? r2 : 1.0;
                phi = fma(keep ? p.w : 0.0, rsqrt_f64(r2), phi);
            }
        } else if (COMPAT && q < b) {  // wrapped: every source has a lower index
            for (int kk = 0; kk < cnt; ++kk) {
                const double4 p = tile[kk];
                const double dx = p.x - xi, dy = p.y - yi, dz = p.z - zi;
                rin += rsqrt_f64(fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2))));
            }
        } else {
            for (int kk = 0; kk < cnt; ++kk) {
                const double4 p = tile[kk];
                const double dx = p.x - xi, dy = p.y - yi, dz = p.z - zi;
                phi = fma(p.w, rsqrt_f64(fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2)))), phi);
            }
        }
    }
    double pot = 0.0, kin = 0.0;
    if (i < n) {
        const double mi = s.m[i];
        pot = COMPAT ? -fma(mi * mi, rin, phi) : -mi * phi;
        const double ux = s.vx[i], uy = s.vy[i], uz = s.vz[i];
        kin = 0.5 * mi * (ux * ux + uy * uy + uz * uz);
    }
    for (int off = 32; off > 0; off >>= 1) {
        pot += __shfl_down(pot, off, 64);
        kin += __shfl_down(kin, off, 64);
    }
    if ((t & 63) == 0) {
        red[0][t >> 6] = kin;
        red[1][t >> 6] = pot;
    }
    __syncthreads();
    if (t == 0) {
        double k = 0.0, p = 0.0;
        for (int w = 0; w < kEBlock / 64; ++w) {
            k += red[0][w];
            p += red[1][w];
        }
        work[2 * b] = k;
        work[2 * b + 1] = p;
    }
}

__global__ __launch_bounds__(256) void k_energy_final(const double* work, int blocks, double* out)
{
    __shared__ double rk[256], rp[256];
    const int t = threadIdx.x;
    double k = 0.0, p = 0.0;
    for (int b = t; b < blocks; b += 256) {
        k += work[2 * b];
        p += work[2 * b + 1];
    }
    rk[t] = k;
    rp[t] = p;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            rk[t] += rk[t + h];
            rp[t] += rp[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = rk[0];
        out[1] = rp[0];
    }
}

__global__ __launch_bounds__(kOBlock) void k_check_finite(int n, Forces f, int* flag)
{
    const int i = blockIdx.x * kOBlock + threadIdx.x;
    if (i >= n) return;
    bool bad = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) bad |= !isfinite(f.c[c][i]);
    if (bad) atomicOr(flag, 1);
}

inline int oblocks(int n) { return (n + kOBlock - 1) / kOBlock; }

template <int BLOCK, int IPL>
hipError_t launch_pair(int precision, bool mask, const PairArgs& a, int jsplit, const void* rec, hipStream_t st)
{
    const dim3 grid((a.n + BLOCK * IPL - 1) / (BLOCK * IPL), jsplit);
    if (!rec) {  // LDS tiles
        if (precision == 0) {
            if (mask)
                k_force_f64_lds<BLOCK, IPL, true><<<grid, BLOCK, 0, st>>>(a);
            else
                k_force_f64_lds<BLOCK, IPL, false><<<grid, BLOCK, 0, st>>>(a);
        } else {
            if (mask)
                k_force_mixed_lds<BLOCK, IPL, true><<<grid, BLOCK, 0, st>>>(a);
            else
                k_force_mixed_lds<BLOCK, IPL, false><<<grid, BLOCK, 0, st>>>(a);
        }
        return hipGetLastError();
    }
    if (precision == 0) {
        const JB64* r = static_cast<const JB64*>(rec);
        if (mask)
            k_force_f64<BLOCK, IPL, true><<<grid, BLOCK, 0, st>>>(a, r);
        else
            k_force_f64<BLOCK, IPL, false><<<grid, BLOCK, 0, st>>>(a, r);
    } else {
        const JBMX* r = static_cast<const JBMX*>(rec);
        if (mask)
            k_force_mixed<BLOCK, IPL, true><<<grid, BLOCK, 0, st>>>(a, r);
        else
            k_force_mixed<BLOCK, IPL, false><<<grid, BLOCK, 0, st>>>(a, r);
    }
    return hipGetLastError();
}


}  // namespace

size_t source_record_bytes(int precision, int n)
{
    return (precision == 0 ? sizeof(JB64) : sizeof(JBMX)) * (size_t)n;
}

hipError_t pack_source_records(int precision, const State& s, int n, const double* vref, void* rec,
                               hipStream_t st)
{
    if (!rec || (precision != 0 && !vref)) return hipErrorInvalidValue;
    if (precision == 0)
        k_pack_f64<<<(n + kOBlock - 1) / kOBlock, kOBlock, 0, st>>>(s, n, static_cast<JB64*>(rec));
    else
        k_pack_mixed<<<(n + kOBlock - 1) / kOBlock, kOBlock, 0, st>>>(s, n, vref, static_cast<JBMX*>(rec));
    return hipGetLastError();
}

hipError_t launch_vref(const State& s, int n, double* work, hipStream_t st)
{
    if (!work || n < 1) return hipErrorInvalidValue;
    const int parts = std::min(kVrefParts, oblocks(n));
    k_vref_part<<<parts, kOBlock, 0, st>>>(s, n, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_vref_final<<<1, kVrefParts, 0, st>>>(work, parts);
    return hipGetLastError();
}

hipError_t launch_force(int precision, const State& s, int n, double eps2,
                        const Shape& shape, const Forces& out, double* partial,
                        void* rec, double* vwork, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1, int nj,
                        bool reduce_splits)
{
    if (precision != 0 && !vwork) return hipErrorInvalidValue;
    PairArgs a;
    a.s = s;
    a.n = n;
    a.nj = nj > 0 ? nj : n;
    a.jchunk = shape.jchunk;
    a.eps2 = eps2;
    a.diag = 0;
    a.vref = precision != 0 ? vwork + vref_offset() : nullptr;
    hipError_t e;
    if (ev0 && (e = hipEventRecord(ev0, st)) != hipSuccess) return e;
    if (precision != 0 && (e = launch_vref(s, n, vwork, st)) != hipSuccess) return e;
    if (shape.sym_chunk > 0) {
        if (nj > 0) return hipErrorInvalidValue;
        return launch_force_sym(precision, s, n, eps2, shape.sym_chunk, out, a.vref, st, nullptr, ev1);
    }
    if (shape.jsplit > 1) {
        // partial plane for (component c, split k) starts at (c*S + k)*n
        for (int c = 0; c < 6; ++c) a.out.c[c] = partial + (size_t)c * shape.jsplit * n;
        a.split_stride = (size_t)n;
    } else {
        a.out = out;
        a.split_stride = 0;
    }
    const bool mask = self_mask(precision, eps2);
    const void* r = shape.scalar ? rec : nullptr;
    if (shape.scalar && (e = pack_source_records(precision, s, a.nj, a.vref, rec, st)) != hipSuccess) return e;
    const int key = shape.block * 10 + shape.ipl;
    switch (key) {
    case 641: e = launch_pair<64, 1>(precision, mask, a, shape.jsplit, r, st); break;
    case 642: e = launch_pair<64, 2>(precision, mask, a, shape.jsplit, r, st); break;
    case 1281: e = launch_pair<128, 1>(precision, mask, a, shape.jsplit, r, st); break;
    case 1282: e = launch_pair<128, 2>(precision, mask, a, shape.jsplit, r, st); break;
    case 2561: e = launch_pair<256, 1>(precision, mask, a, shape.jsplit, r, st); break;
    case 2562: e = launch_pair<256, 2>(precision, mask, a, shape.jsplit, r, st); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, st)) != hipSuccess) return e;
    if (shape.jsplit > 1 && reduce_splits) {
        k_reduce_splits<<<dim3(oblocks(n), 6), kOBlock, 0, st>>>(partial, n, shape.jsplit, out);
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_force_blockdiag(int precision, const State& s, int n, double eps2, int B,
                                  const Forces& out, const double* vref, hipStream_t st)
{
    constexpr int BLOCK = 256;
    if (B <= 0 || B % BLOCK != 0 || (precision != 0 && !vref)) return hipErrorInvalidValue;
    PairArgs a;
    a.s = s;
    a.n = n;
    a.nj = n;
    a.jchunk = B;
    a.eps2 = eps2;
    a.out = out;
    a.split_stride = 0;
    a.diag = B;
    a.vref = vref;
    return launch_pair<BLOCK, 1>(precision, self_mask(precision, eps2), a, 1, nullptr, st);  // LDS tiles
}

hipError_t launch_predict(int n, const double* dt, double* x, double* y, double* z,
                          double* vx, double* vy, double* vz, const Forces& f0,
                          hipStream_t st)
{
    k_predict<<<oblocks(n), kOBlock, 0, st>>>(n, dt, x, y, z, vx, vy, vz, f0);
    return hipGetLastError();
}

hipError_t launch_correct(int n, const double* dt, double* x, double* y, double* z,
                          double* vx, double* vy, double* vz, const Forces& f0,
                          const Forces& f1, unsigned long long* dt_next, double eta,
                          hipStream_t st)
{
    k_correct<<<oblocks(n), kOBlock, 0, st>>>(n, dt, x, y, z, vx, vy, vz, f0, f1, dt_next, eta);
    return hipGetLastError();
}

hipError_t launch_dt_advance(double* dt, unsigned long long* dt_next, hipStream_t st)
{
    k_dt_advance<<<1, 64, 0, st>>>(dt, dt_next);
    return hipGetLastError();
}

hipError_t launch_initial_dt(int n, const Forces& f, double eta0, unsigned long long* out,
                             hipStream_t st)
{
    k_initial_dt<<<oblocks(n), kOBlock, 0, st>>>(n, f, eta0, out);
    return hipGetLastError();
}

hipError_t launch_fill_inf(unsigned long long* p, hipStream_t st)
{
    k_fill_inf<<<1, 64, 0, st>>>(p);
    return hipGetLastError();
}

int energy_blocks(int n) { return (n + kEBlock - 1) / kEBlock; }

hipError_t launch_energy(const State& s, int n, double eps2, int compat, double* work,
                         double* out, hipStream_t st)
{
    const int nb = energy_blocks(n);
    if (compat)
        k_energy<true><<<nb, kEBlock, 0, st>>>(s, n, eps2, work);
    else
        k_energy<false><<<nb, kEBlock, 0, st>>>(s, n, eps2, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_energy_final<<<1, 256, 0, st>>>(work, nb, out);
    return hipGetLastError();
}

hipError_t launch_check_finite(int n, const Forces& f, int* flag, hipStream_t st)
{
    k_check_finite<<<oblocks(n), kOBlock, 0, st>>>(n, f, flag);
    return hipGetLastError();
}

}  // namespace nbh
