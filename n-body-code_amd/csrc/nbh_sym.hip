// nbh_sym.hip -- pair kernels that evaluate each unordered pair once and
// apply it to both bodies (Newton's third law), as the reference's CPU loops
// do with their j < i triangle (src/kernel.cpp:1126 for _F32, :1187 for
// FP64).  Two precisions share the schedule: mixed (FP64 position
// differences rounded to FP32, FP32 from there on) and FP64 (the directed
// FP64 kernel's arithmetic).
//
// Layout of the work:
//   * bodies are cut into nr = n / B full blocks of B = 64*K*W (K targets per
//     lane, W waves per workgroup) and, when B does not divide n, a partial
//     block nr of the n - nr*B remaining bodies.  Full block P is the target
//     block of the pairs with its source list: blocks P+1 .. P+H (mod nr),
//     H = (nr-1)/2, plus P+nr/2 for P < nr/2 when nr is even, then the
//     partial block.  Every unordered block pair is taken exactly once and
//     every full block has the same work; the partial block is only ever a
//     source, and its tiles past n are not visited (the last visited one is
//     padded with massless bodies far away);
//   * a workgroup takes `chunk` consecutive entries of one source list; the
//     ensemble drives the same body from a list of (system, block, entry
//     range) items (k_ens_sym);
//   * pairs inside a block are evaluated one-sided by the directed kernel
//     (launch_force_blockdiag), which also initialises the outputs; these
//     kernels then add their contributions with FP64 atomics;
//   * source bodies are staged in LDS in tiles of 64 = 4 groups of 16.  Each
//     group is stored twice in a row (32 records), so at step r the lane at
//     position p of its 16-lane row reads record p + r: the rows pair their
//     lanes with the group's 16 sources in rotated order, reading distinct
//     records at a compile-time offset;
//   * the reaction on the source (minus m_i r^3 times the same d and
//     u - 3 (d.u) r^2 d) is summed over the K targets of the lane, then
//     added to a per-lane accumulator after a DPP row rotation (row_ror:r)
//     that brings it to the lane holding that source: one DPP add per
//     component per step and K pairs (FP32: the rotation folds into the add;
//     FP64: two 32-bit DPP moves and an add).  After 16 steps the four rows
//     are summed with cross-lane shuffles, the waves' sums are combined in
//     LDS in wave order, and the tile's 64 x 6 source totals are added to the
//     outputs with FP64 atomics;
//   * target sums run over the sources of FL consecutive tiles and are then
//     added to the outputs with FP64 atomics (the single-system kernel, FL =
//     4 mixed, 8 FP64: no totals held in registers, which leaves room for
//     three targets per lane at five waves per SIMD in mixed), or over one tile and then into
//     per-lane totals that go out once the workgroup's block pairs are done
//     (the ensemble's kernel, FL = 0; mixed: FP32 sums folded into an FP32
//     (hi, lo) pair with Knuth's two-sum).
//
// The FP64 atomic adds make the summation order depend on scheduling: two
// runs agree to rounding of FP64 sums (about 1e-16 relative), not bit for
// bit.  The directed kernels remain the deterministic choice.
#include "nbh_internal.h"
#include "nbh_pair.h"

#include <utility>

namespace nbh {
namespace {

constexpr int kSymTile = 64;  // sources per LDS tile (4 groups of 16)

// Source records past the last body (in the last tile of the partial block)
// hold pad bodies: zero mass, at rest, 1e18 away along -x, each at its own
// distance.  Every pair with a pad then has r^3 m = 0 exactly and all
// intermediates finite (s ~ 1e36 fits FP32), so pads add nothing to real
// bodies; their own sums are never written.
__device__ __forceinline__ double pad_x(int j)
{
    return -1e18 * (1.0 + 1e-3 * (j & 1023));
}

// FP64 atomic add to element idx of one output array.  The base is a kernel
// argument (wave-uniform) and the offset a 32-bit byte count, so the address
// is an SGPR pair plus one VGPR: no 64-bit address registers are kept across
// the step loop (they were spilled, and each reload waited for every atomic
// in flight).  Needs n <= 2^29 (sym_supported).
__device__ __forceinline__ void out_add(double* base, unsigned idx, double v)
{
    unsigned off = idx * 8u;
    // opaque to the optimiser: the address is formed here, at the add, and
    // not hoisted out of the loops into 64-bit registers
    asm volatile("" : "+v"(off));
    unsafeAtomicAdd(reinterpret_cast<double*>(reinterpret_cast<char*>(base) + off), v);
}

// Lane index within the wave, recomputed where it is used (two VALU
// instructions): the compiler cannot fold it into a value computed before
// the step loop and keep that value live, or spilled, across the loop.
__device__ __forceinline__ int lane_now()
{
    int l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

// Row rotation by R lanes: lane k of each 16-lane row receives the value of
// lane (k - R) mod 16 of the same row.
template <int R>
__device__ __forceinline__ float row_rot(float v)
{
    if constexpr (R == 0)
        return v;
    else
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + R, 0xF, 0xF, false));
}

template <int R>
__device__ __forceinline__ double row_rot(double v)
{
    if constexpr (R == 0) {
        return v;
    } else {
        const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x120 + R, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x120 + R, 0xF, 0xF, false);
        return __hiloint2double(hi, lo);
    }
}

// Mixed precision: FP64 position differences rounded to FP32, FP32 from
// there on.  Target sums: FP32 runs of kFlushTiles tiles added to the
// outputs in FP64 (one system), or per-tile FP32 sums folded into a two-sum
// (hi, lo) pair (the ensemble's kernel).
template <int K>
struct SymMixed {
    static constexpr int kWaves = 5;
    // tiles per flush of the target sums in the single-system kernel: FP32
    // runs of 256 sources between FP64 atomic adds.  Measured at n = 2^20
    // (three targets per lane, five waves, eps2 = 1e-6) against two tiles:
    // 2.2% faster (three tiles 1.4%), the largest acceleration error against
    // FP64 7.2e-7 instead of 6.3e-7 (n = 65536: 1.3e-6 instead of 9.0e-7),
    // the median 2.25e-8 instead of 2.23e-8.  One tile ran no faster than two
    // targets per lane with two-sum totals in registers.
    static constexpr int kFlushTiles = 4;
    using Rec = JBMX;
    using Sum = float;
    using Eps = float;
    struct Targets {
        double x[K], y[K], z[K];
        float vx[K], vy[K], vz[K], m[K];
    };
    struct Tot {
        float hi[K][6], lo[K][6];
    };
    static __device__ __forceinline__ Rec load_rec(const State& s, int j, int n, const VRef& vr)
    {
        if (j < n) return load_jbmx(s, j, vr);
        Rec b{};
        b.x = pad_x(j);
        return b;
    }
    static __device__ __forceinline__ void load_targets(const State& s, int ibase, Targets& I, const VRef& vr)
    {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = ibase + k * 64;
            I.x[k] = s.x[i]; I.y[k] = s.y[i]; I.z[k] = s.z[i];
            I.m[k] = (float)s.m[i];
            I.vx[k] = (float)(s.vx[i] - vr.x);
            I.vy[k] = (float)(s.vy[i] - vr.y);
            I.vz[k] = (float)(s.vz[i] - vr.z);
        }
    }
    // One source against the K targets, written one operation at a time
    // across the targets so that their K dependency chains (FP64 difference,
    // convert, |d|^2, rsqrt, r^3, ...) reach the scheduler interleaved.  With
    // the chains written target after target, the compiler, held to 96 VGPRs
    // for five waves, ran the first target's chain nearly serially (one FP64
    // temporary, every instruction waiting on the last): the interleaved
    // form is 6% faster at n = 2^20 with the same 96 VGPRs.
    template <int R>
    static __device__ __forceinline__ void step(const Rec* __restrict__ row, const Targets& I, Eps eps2,
                                                Sum (&ai)[K][6], Sum (&aj)[6])
    {
        const Rec b = row[R];
        float dx[K], dy[K], dz[K], ux[K], uy[K], uz[K], sq[K], vr[K], r[K], r2[K], r3[K], al[K];
        // velocity differences first: FP32 work the scheduler can pair while
        // the FP64 differences below are in flight (1.4% at n = 2^20)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ux[k] = b.vx - I.vx[k];
            uy[k] = b.vy - I.vy[k];
            uz[k] = b.vz - I.vz[k];
        }
#pragma unroll
        for (int k = 0; k < K; ++k) dx[k] = (float)(b.x - I.x[k]);
#pragma unroll
        for (int k = 0; k < K; ++k) dy[k] = (float)(b.y - I.y[k]);
#pragma unroll
        for (int k = 0; k < K; ++k) dz[k] = (float)(b.z - I.z[k]);
#pragma unroll
        for (int k = 0; k < K; ++k) sq[k] = fmaf(dz[k], dz[k], eps2);
#pragma unroll
        for (int k = 0; k < K; ++k) sq[k] = fmaf(dy[k], dy[k], sq[k]);
#pragma unroll
        for (int k = 0; k < K; ++k) sq[k] = fmaf(dx[k], dx[k], sq[k]);
#pragma unroll
        for (int k = 0; k < K; ++k) r[k] = __builtin_amdgcn_rsqf(sq[k]);
#pragma unroll
        for (int k = 0; k < K; ++k) vr[k] = dz[k] * uz[k];
#pragma unroll
        for (int k = 0; k < K; ++k) vr[k] = fmaf(dy[k], uy[k], vr[k]);
#pragma unroll
        for (int k = 0; k < K; ++k) vr[k] = fmaf(dx[k], ux[k], vr[k]);
#pragma unroll
        for (int k = 0; k < K; ++k) r2[k] = r[k] * r[k];
#pragma unroll
        for (int k = 0; k < K; ++k) r3[k] = r[k] * r2[k];
#pragma unroll
        for (int k = 0; k < K; ++k) al[k] = (-3.0f * r2[k]) * vr[k];
        // u becomes w = u - 3 (d.u) r^2 d in place
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ux[k] = fmaf(al[k], dx[k], ux[k]);
            uy[k] = fmaf(al[k], dy[k], uy[k]);
            uz[k] = fmaf(al[k], dz[k], uz[k]);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float mj3 = b.m * r3[k];
            ai[k][0] = fmaf(mj3, dx[k], ai[k][0]);
            ai[k][1] = fmaf(mj3, dy[k], ai[k][1]);
            ai[k][2] = fmaf(mj3, dz[k], ai[k][2]);
            ai[k][3] = fmaf(mj3, ux[k], ai[k][3]);
            ai[k][4] = fmaf(mj3, uy[k], ai[k][4]);
            ai[k][5] = fmaf(mj3, uz[k], ai[k][5]);
        }
        float cs[6];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float mi3 = I.m[k] * r3[k];
            if (k == 0) {
                cs[0] = mi3 * dx[k]; cs[1] = mi3 * dy[k]; cs[2] = mi3 * dz[k];
                cs[3] = mi3 * ux[k]; cs[4] = mi3 * uy[k]; cs[5] = mi3 * uz[k];
            } else {
                cs[0] = fmaf(mi3, dx[k], cs[0]); cs[1] = fmaf(mi3, dy[k], cs[1]); cs[2] = fmaf(mi3, dz[k], cs[2]);
                cs[3] = fmaf(mi3, ux[k], cs[3]); cs[4] = fmaf(mi3, uy[k], cs[4]); cs[5] = fmaf(mi3, uz[k], cs[5]);
            }
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) aj[c] -= row_rot<R>(cs[c]);
    }
    static __device__ __forceinline__ void zero(Tot& t)
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < 6; ++c) t.hi[k][c] = t.lo[k][c] = 0.f;
    }
    static __device__ __forceinline__ void fold(Tot& t, const Sum (&ai)[K][6])
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < 6; ++c) two_sum_fold(t.hi[k][c], t.lo[k][c], ai[k][c]);
    }
    static __device__ __forceinline__ double total(const Tot& t, int k, int c)
    {
        return (double)t.hi[k][c] + (double)t.lo[k][c];
    }
};

// FP64 source record of the symmetric kernel: 80 bytes, so that the 16
// lanes of a row, reading 16 consecutive records, start on 16 different LDS
// banks (20 dwords apart; a 64-byte record would put lanes p and p + 4 on
// the same bank).  pad is never read.
struct __align__(16) JB80 {
    double x, y, z, m, vx, vy, vz, pad[3];
};

// FP64: the directed FP64 kernel's arithmetic (pair_f64) with the reaction
// added; per-tile sums added to FP64 totals.
template <int K>
struct SymF64 {
    static constexpr int kWaves = 3;
    // FP64 runs lose nothing by being long: 8 tiles per flush (512 sources)
    // were 2% faster than 2 at n = 2^20
    static constexpr int kFlushTiles = 8;
    using Rec = JB80;
    using Sum = double;
    using Eps = double;
    struct Targets {
        double x[K], y[K], z[K], vx[K], vy[K], vz[K], m[K];
    };
    struct Tot {
        double t[K][6];
    };
    static __device__ __forceinline__ Rec load_rec(const State& s, int j, int n, const VRef&)
    {
        Rec b{};
        if (j < n) {
            b.x = s.x[j]; b.y = s.y[j]; b.z = s.z[j]; b.m = s.m[j];
            b.vx = s.vx[j]; b.vy = s.vy[j]; b.vz = s.vz[j];
        } else {
            b.x = pad_x(j);
        }
        return b;
    }
    static __device__ __forceinline__ void load_targets(const State& s, int ibase, Targets& I, const VRef&)
    {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = ibase + k * 64;
            I.x[k] = s.x[i]; I.y[k] = s.y[i]; I.z[k] = s.z[i];
            I.m[k] = s.m[i];
            I.vx[k] = s.vx[i]; I.vy[k] = s.vy[i]; I.vz[k] = s.vz[i];
        }
    }
    template <int R>
    static __device__ __forceinline__ void step(const Rec* __restrict__ row, const Targets& I, Eps eps2,
                                                Sum (&ai)[K][6], Sum (&aj)[6])
    {
        const Rec b = row[R];
        double cs[6];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double dx = b.x - I.x[k];
            const double dy = b.y - I.y[k];
            const double dz = b.z - I.z[k];
            const double ux = b.vx - I.vx[k];
            const double uy = b.vy - I.vy[k];
            const double uz = b.vz - I.vz[k];
            const double s = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2)));
            const double vr = fma(dx, ux, fma(dy, uy, dz * uz));
            const double r = rsqrt_f64(s);
            const double r2 = r * r;
            const double r3 = r * r2;
            const double mj3 = b.m * r3;
            const double mi3 = I.m[k] * r3;
            const double al = (-3.0 * r2) * vr;
            const double wx = fma(al, dx, ux);
            const double wy = fma(al, dy, uy);
            const double wz = fma(al, dz, uz);
            ai[k][0] = fma(mj3, dx, ai[k][0]);
            ai[k][1] = fma(mj3, dy, ai[k][1]);
            ai[k][2] = fma(mj3, dz, ai[k][2]);
            ai[k][3] = fma(mj3, wx, ai[k][3]);
            ai[k][4] = fma(mj3, wy, ai[k][4]);
            ai[k][5] = fma(mj3, wz, ai[k][5]);
            if (k == 0) {
                cs[0] = mi3 * dx; cs[1] = mi3 * dy; cs[2] = mi3 * dz;
                cs[3] = mi3 * wx; cs[4] = mi3 * wy; cs[5] = mi3 * wz;
            } else {
                cs[0] = fma(mi3, dx, cs[0]); cs[1] = fma(mi3, dy, cs[1]); cs[2] = fma(mi3, dz, cs[2]);
                cs[3] = fma(mi3, wx, cs[3]); cs[4] = fma(mi3, wy, cs[4]); cs[5] = fma(mi3, wz, cs[5]);
            }
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) aj[c] -= row_rot<R>(cs[c]);
    }
    static __device__ __forceinline__ void zero(Tot& t)
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < 6; ++c) t.t[k][c] = 0.0;
    }
    static __device__ __forceinline__ void fold(Tot& t, const Sum (&ai)[K][6])
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < 6; ++c) t.t[k][c] += ai[k][c];
    }
    static __device__ __forceinline__ double total(const Tot& t, int k, int c) { return t.t[k][c]; }
};

template <class T, int K, int... R>
__device__ __forceinline__ void sym_group(const typename T::Rec* __restrict__ row, const typename T::Targets& I,
                                          typename T::Eps eps2, typename T::Sum (&ai)[K][6],
                                          typename T::Sum (&aj)[6], std::integer_sequence<int, R...>)
{
    (T::template step<R>(row, I, eps2, ai, aj), ...);
}

struct SymArgs {
    State s;
    int n;        // bodies, >= 2 full blocks
    int chunk;    // source-list entries per workgroup
    double eps2;
    Forces out;   // initialised by the block-diagonal pass
    const double* vref;  // mixed: reference velocity (3 doubles)
};

// Full block P of one system (bodies 0..n-1 of s and out) against entries
// i0 .. i1 of its source list: the body of both symmetric kernels.  FL > 0:
// target sums go to the outputs every FL tiles; FL = 0: into per-lane totals.
template <template <int> class TT, int K, int W, int FL>
__device__ __forceinline__ void sym_block_pairs(const State& s, const Forces& out, int n, double eps2d, int P,
                                                int i0, int i1, const VRef& vr)
{
    using T = TT<K>;
    using Sum = typename T::Sum;
    constexpr int B = 64 * K * W;
    constexpr int TPB = B / kSymTile;  // tiles per block
    __shared__ typename T::Rec dup[2][4][32];
    __shared__ Sum part[2][W][kSymTile][6];
    const int nr = n / B;
    const int dmax = sym_list_len(nr, P, false);  // entry dmax, if listed, is the partial block
    const int vlast = (n - nr * B + kSymTile - 1) / kSymTile;  // its tiles holding bodies
    const int ntiles = i1 < dmax ? (i1 - i0 + 1) * TPB : (i1 - i0) * TPB + vlast;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);  // wave-uniform: an SGPR
    const int p = lane & 15;
    const typename T::Eps eps2 = (typename T::Eps)eps2d;
    const int ibase = P * B + w * 64 * K + lane;

    typename T::Targets I;
    T::load_targets(s, ibase, I, vr);
    // first source of tile tt of this workgroup's range
    auto tile_src = [&](int tt) {
        const int i = i0 + tt / TPB;
        int q = P + 1 + i;
        if (q >= nr) q -= nr;
        if (i == dmax) q = nr;
        return q * B + (tt % TPB) * kSymTile;
    };

    typename T::Rec st{};
    if (t < kSymTile) {
        st = T::load_rec(s, tile_src(0) + t, n, vr);
        dup[0][t >> 4][t & 15] = st;
        dup[0][t >> 4][(t & 15) + 16] = st;
    }
    __syncthreads();

    typename T::Tot tot;  // unused (and optimised away) when FL > 0
    if constexpr (FL == 0) T::zero(tot);
    Sum ai[K][6];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 6; ++c) ai[k][c] = Sum(0);
    for (int tt = 0; tt < ntiles; ++tt) {
        const int buf = tt & 1;
        const bool more = tt + 1 < ntiles;
#pragma unroll 1
        for (int g = 0; g < 4; ++g) {
            Sum aj[6] = {Sum(0), Sum(0), Sum(0), Sum(0), Sum(0), Sum(0)};
            sym_group<T, K>(&dup[buf][g][p], I, eps2, ai, aj, std::make_integer_sequence<int, 16>{});
            // rows summed; every row then holds the same totals and all four
            // store them (an unconditional store keeps the step code free of
            // branches, which lets the compiler fold each rotation into its add)
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                Sum v = aj[c];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                part[buf][w][g * 16 + p][c] = v;
            }
        }
        // the next tile is loaded after the arithmetic, not held in registers
        // across it (12 VGPRs); the other resident workgroups cover the load.
        // It comes before the target atomics below, since a wait for a load
        // also waits for every memory operation issued before it
        // (lane indices past the step loop come from lane_now(), so that none
        // is held in a register, or spilled, across it)
        if (more && w == 0) {
            const int l = lane_now();
            st = T::load_rec(s, tile_src(tt + 1) + l, n, vr);
            dup[buf ^ 1][l >> 4][l & 15] = st;
            dup[buf ^ 1][l >> 4][(l & 15) + 16] = st;
        }
        if constexpr (FL == 0) {
            T::fold(tot, ai);
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int c = 0; c < 6; ++c) ai[k][c] = Sum(0);
        } else if (tt % FL == FL - 1 || !more) {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    out_add(out.c[c], (unsigned)(P * B + w * 64 * K + lane_now() + k * 64), (double)ai[k][c]);
                    ai[k][c] = Sum(0);
                }
        }
        __syncthreads();
        // the tile's source totals, waves combined in wave order: wave w adds
        // the components c = w (mod W) of the tile's 64 sources, lane j source j
        const int j0 = tile_src(tt);
        const int l = lane_now();
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            if (c % W == w && j0 + l < n) {
                double sum = 0.0;
#pragma unroll
                for (int v = 0; v < W; ++v) sum += (double)part[buf][v][l][c];
                out_add(out.c[c], (unsigned)(j0 + l), sum);
            }
        }
    }
    if constexpr (FL == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < 6; ++c) out_add(out.c[c], (unsigned)(ibase + k * 64), T::total(tot, k, c));
    }
}

// Resident waves per SIMD asked of the compiler: T::kWaves (mixed 5: 96
// VGPRs with three targets per lane, its few spills outside the step loop;
// FP64 3: 166 VGPRs), or fewer where the workgroups' LDS (160 KiB per CU)
// holds fewer (mixed with 2 waves: 18 KiB, 8 workgroups, 4 waves).  gfx950
// issues FP32 at twice the one-wave rate only by pairing instructions of two
// waves, so more resident waves raise the dual-issue share (mixed, two
// targets per lane with two-sum totals: +1.6% at n = 2^20 over 4 waves;
// three targets per lane with flushed sums: 5 waves 5% faster than 4).
template <class T, int W>
constexpr int sym_waves()
{
    constexpr int lds = 2 * 4 * 32 * sizeof(typename T::Rec) + 2 * W * kSymTile * 6 * sizeof(typename T::Sum);
    constexpr int by_lds = (160 * 1024 / lds) * W / 4;
    return by_lds < T::kWaves ? by_lds : T::kWaves;
}

template <template <int> class TT, int K, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(sym_waves<TT<K>, W>(), 8))) void k_force_sym(
    SymArgs a)
{
    constexpr int B = 64 * K * W;
    const int P = blockIdx.x;
    const int len = sym_list_len(a.n / B, P, a.n % B != 0);
    const int i0 = blockIdx.y * a.chunk;
    if (i0 >= len) return;  // uniform over the workgroup, before any barrier
    const int i1 = min(len, i0 + a.chunk) - 1;
    const VRef vr = a.vref ? load_vref(a.vref) : VRef{0.0, 0.0, 0.0};
    sym_block_pairs<TT, K, W, TT<K>::kFlushTiles>(a.s, a.out, a.n, a.eps2, P, i0, i1, vr);
}

// Many systems: one workgroup per item of a list (system offset and size,
// block P and a range of its source list), systems' arrays concatenated.
template <template <int> class TT, int K, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(sym_waves<TT<K>, W>(), 8))) void k_ens_sym(
    State s, Forces out, const SymItem* __restrict__ items, const double* __restrict__ vref)
{
    const SymItem it = items[blockIdx.x];
    const VRef vr = vref ? load_vref(vref + 3 * it.sys) : VRef{0.0, 0.0, 0.0};
    State q = s;
    q.x += it.off; q.y += it.off; q.z += it.off; q.m += it.off;
    q.vx += it.off; q.vy += it.off; q.vz += it.off;
    Forces o = out;
#pragma unroll
    for (int c = 0; c < 6; ++c) o.c[c] += it.off;
    sym_block_pairs<TT, K, W, 0>(q, o, it.n, it.eps2, it.P, it.i0, it.i1, vr);
}

// Launch shape per precision: targets per lane and waves per workgroup of
// one system's kernel (blocks of 768) and of the ensemble's (targets in
// registers: blocks of 512 in FP64, 256 mixed, since its systems are small
// and their diagonal blocks a larger share of the pairs).
constexpr int kMixK = 3, kMixW = 4, kMixEnsK = 2, kMixEnsW = 2;
constexpr int kF64K = 3, kF64W = 4, kF64EnsK = 2, kF64EnsW = 4;

}  // namespace

int sym_block(int precision) { return precision == 0 ? 64 * kF64K * kF64W : 64 * kMixK * kMixW; }

int sym_targets_per_lane(int precision) { return precision == 0 ? kF64K : kMixK; }

// n <= 2^29: the kernels address the outputs with 32-bit byte offsets
bool sym_supported(int precision, int n) { return n >= 2 * sym_block(precision) && n <= (1 << 29); }

int ens_sym_block(int precision) { return precision == 0 ? 64 * kF64EnsK * kF64EnsW : 64 * kMixEnsK * kMixEnsW; }

bool ens_sym_supported(int precision, int n) { return n >= 2 * ens_sym_block(precision); }

int sym_list_max(int precision, int n)
{
    const int b = sym_block(precision);
    return sym_list_len(n / b, 0, n % b != 0);
}

int sym_chunk(int precision, int n, int cus)
{
    const int b = sym_block(precision);
    const int per = sym_list_max(precision, n);
    // about 128 workgroups per CU in all (some 25 rounds of resident ones),
    // so the last round is a small part; measured at n = 2^18 and 2^20
    // against 16..256 per CU: within 0.5% of the best (64: 1% slower at 2^18)
    const long want = 128L * (cus > 0 ? cus : 256);
    long c = ((long)(n / b) * per + want - 1) / want;
    if (c < 1) c = 1;
    if (c > per) c = per;
    return (int)c;
}

long sym_workgroups(int precision, int n, int chunk)
{
    const int per = sym_list_max(precision, n);
    return (long)(n / sym_block(precision)) * ((per + chunk - 1) / chunk);
}

hipError_t launch_ens_sym(int precision, const State& s, const Forces& out, const SymItem* items, int nitems,
                          const double* vref, hipStream_t st)
{
    if (nitems <= 0) return hipSuccess;
    if (precision != 0 && !vref) return hipErrorInvalidValue;
    if (precision == 0)
        k_ens_sym<SymF64, kF64EnsK, kF64EnsW><<<nitems, 64 * kF64EnsW, 0, st>>>(s, out, items, nullptr);
    else
        k_ens_sym<SymMixed, kMixEnsK, kMixEnsW><<<nitems, 64 * kMixEnsW, 0, st>>>(s, out, items, vref);
    return hipGetLastError();
}

hipError_t launch_force_sym(int precision, const State& s, int n, double eps2, int chunk, const Forces& out,
                            const double* vref, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1)
{
    if (!sym_supported(precision, n) || chunk < 1 || (precision != 0 && !vref)) return hipErrorInvalidValue;
    const int B = sym_block(precision);
    hipError_t e;
    if (ev0 && (e = hipEventRecord(ev0, st)) != hipSuccess) return e;
    if ((e = launch_force_blockdiag(precision, s, n, eps2, B, out, vref, st)) != hipSuccess) return e;
    SymArgs a;
    a.s = s;
    a.n = n;
    a.chunk = chunk;
    a.eps2 = eps2;
    a.out = out;
    a.vref = vref;
    const int per = sym_list_max(precision, n);
    const dim3 grid(n / B, (per + chunk - 1) / chunk);
    if (precision == 0)
        k_force_sym<SymF64, kF64K, kF64W><<<grid, 64 * kF64W, 0, st>>>(a);
    else
        k_force_sym<SymMixed, kMixK, kMixW><<<grid, 64 * kMixW, 0, st>>>(a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, st)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace nbh
