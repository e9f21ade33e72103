// nbh_sym.hip -- mixed-precision pair kernel that evaluates each unordered
// pair once and applies it to both bodies (Newton's third law), as the
// reference's CPU loops do with their j < i triangle
// (src/kernel.cpp:1101-1163 for _F32, :1164-1224 for FP64).
//
// Layout of the work:
//   * bodies are cut into blocks of B = 64*K*W (K targets per lane, W waves
//     per workgroup); block P's pairs with block Q != P are evaluated by one
//     workgroup, for Q = P+1 .. P+H (mod nb), H = (nb-1)/2, plus Q = P+nb/2
//     for P < nb/2 when nb is even, so every unordered block pair is taken
//     exactly once and every block has the same work;
//   * pairs inside a block are evaluated one-sided by the directed kernel
//     (launch_force_blockdiag), which also initialises the outputs; this
//     kernel then adds its contributions with FP64 atomics;
//   * source bodies are staged in LDS in tiles of 64 = 4 groups of 16.  Each
//     group is stored twice in a row (32 records), so at step r the lane at
//     position p of its 16-lane row reads record p + r: the rows pair their
//     lanes with the group's 16 sources in rotated order, reading distinct
//     records at a compile-time offset;
//   * the reaction on the source (minus m_i r^3 times the same d and
//     u - 3 (d.u) r^2 d) is summed over the K targets of the lane, then
//     added to a per-lane accumulator after a DPP row rotation (row_ror:r)
//     that brings it to the lane holding that source: one DPP add per
//     component per step and K pairs.  After 16 steps the four rows are
//     summed with cross-lane shuffles, the waves' sums are combined in LDS
//     in wave order, and the tile's 64 x 6 source totals are added to the
//     outputs with FP64 atomics;
//   * target sums are FP32 over the 64 sources of a tile, folded into an
//     FP32 (hi, lo) pair with Knuth's two-sum, and added with FP64 atomics
//     when the workgroup's block pairs are done.
// Arithmetic per pair is that of the directed mixed kernel: position
// differences in FP64, rounded to FP32; FP32 from there on.
//
// The FP64 atomic adds make the summation order depend on scheduling: two
// runs agree to rounding of FP64 sums (about 1e-16 relative), not bit for
// bit.  The directed kernels remain the deterministic choice.
#include "nbh_internal.h"
#include "nbh_pair.h"

#include <utility>

namespace nbh {
namespace {

constexpr int kSymK = 2;     // targets per lane
constexpr int kSymW = 4;     // waves per workgroup
constexpr int kSymTile = 64; // sources per LDS tile (4 groups of 16)

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

template <int K>
struct SymTargets {
    double x[K], y[K], z[K];
    float vx[K], vy[K], vz[K], m[K];
};

template <int K, int R>
__device__ __forceinline__ void sym_step(const JBMX* __restrict__ row, const SymTargets<K>& I, float eps2,
                                         float (&ai)[K][6], float (&aj)[6])
{
    const JBMX b = row[R];
    float cs[6];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float dx = (float)(b.x - I.x[k]);
        const float dy = (float)(b.y - I.y[k]);
        const float dz = (float)(b.z - I.z[k]);
        const float ux = b.vx - I.vx[k];
        const float uy = b.vy - I.vy[k];
        const float uz = b.vz - I.vz[k];
        const float s = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const float vr = fmaf(dx, ux, fmaf(dy, uy, dz * uz));
        const float r = __builtin_amdgcn_rsqf(s);
        const float r2 = r * r;
        const float r3 = r * r2;
        const float mj3 = b.m * r3;
        const float mi3 = I.m[k] * r3;
        const float al = (-3.0f * r2) * vr;
        const float wx = fmaf(al, dx, ux);
        const float wy = fmaf(al, dy, uy);
        const float wz = fmaf(al, dz, uz);
        ai[k][0] = fmaf(mj3, dx, ai[k][0]);
        ai[k][1] = fmaf(mj3, dy, ai[k][1]);
        ai[k][2] = fmaf(mj3, dz, ai[k][2]);
        ai[k][3] = fmaf(mj3, wx, ai[k][3]);
        ai[k][4] = fmaf(mj3, wy, ai[k][4]);
        ai[k][5] = fmaf(mj3, wz, ai[k][5]);
        if (k == 0) {
            cs[0] = mi3 * dx; cs[1] = mi3 * dy; cs[2] = mi3 * dz;
            cs[3] = mi3 * wx; cs[4] = mi3 * wy; cs[5] = mi3 * wz;
        } else {
            cs[0] = fmaf(mi3, dx, cs[0]); cs[1] = fmaf(mi3, dy, cs[1]); cs[2] = fmaf(mi3, dz, cs[2]);
            cs[3] = fmaf(mi3, wx, cs[3]); cs[4] = fmaf(mi3, wy, cs[4]); cs[5] = fmaf(mi3, wz, cs[5]);
        }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) aj[c] -= row_rot<R>(cs[c]);
}

template <int K, int... R>
__device__ __forceinline__ void sym_group(const JBMX* __restrict__ row, const SymTargets<K>& I, float eps2,
                                          float (&ai)[K][6], float (&aj)[6], std::integer_sequence<int, R...>)
{
    (sym_step<K, R>(row, I, eps2, ai, aj), ...);
}

struct SymArgs {
    State s;
    int nb;      // blocks of kSymK*kSymW*64 bodies (n = nb * block)
    int half;    // H = (nb - 1) / 2
    int chunk;   // block pairs per workgroup
    float eps2;
    Forces out;  // initialised by the block-diagonal pass
};

template <int K, int W>
__global__ __launch_bounds__(64 * W) void k_force_mixed_sym(SymArgs a)
{
    constexpr int B = 64 * K * W;
    constexpr int TPB = B / kSymTile;  // tiles per block
    __shared__ JBMX dup[2][4][32];
    __shared__ float part[2][W][kSymTile][6];

    const int P = blockIdx.x;
    const int dmax = a.half + ((a.nb % 2 == 0 && P < a.nb / 2) ? 1 : 0);
    const int d0 = 1 + blockIdx.y * a.chunk;
    if (d0 > dmax) return;  // uniform over the workgroup, before any barrier
    const int d1 = min(dmax, d0 + a.chunk - 1);
    const int ntiles = (d1 - d0 + 1) * TPB;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int p = lane & 15;
    const float eps2 = a.eps2;
    const int ibase = P * B + w * 64 * K + lane;

    SymTargets<K> I;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = ibase + k * 64;
        I.x[k] = a.s.x[i]; I.y[k] = a.s.y[i]; I.z[k] = a.s.z[i];
        I.m[k] = (float)a.s.m[i];
        I.vx[k] = (float)a.s.vx[i]; I.vy[k] = (float)a.s.vy[i]; I.vz[k] = (float)a.s.vz[i];
    }
    // first source of tile tt of this workgroup's range
    auto tile_src = [&](int tt) {
        int q = P + d0 + tt / TPB;
        if (q >= a.nb) q -= a.nb;
        return q * B + (tt % TPB) * kSymTile;
    };

    JBMX st{};
    if (t < kSymTile) {
        st = load_jbmx(a.s, tile_src(0) + t);
        dup[0][t >> 4][t & 15] = st;
        dup[0][t >> 4][(t & 15) + 16] = st;
    }
    __syncthreads();

    float hi[K][6], lo[K][6];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 6; ++c) hi[k][c] = lo[k][c] = 0.f;

    for (int tt = 0; tt < ntiles; ++tt) {
        const int buf = tt & 1;
        const bool more = tt + 1 < ntiles;
        if (more && t < kSymTile) st = load_jbmx(a.s, tile_src(tt + 1) + t);
        float ai[K][6];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < 6; ++c) ai[k][c] = 0.f;
#pragma unroll 1
        for (int g = 0; g < 4; ++g) {
            float aj[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            sym_group<K>(&dup[buf][g][p], I, eps2, ai, aj, std::make_integer_sequence<int, 16>{});
            // rows summed; every row then holds the same totals and all four
            // store them (an unconditional store keeps the step code free of
            // branches, which lets the compiler fold each rotation into its add)
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                float v = aj[c];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                part[buf][w][g * 16 + p][c] = v;
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < 6; ++c) two_sum_fold(hi[k][c], lo[k][c], ai[k][c]);
        if (more && t < kSymTile) {
            dup[buf ^ 1][t >> 4][t & 15] = st;
            dup[buf ^ 1][t >> 4][(t & 15) + 16] = st;
        }
        __syncthreads();
        // the tile's source totals: waves combined in wave order
        const int j0 = tile_src(tt);
        for (int e = t; e < kSymTile * 6; e += 64 * W) {
            const int c = e / kSymTile;
            const int j = e % kSymTile;
            double sum = 0.0;
#pragma unroll
            for (int v = 0; v < W; ++v) sum += (double)part[buf][v][j][c];
            unsafeAtomicAdd(&a.out.c[c][j0 + j], sum);
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 6; ++c)
            unsafeAtomicAdd(&a.out.c[c][ibase + k * 64], (double)hi[k][c] + (double)lo[k][c]);
}

}  // namespace

int sym_block() { return 64 * kSymK * kSymW; }

bool sym_supported(int n) { return n % sym_block() == 0 && n >= 2 * sym_block(); }

int sym_chunk(int n, int cus)
{
    const int nb = n / sym_block();
    const int pairs_per_block = (nb - 1) / 2 + 1;
    // about 32 workgroups per CU in all, so the last round is a small part
    const long want = 32L * (cus > 0 ? cus : 256);
    long c = ((long)nb * pairs_per_block + want - 1) / want;
    if (c < 1) c = 1;
    if (c > pairs_per_block) c = pairs_per_block;
    return (int)c;
}

hipError_t launch_force_sym(const State& s, int n, double eps2, int chunk, const Forces& out,
                            hipStream_t st, hipEvent_t ev0, hipEvent_t ev1)
{
    if (!sym_supported(n) || chunk < 1) return hipErrorInvalidValue;
    hipError_t e;
    if (ev0 && (e = hipEventRecord(ev0, st)) != hipSuccess) return e;
    if ((e = launch_force_blockdiag(s, n, eps2, sym_block(), out, st)) != hipSuccess) return e;
    SymArgs a;
    a.s = s;
    a.nb = n / sym_block();
    a.half = (a.nb - 1) / 2;
    a.chunk = chunk;
    a.eps2 = (float)eps2;
    a.out = out;
    const int per = a.half + 1;
    const dim3 grid(a.nb, (per + chunk - 1) / chunk);
    k_force_mixed_sym<kSymK, kSymW><<<grid, 64 * kSymW, 0, st>>>(a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, st)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace nbh
