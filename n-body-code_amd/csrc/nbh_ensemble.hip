// nbh_ensemble.hip -- many independent systems advanced together.
//
// Layout: the systems' SoA arrays are concatenated in system order (system k
// occupies [off_k, off_k + n_k)).  A step is the force evaluation at the
// predicted state, whose kernels also run the corrector and the next
// prediction (and its source records) for the targets they complete, with
// the single-system kernels' arithmetic; predictions, records and forces
// alternate between two buffers from step to step.
//
// Force evaluation: one persistent launch of one-wave (64-lane) workgroups.
// Each wave takes work items from a device queue with an atomic counter:
// item = (system, 64 or 128 consecutive targets (one or two per lane, the
// same for every item of a list), a source range, an LDS tile length, an
// output slot).  Items are sorted by cost, largest first, so the
// long items start early and the short ones fill the tail.  Systems whose
// bucket uses a j-split write per-item partial sums into slots that a second
// kernel adds in split order (deterministic).  Every wave leaves the loop
// once the counter passes the item count, so the grid always drains.
#include "nbh_ctx.h"
#include "nbh_pair.h"
#include "nbh_hermite.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

namespace nbh {
namespace {

constexpr int kLanes = 64;
constexpr int kMaxTargets = 128;  // targets per item: 64 per lane-set, up to two sets
const int kBucketEdge[NBH_ENS_BUCKETS] = {64, 256, 1024, INT_MAX};
// LDS tile lengths of the persistent kernel; 0 selects the scalar-load
// kernel (k_ens_big) for a bucket's systems instead
const int kTiles[4] = {0, 64, 128, 256};
const int kSplits[4] = {1, 2, 4, 8};

struct EnsSys {
    int off;
    int n;
    double eps2;
};

struct EnsItem {
    int sys;
    int i0;      // first target (global index)
    int icount;  // targets in this item (<= 64 * ipl)
    int j0, j1;  // source range (global indices)
    int tile;    // LDS tile length for this item
    int slot;    // partial-sum slot, or -1: write the forces directly
    int ipl;     // targets per lane of the list: lane t holds i0 + t and, if 2, i0 + 64 + t
};

// One workgroup of the scalar-load kernel: up to 256 consecutive targets of
// system `sys` against its sources [j0, j1) (global indices).
struct EnsBig {
    int sys, i0, icount, j0, j1;
};

struct EnsReduce {
    int i0;
    int icount;
    int slot0;
    int splits;
};

// Where a step puts the corrected state, the forces and the next prediction
// (the prediction buffer it does not read, with its records): see the step
// kernels below.
struct EnsNext {
    double* cp[3];  // corrected positions (the current state)
    double* cv[3];  // corrected velocities
    double* np[3];  // next predicted positions
    double* nv[3];  // next predicted velocities
    const double* m;
    Forces f0;      // forces at the start of the step (read)
    Forces f1;      // this step's forces, at the prediction (written; the next step's f0)
    const double* dt;  // null: a force evaluation only, write `out`
    void* rec;      // next records (scalar-load systems), indexed like the bodies
};

struct EnsArgs {
    State s;
    Forces out;
    const EnsSys* sys;
    const EnsItem* items;
    int nitems;
    unsigned* counter;
    double* partial;  // [slot][6][kMaxTargets]
    const double* vref;  // mixed: reference velocity of each system, [3 * nsys]
    EnsNext nx;          // a step (nx.dt set): unsplit items of directed systems advance their targets
    const int* bodysys;  // [total]: system on the scalar-load kernel, -1 other directed, -2 symmetric
};

// Corrector of one body from its prediction (p, u) and this step's forces
// f1 at it (stored into nx.f1 with STORE, else already there), then the next
// prediction; returns the corrected state in (p, u) replaced by the next
// prediction.
template <bool STORE = true>
__device__ __forceinline__ void advance_body(const EnsNext& nx, int i, double dt, double invdt2, double invdt3,
                                             const double (&f1)[6], double (&p)[3], double (&u)[3])
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double a2, a3;
        hermite_correct_axis(dt, invdt2, invdt3, nx.f0.c[c][i], nx.f0.c[3 + c][i], f1[c], f1[3 + c], p[c], u[c],
                             a2, a3);
        nx.cp[c][i] = p[c];
        nx.cv[c][i] = u[c];
        if constexpr (STORE) {
            nx.f1.c[c][i] = f1[c];
            nx.f1.c[3 + c][i] = f1[3 + c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        hermite_predict_axis(dt, f1[c], f1[3 + c], p[c], u[c]);
        nx.np[c][i] = p[c];
        nx.nv[c][i] = u[c];
    }
}

int bucket_of(int n)
{
    for (int b = 0; b < NBH_ENS_BUCKETS; ++b)
        if (n <= kBucketEdge[b]) return b;
    return NBH_ENS_BUCKETS - 1;
}

// Lane 0 takes the next item; the index is made wave-uniform (an SGPR), so
// the item and its loop bounds are scalar loads and scalar branches.
__device__ __forceinline__ int take_item(unsigned* counter)
{
    int w = 0;
    if (threadIdx.x == 0) w = (int)atomicAdd(counter, 1u);
    return __builtin_amdgcn_readfirstlane(w);
}

// One work item with IPL targets per lane: sources streamed through the LDS
// tile, broadcast to the lanes, each record feeding IPL pair evaluations.
template <bool MIXED, int IPL, bool MASK, class JB>
__device__ __forceinline__ void ens_item(const EnsArgs& a, const EnsItem& it, JB* tile)
{
    constexpr int UNROLL = IPL == 1 ? 4 : 2;
    const int lane = threadIdx.x;
    const double eps2 = a.sys[it.sys].eps2;
    int ti[IPL];
    double xi[IPL], yi[IPL], zi[IPL];
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        ti[q] = it.i0 + q * kLanes + lane;
        const int ic = q * kLanes + lane < it.icount ? ti[q] : it.i0;
        xi[q] = a.s.x[ic]; yi[q] = a.s.y[ic]; zi[q] = a.s.z[ic];
    }
    double out[IPL][6];
    if constexpr (MIXED) {
        const VRef vr = load_vref(a.vref + 3 * it.sys);
        float vxi[IPL], vyi[IPL], vzi[IPL];
#pragma unroll
        for (int q = 0; q < IPL; ++q) {
            const int ic = q * kLanes + lane < it.icount ? ti[q] : it.i0;
            vxi[q] = (float)(a.s.vx[ic] - vr.x);
            vyi[q] = (float)(a.s.vy[ic] - vr.y);
            vzi[q] = (float)(a.s.vz[ic] - vr.z);
        }
        const float e2 = (float)eps2;
        float hi[IPL][6], lo[IPL][6];
#pragma unroll
        for (int q = 0; q < IPL; ++q)
#pragma unroll
            for (int c = 0; c < 6; ++c) hi[q][c] = lo[q][c] = 0.f;
        for (int jt = it.j0; jt < it.j1; jt += it.tile) {
            const int cnt = min(it.tile, it.j1 - jt);
            __syncthreads();
            for (int k = lane; k < cnt; k += kLanes) tile[k] = load_jbmx(a.s, jt + k, vr);
            __syncthreads();
            for (int k0 = 0; k0 < cnt; k0 += kFold) {
                float acc[IPL][6];
#pragma unroll
                for (int q = 0; q < IPL; ++q)
#pragma unroll
                    for (int c = 0; c < 6; ++c) acc[q][c] = 0.f;
                const int kend = min(cnt, k0 + kFold);
                if (kend - k0 == kFold) {
#pragma unroll UNROLL
                    for (int k = k0; k < k0 + kFold; ++k) {
                        const JBMX b = tile[k];
#pragma unroll
                        for (int q = 0; q < IPL; ++q)
                            pair_mx<MASK>(b, jt + k, ti[q], xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q], e2, acc[q]);
                    }
                } else {
                    for (int k = k0; k < kend; ++k) {
                        const JBMX b = tile[k];
#pragma unroll
                        for (int q = 0; q < IPL; ++q)
                            pair_mx<MASK>(b, jt + k, ti[q], xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q], e2, acc[q]);
                    }
                }
#pragma unroll
                for (int q = 0; q < IPL; ++q)
#pragma unroll
                    for (int c = 0; c < 6; ++c) two_sum_fold(hi[q][c], lo[q][c], acc[q][c]);
            }
        }
#pragma unroll
        for (int q = 0; q < IPL; ++q)
#pragma unroll
            for (int c = 0; c < 6; ++c) out[q][c] = (double)hi[q][c] + (double)lo[q][c];
    } else {
        double vxi[IPL], vyi[IPL], vzi[IPL];
#pragma unroll
        for (int q = 0; q < IPL; ++q) {
            const int ic = q * kLanes + lane < it.icount ? ti[q] : it.i0;
            vxi[q] = a.s.vx[ic]; vyi[q] = a.s.vy[ic]; vzi[q] = a.s.vz[ic];
        }
        double acc[IPL][6];
#pragma unroll
        for (int q = 0; q < IPL; ++q)
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[q][c] = 0.0;
        for (int jt = it.j0; jt < it.j1; jt += it.tile) {
            const int cnt = min(it.tile, it.j1 - jt);
            __syncthreads();
            for (int k = lane; k < cnt; k += kLanes) tile[k] = load_jb64(a.s, jt + k);
            __syncthreads();
            if (cnt == it.tile) {
#pragma unroll UNROLL
                for (int k = 0; k < cnt; ++k) {
                    const JB64 b = tile[k];
#pragma unroll
                    for (int q = 0; q < IPL; ++q)
                        pair_f64<MASK>(b, jt + k, ti[q], xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q], eps2, acc[q]);
                }
            } else {
                for (int k = 0; k < cnt; ++k) {
                    const JB64 b = tile[k];
#pragma unroll
                    for (int q = 0; q < IPL; ++q)
                        pair_f64<MASK>(b, jt + k, ti[q], xi[q], yi[q], zi[q], vxi[q], vyi[q], vzi[q], eps2, acc[q]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < IPL; ++q)
#pragma unroll
            for (int c = 0; c < 6; ++c) out[q][c] = acc[q][c];
    }
    // a step advances the targets of an unsplit item of a directed system
    // (the forces are complete here); a symmetric system's diagonal block
    // still gets the off-diagonal pairs added to `out`
    const bool adv = a.nx.dt && it.slot < 0 && a.bodysys[it.i0] == -1;
    const double dt = adv ? *a.nx.dt : 0.0;
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int t = q * kLanes + lane;
        if (t < it.icount) {
            if (adv) {
                double p[3] = {xi[q], yi[q], zi[q]};
                double u[3] = {a.s.vx[ti[q]], a.s.vy[ti[q]], a.s.vz[ti[q]]};
                advance_body(a.nx, ti[q], dt, 1.0 / (dt * dt), 1.0 / (dt * dt * dt), out[q], p, u);
            } else if (it.slot < 0) {
#pragma unroll
                for (int c = 0; c < 6; ++c) a.out.c[c][ti[q]] = out[q][c];
            } else {
                double* p = a.partial + (size_t)it.slot * 6 * kMaxTargets + t;
#pragma unroll
                for (int c = 0; c < 6; ++c) p[c * kMaxTargets] = out[q][c];
            }
        }
    }
}

// One instance per targets-per-lane count, so that the one-target kernel
// keeps its lower register count (and occupancy).
template <bool MIXED, int MAXT, bool MASK, int IPL>
__global__ __launch_bounds__(kLanes) void k_ens_force(EnsArgs a)
{
    using JB = typename std::conditional<MIXED, JBMX, JB64>::type;
    __shared__ JB tile[MAXT];
    for (;;) {
        const int w = take_item(a.counter);
        if (w >= a.nitems) break;
        ens_item<MIXED, IPL, MASK>(a, a.items[w], tile);
    }
}

// Split items' partial sums added in split order; in a step (nx.dt set),
// the targets are then advanced (split items are always directed systems).
__global__ __launch_bounds__(kLanes) void k_ens_reduce(const EnsReduce* r, int nr,
                                                     const double* __restrict__ partial, Forces out, State s,
                                                     EnsNext nx)
{
    const int b = blockIdx.x;
    if (b >= nr) return;
    const EnsReduce q = r[b];
    for (int t = threadIdx.x; t < q.icount; t += kLanes) {
        double f[6];
        for (int c = 0; c < 6; ++c) {
            double v = 0.0;
            for (int sp = 0; sp < q.splits; ++sp) v += partial[((size_t)(q.slot0 + sp) * 6 + c) * kMaxTargets + t];
            f[c] = v;
        }
        const int i = q.i0 + t;
        if (nx.dt) {
            const double dt = *nx.dt;
            double p[3] = {s.x[i], s.y[i], s.z[i]};
            double u[3] = {s.vx[i], s.vy[i], s.vz[i]};
            advance_body(nx, i, dt, 1.0 / (dt * dt), 1.0 / (dt * dt * dt), f, p, u);
        } else {
            for (int c = 0; c < 6; ++c) out.c[c][i] = f[c];
        }
    }
}

__global__ void k_zero_u32(unsigned* p)
{
    if (threadIdx.x == 0) *p = 0u;
}

// Centre-of-mass velocity of every system (mixed precision's reference
// velocity, subtracted before velocities are rounded to FP32): one wave per
// system, lane sums combined by a fixed butterfly (deterministic); 0 for a
// system without mass.
__global__ __launch_bounds__(kLanes) void k_ens_vref(State s, const EnsSys* __restrict__ sys, int nsys,
                                                   double* __restrict__ vref)
{
    const int k = blockIdx.x;
    if (k >= nsys) return;
    const EnsSys q = sys[k];
    double m = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    for (int i = q.off + (int)threadIdx.x; i < q.off + q.n; i += kLanes) {
        const double mi = s.m[i];
        m += mi;
        px = fma(mi, s.vx[i], px);
        py = fma(mi, s.vy[i], py);
        pz = fma(mi, s.vz[i], pz);
    }
    for (int off = kLanes / 2; off > 0; off >>= 1) {
        m += __shfl_xor(m, off, kLanes);
        px += __shfl_xor(px, off, kLanes);
        py += __shfl_xor(py, off, kLanes);
        pz += __shfl_xor(pz, off, kLanes);
    }
    if (threadIdx.x == 0) {
        vref[3 * k + 0] = m > 0.0 ? px / m : 0.0;
        vref[3 * k + 1] = m > 0.0 ? py / m : 0.0;
        vref[3 * k + 2] = m > 0.0 ? pz / m : 0.0;
    }
}

// Source records of the systems the scalar-load kernel covers (list), one
// workgroup per system: record j of body j, velocities less the system's
// reference velocity in mixed precision.
constexpr int kBigBlock = 256;

template <bool MIXED>
__global__ __launch_bounds__(kBigBlock) void k_ens_pack(State s, const EnsSys* __restrict__ sys,
                                                      const int* __restrict__ list, const double* __restrict__ vref,
                                                      void* __restrict__ rec)
{
    const int k = list[blockIdx.x];
    const EnsSys q = sys[k];
    if constexpr (MIXED) {
        const VRef vr = load_vref(vref + 3 * k);
        for (int j = q.off + (int)threadIdx.x; j < q.off + q.n; j += kBigBlock)
            static_cast<JBMX*>(rec)[j] = load_jbmx(s, j, vr);
    } else {
        for (int j = q.off + (int)threadIdx.x; j < q.off + q.n; j += kBigBlock)
            static_cast<JB64*>(rec)[j] = load_jb64(s, j);
    }
}

// The ensemble's step after its force evaluation.  The prediction of a step
// is double-buffered (pred[0], pred[1], and likewise the records of the
// scalar-load kernel's systems and the forces): step n reads buffer p and
// writes its forces and the next prediction into 1 - p, so a kernel can
// write the next prediction of its own targets while other workgroups of
// the same system still read this step's sources, and no forces are copied.
//   * Systems on the scalar-load kernel (k_ens_big with ADVANCE): each
//     workgroup, once its targets' forces are summed, stores them, runs the
//     corrector and the next prediction for those targets, and packs their
//     next records: no pass over memory of its own.
//   * Other directed systems likewise in the persistent kernel's unsplit
//     items (ens_item) and, for split ones, in k_ens_reduce.
//   * Symmetric systems, whose forces are complete only after the symmetric
//     kernel: k_ens_advance, one thread per body.
//   * The first step of a run: k_ens_prologue predicts from the current
//     state and f0, and packs the records.
// The arithmetic is the single-system kernels' (nbh_hermite.h), so the
// states are bitwise those of k_predict and k_correct.  The reference
// velocity of each system (mixed precision) is the centre-of-mass velocity
// of the state the last force evaluation with preparation was made on
// (upload, autotune): pairwise forces conserve it, so it stays the
// centre-of-mass velocity to rounding, without a reduction per step.
constexpr int kAdvBlock = 256;

// One thread per body of the symmetric systems (bodysys[i] == -2), whose
// forces are complete only after the symmetric kernel's atomic adds: the
// prediction and this step's forces f1 at it (already in nx.f1).
__global__ __launch_bounds__(kAdvBlock) void k_ens_advance(EnsNext nx, State pred, Forces f1,
                                                         const int* __restrict__ bodysys, int total)
{
    const int i = blockIdx.x * kAdvBlock + threadIdx.x;
    if (i >= total || bodysys[i] != -2) return;
    const double dt = *nx.dt;
    const double invdt2 = 1.0 / (dt * dt);
    const double invdt3 = 1.0 / (dt * dt * dt);
    double f[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) f[c] = f1.c[c][i];
    double p[3] = {pred.x[i], pred.y[i], pred.z[i]};
    double u[3] = {pred.vx[i], pred.vy[i], pred.vz[i]};
    advance_body<false>(nx, i, dt, invdt2, invdt3, f, p, u);
}

// First prediction of a run, one thread per body: from the current state
// (nx.cp, nx.cv) and f0 into nx.np, nx.nv, and the records of the bodies
// whose system (bodysys[i] >= 0) is on the scalar-load kernel.
template <bool MIXED>
__global__ __launch_bounds__(kAdvBlock) void k_ens_prologue(EnsNext nx, const int* __restrict__ bodysys,
                                                          const double* __restrict__ vref, int total)
{
    const int i = blockIdx.x * kAdvBlock + threadIdx.x;
    if (i >= total) return;
    const double dt = *nx.dt;
    double p[3], u[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        p[c] = nx.cp[c][i];
        u[c] = nx.cv[c][i];
        hermite_predict_axis(dt, nx.f0.c[c][i], nx.f0.c[3 + c][i], p[c], u[c]);
        nx.np[c][i] = p[c];
        nx.nv[c][i] = u[c];
    }
    const int k = bodysys[i];
    if (k >= 0 && nx.rec) {
        if constexpr (MIXED)
            static_cast<JBMX*>(nx.rec)[i] = make_jbmx(p[0], p[1], p[2], nx.m[i], u[0], u[1], u[2],
                                                      load_vref(vref + 3 * k));
        else
            static_cast<JB64*>(nx.rec)[i] = make_jb64(p[0], p[1], p[2], nx.m[i], u[0], u[1], u[2]);
    }
}

// Scalar-load kernel for the large systems: the single-system directed
// kernel's structure (one target per lane, four waves per workgroup, the
// source record at a wave-uniform index so that it is read with scalar
// loads into SGPRs, no LDS), here over one system's targets and sources per
// workgroup.  Every wave of a workgroup streams the same records, which the
// scalar cache then serves to all four.  Each target sums its sources in
// order from j0, FP32 runs of kFold folded into the two-sum (hi, lo) totals
// in mixed precision: the same sums, bit for bit, as the persistent kernel
// over the same source range.  ADVANCE (a step): instead of writing the
// forces to `out`, each target is corrected and predicted for the next step
// (advance_body) and its next record packed, into nx.
template <bool MIXED, bool MASK, bool ADVANCE>
__global__ __launch_bounds__(kBigBlock) void k_ens_big(State s, Forces out, const EnsSys* __restrict__ sys,
                                                     const EnsBig* __restrict__ items, const void* __restrict__ rec,
                                                     const double* __restrict__ vref, EnsNext nx)
{
    const EnsBig it = items[blockIdx.x];
    const int t = threadIdx.x;
    // a system's last workgroup holds fewer than 256 targets: waves with
    // none of them leave at once (wave-uniform, and the kernel has no
    // barrier), so only the partial wave's idle lanes repeat a target
    if ((t & ~63) >= it.icount) return;
    const int i = it.i0 + (t < it.icount ? t : 0);
    const double eps2 = sys[it.sys].eps2;
    const double xi = s.x[i], yi = s.y[i], zi = s.z[i];
    double f[6];
    VRef vr{0.0, 0.0, 0.0};
    if constexpr (MIXED) {
        vr = load_vref(vref + 3 * it.sys);
        const float vxi = (float)(s.vx[i] - vr.x), vyi = (float)(s.vy[i] - vr.y), vzi = (float)(s.vz[i] - vr.z);
        const float e2 = (float)eps2;
        const JBMX* __restrict__ r = static_cast<const JBMX*>(rec);
        float hi[6] = {0, 0, 0, 0, 0, 0}, lo[6] = {0, 0, 0, 0, 0, 0};
        for (int j0 = it.j0; j0 < it.j1; j0 += kFold) {
            float acc[6] = {0, 0, 0, 0, 0, 0};
            if (j0 + kFold <= it.j1) {
                // 16 sources per unrolled body: 1.6% faster over the 8192
                // systems than 8 (4: 2.6% slower; 32 and 64 take 119 and 154
                // VGPRs against 54)
#pragma unroll 16
                for (int j = j0; j < j0 + kFold; ++j)
                    pair_mx<MASK>(r[j], j, i, xi, yi, zi, vxi, vyi, vzi, e2, acc);
            } else {
                for (int j = j0; j < it.j1; ++j) pair_mx<MASK>(r[j], j, i, xi, yi, zi, vxi, vyi, vzi, e2, acc);
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) two_sum_fold(hi[c], lo[c], acc[c]);
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) f[c] = (double)hi[c] + (double)lo[c];
    } else {
        const double vxi = s.vx[i], vyi = s.vy[i], vzi = s.vz[i];
        const JB64* __restrict__ r = static_cast<const JB64*>(rec);
        double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4
        for (int j = it.j0; j < it.j1; ++j) pair_f64<MASK>(r[j], j, i, xi, yi, zi, vxi, vyi, vzi, eps2, acc);
#pragma unroll
        for (int c = 0; c < 6; ++c) f[c] = acc[c];
    }
    if (t < it.icount) {
        if constexpr (ADVANCE) {
            const double dt = *nx.dt;
            double p[3] = {xi, yi, zi};
            double u[3] = {s.vx[i], s.vy[i], s.vz[i]};
            advance_body(nx, i, dt, 1.0 / (dt * dt), 1.0 / (dt * dt * dt), f, p, u);
            if constexpr (MIXED)
                static_cast<JBMX*>(nx.rec)[i] = make_jbmx(p[0], p[1], p[2], s.m[i], u[0], u[1], u[2], vr);
            else
                static_cast<JB64*>(nx.rec)[i] = make_jb64(p[0], p[1], p[2], s.m[i], u[0], u[1], u[2]);
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) out.c[c][it.i0 + t] = f[c];
        }
    }
}

// Per-system kinetic and softened potential energy (FP64), one workgroup per
// system, block sums in a fixed order.
constexpr int kEnBlock = 256;
__global__ __launch_bounds__(kEnBlock) void k_ens_energy(State s, const EnsSys* sys, double* kin,
                                                         double* pot)
{
    __shared__ double rk[kEnBlock], rp[kEnBlock];
    const EnsSys q = sys[blockIdx.x];
    const int t = threadIdx.x;
    double k = 0.0, p = 0.0;
    for (int i = q.off + t; i < q.off + q.n; i += kEnBlock) {
        const double xi = s.x[i], yi = s.y[i], zi = s.z[i];
        double phi = 0.0;
        for (int j = q.off; j < q.off + q.n; ++j) {
            const double dx = s.x[j] - xi, dy = s.y[j] - yi, dz = s.z[j] - zi;
            double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, q.eps2)));
            const bool keep = j != i;
            r2 = keep ? r2 : 1.0;
            phi = fma(keep ? s.m[j] : 0.0, rsqrt_f64(r2), phi);
        }
        const double mi = s.m[i];
        p += -0.5 * mi * phi;
        k += 0.5 * mi * (s.vx[i] * s.vx[i] + s.vy[i] * s.vy[i] + s.vz[i] * s.vz[i]);
    }
    rk[t] = k;
    rp[t] = p;
    __syncthreads();
    for (int h = kEnBlock / 2; h > 0; h >>= 1) {
        if (t < h) {
            rk[t] += rk[t + h];
            rp[t] += rp[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        kin[blockIdx.x] = rk[0];
        pot[blockIdx.x] = rp[0];
    }
}

template <bool MIXED, int MAXT, bool MASK, int IPL>
const void* ens_kernel_ptr()
{
    return reinterpret_cast<const void*>(&k_ens_force<MIXED, MAXT, MASK, IPL>);
}

template <bool MIXED, int MAXT>
const void* ens_kernel_for(bool mask, int ipl)
{
    if (ipl == 2) return mask ? ens_kernel_ptr<MIXED, MAXT, true, 2>() : ens_kernel_ptr<MIXED, MAXT, false, 2>();
    return mask ? ens_kernel_ptr<MIXED, MAXT, true, 1>() : ens_kernel_ptr<MIXED, MAXT, false, 1>();
}

const void* pick_ens_kernel(bool mixed, int maxt, bool mask, int ipl)
{
    if (ipl != 1 && ipl != 2) return nullptr;
#define NBH_ENS_K(M, T) \
    if (mixed == M && maxt == T) return ens_kernel_for<M, T>(mask, ipl);
    NBH_ENS_K(false, 64)
    NBH_ENS_K(false, 128)
    NBH_ENS_K(false, 256)
    NBH_ENS_K(true, 64)
    NBH_ENS_K(true, 128)
    NBH_ENS_K(true, 256)
#undef NBH_ENS_K
    return nullptr;
}

}  // namespace
}  // namespace nbh

using nbh::DevBuf;
using nbh::fail;

struct nbh_ensemble {
    nbh_ctx* ctx = nullptr;
    int nsys = 0;
    int total = 0;
    std::vector<int> off, cnt;
    std::vector<double> eps2;
    // some system's eps2 needs the self pair dropped (self_mask), per precision
    bool any_mask[2] = {false, false};
    int tile[NBH_ENS_BUCKETS] = {128, 128, 128, 128};
    int split[NBH_ENS_BUCKETS] = {1, 1, 1, 1};
    int ipl = 1;  // targets per lane, every work item

    bool sym = false;  // systems of >= 2 symmetric blocks use symmetric pairing

    // device
    // force[parity] holds the forces at the current state; a step writes its
    // forces, at the prediction, into force[1 - parity]
    DevBuf force[2];
    DevBuf sys, state, items, reduces, partial, counter, dt, ekin, epot, symitems, vref;
    // scalar-load kernel: items, its systems, and two buffers of packed
    // records (one per prediction buffer), each indexed like the bodies
    DevBuf bigitems, biglist, rec;
    // [total] the body's system if that is on the scalar-load kernel, -2 if
    // it is symmetric, -1 for any other directed system
    DevBuf bodysys;
    // two buffers of the predicted state (x, y, z, vx, vy, vz; masses are the
    // state's): pred[parity] holds this step's prediction, the step writes
    // the next one into pred[1 - parity]
    DevBuf pred;
    int parity = 0;
    int nflat = 0;  // bodies of symmetric systems (advanced by k_ens_advance)
    int nitems = 0, nreduce = 0, nslots = 0, maxtile = 64, grid = 0, nsym = 0, nbig = 0, nbigsys = 0;
    int list_ipl = 1;  // targets per lane of the uploaded work list
    bool uploaded = false, dt_set = false;
    // pred[parity] and its records hold the prediction of the next step from
    // the current state, f0, dt and the reference velocities (cleared by
    // anything that changes one of them, or the records' coverage)
    bool pred_valid = false;
    hipGraphExec_t gexec = nullptr;
    int g_steps = 0;
    int g_parity = 0;  // parity the captured steps start from (and end at)
    // second stream (default priority) on which the persistent launch runs
    // beside the scalar-load launch, forked from and joined to the
    // context's stream with events (captured into the step graph as two
    // branches)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace {

using nbh::EnsBig;
using nbh::EnsItem;
using nbh::EnsReduce;

void drop_graph(nbh_ensemble* e)
{
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    e->gexec = nullptr;
    e->g_steps = 0;
}

// Work list for the systems selected by `use` (all when empty), with the
// per-bucket tile/split of `tile`/`split` and `ipl` targets per lane.  With
// symmetric pairing, a system of at least two symmetric blocks gets directed
// items for its diagonal blocks only (each target against its own block,
// written directly) and symmetric items for its block pairs, which the
// symmetric kernel adds afterwards.
// A bucket with tile 0 puts its (directed) systems on the scalar-load
// kernel instead: one item per 256 targets, the whole system as sources.
struct WorkList {
    std::vector<EnsItem> items;
    std::vector<EnsReduce> reds;
    std::vector<nbh::SymItem> sym;
    std::vector<EnsBig> big;
    std::vector<int> bigsys;
    std::vector<int> symsys;  // systems with symmetric pairing
    int nslots = 0;
    int maxtile() const
    {
        int t = 64;
        for (const auto& it : items) t = std::max(t, it.tile);
        return t;
    }
};

void build_items(const nbh_ensemble* e, const int* tile, const int* split, int ipl,
                 const std::vector<char>& use, WorkList& w)
{
    w = WorkList{};
    std::vector<EnsItem>& items = w.items;
    std::vector<EnsReduce>& reds = w.reds;
    std::vector<nbh::SymItem>& sym = w.sym;
    int& nslots = w.nslots;
    const int prec = e->ctx->cfg.precision;
    const int SB = nbh::ens_sym_block(prec);
    for (int k = 0; k < e->nsys; ++k) {
        if (!use.empty() && !use[k]) continue;
        const int n = e->cnt[k], off = e->off[k];
        const int b = nbh::bucket_of(n);
        const int T = tile[b];
        const int Q = ipl;
        const int per = nbh::kLanes * Q;
        if (e->sym && nbh::ens_sym_supported(prec, n)) {
            // diagonal blocks on the persistent LDS kernel (a bucket set to
            // the scalar-load kernel, tile 0, takes tiles of 256 here)
            const int TD = T > 0 ? T : 256;
            const int nb = (n + SB - 1) / SB, nr = n / SB;
            for (int P = 0; P < nb; ++P) {
                const int b0 = P * SB, b1 = std::min(n, b0 + SB);
                for (int ib = b0; ib < b1; ib += per)
                    items.push_back({k, off + ib, std::min(per, b1 - ib), off + b0, off + b1, TD, -1, Q});
                const int len = P < nr ? nbh::sym_list_len(nr, P, nb > nr) : 0;
                if (len >= 1) sym.push_back({off, n, P, 0, len - 1, k, e->eps2[k]});
            }
            w.symsys.push_back(k);
            continue;
        }
        if (T == 0) {
            for (int ib = 0; ib < n; ib += nbh::kBigBlock)
                w.big.push_back({k, off + ib, std::min(nbh::kBigBlock, n - ib), off, off + n});
            w.bigsys.push_back(k);
            continue;
        }
        int S = std::min(split[b], (n + T - 1) / T);
        S = std::max(S, 1);
        const int chunk = (n + S - 1) / S;
        S = (n + chunk - 1) / chunk;
        for (int ib = 0; ib < n; ib += per) {
            const int ic = std::min(per, n - ib);
            if (S == 1) {
                items.push_back({k, off + ib, ic, off, off + n, T, -1, Q});
            } else {
                reds.push_back({off + ib, ic, nslots, S});
                for (int s = 0; s < S; ++s)
                    items.push_back({k, off + ib, ic, off + s * chunk, off + std::min(n, (s + 1) * chunk), T,
                                     nslots + s, Q});
                nslots += S;
            }
        }
    }
    // most pairs first: sources listed in the item's entries (the partial
    // block, if listed, is the last entry of its list)
    auto work = [SB](const nbh::SymItem& it) {
        const int nr = it.n / SB;
        const int full = nbh::sym_list_len(nr, it.P, false);  // entries before the partial block
        double w = std::min(it.i1 + 1, full) - it.i0;
        if (it.i1 >= full) w += (double)(it.n - nr * SB) / SB;
        return w;
    };
    std::stable_sort(sym.begin(), sym.end(),
                     [&](const nbh::SymItem& a, const nbh::SymItem& b) { return work(a) > work(b); });
    // largest source range first; stable keeps system order among equals
    std::stable_sort(items.begin(), items.end(),
                     [](const EnsItem& a, const EnsItem& b) { return (a.j1 - a.j0) > (b.j1 - b.j0); });
    std::stable_sort(w.big.begin(), w.big.end(),
                     [](const EnsBig& a, const EnsBig& b) { return (a.j1 - a.j0) > (b.j1 - b.j0); });
}

int upload_items(nbh_ensemble* e, const WorkList& w, int ipl)
{
    nbh_ctx* c = e->ctx;
    const std::vector<EnsItem>& items = w.items;
    const std::vector<EnsReduce>& reds = w.reds;
    const std::vector<nbh::SymItem>& sym = w.sym;
    const int nslots = w.nslots;
    const int maxtile = w.maxtile();
    // the persistent kernel steps through an item's sources a tile at a time:
    // every item must have a tile and a non-empty range within its system
    for (const auto& it : items)
        if (!(it.tile == 64 || it.tile == 128 || it.tile == 256) || it.j1 <= it.j0 || it.icount < 1 ||
            it.icount > nbh::kLanes * it.ipl)
            return fail(NBH_ERR_STATE, "ensemble work item of system %d has tile %d, sources [%d, %d), %d targets",
                        it.sys, it.tile, it.j0, it.j1, it.icount);
    for (const auto& b : w.big)
        if (b.j1 <= b.j0 || b.icount < 1 || b.icount > nbh::kBigBlock)
            return fail(NBH_ERR_STATE, "scalar-load item of system %d has sources [%d, %d), %d targets", b.sys,
                        b.j0, b.j1, b.icount);
    if (!w.big.empty()) {
        if (e->bigitems.reserve(sizeof(EnsBig) * w.big.size()) != hipSuccess ||
            e->biglist.reserve(sizeof(int) * w.bigsys.size()) != hipSuccess ||
            e->rec.reserve(2 * sizeof(nbh::JB64) * (size_t)e->total) != hipSuccess)
            return fail(NBH_ERR_ALLOC, "ensemble scalar-load buffers allocation failed");
        NBH_HIP(hipMemcpyAsync(e->bigitems.p, w.big.data(), sizeof(EnsBig) * w.big.size(), hipMemcpyHostToDevice,
                               c->stream));
        NBH_HIP(hipMemcpyAsync(e->biglist.p, w.bigsys.data(), sizeof(int) * w.bigsys.size(),
                               hipMemcpyHostToDevice, c->stream));
    }
    {
        // each body's system if that is on the scalar-load kernel, -2 for a
        // symmetric system, -1 for any other
        std::vector<int> bsys(e->total, -1);
        for (int k : w.bigsys) std::fill(bsys.begin() + e->off[k], bsys.begin() + e->off[k] + e->cnt[k], k);
        int nflat = 0;
        for (int k : w.symsys) {
            std::fill(bsys.begin() + e->off[k], bsys.begin() + e->off[k] + e->cnt[k], -2);
            nflat += e->cnt[k];
        }
        if (e->bodysys.reserve(sizeof(int) * (size_t)e->total) != hipSuccess)
            return fail(NBH_ERR_ALLOC, "ensemble body index allocation failed");
        NBH_HIP(hipMemcpyAsync(e->bodysys.p, bsys.data(), sizeof(int) * (size_t)e->total, hipMemcpyHostToDevice,
                               c->stream));
        NBH_HIP(hipStreamSynchronize(c->stream));
        e->nflat = nflat;
    }
    e->nbig = (int)w.big.size();
    e->nbigsys = (int)w.bigsys.size();
    e->pred_valid = false;
    if (e->symitems.reserve(sizeof(nbh::SymItem) * std::max<size_t>(sym.size(), 1)) != hipSuccess ||
        e->items.reserve(sizeof(EnsItem) * std::max<size_t>(items.size(), 1)) != hipSuccess ||
        e->reduces.reserve(sizeof(EnsReduce) * std::max<size_t>(reds.size(), 1)) != hipSuccess ||
        e->partial.reserve(sizeof(double) * 6 * nbh::kMaxTargets * std::max(nslots, 1)) != hipSuccess)
        return fail(NBH_ERR_ALLOC, "ensemble work-list allocation failed");
    if (!items.empty())
        NBH_HIP(hipMemcpyAsync(e->items.p, items.data(), sizeof(EnsItem) * items.size(), hipMemcpyHostToDevice,
                               c->stream));
    if (!reds.empty())
        NBH_HIP(hipMemcpyAsync(e->reduces.p, reds.data(), sizeof(EnsReduce) * reds.size(),
                               hipMemcpyHostToDevice, c->stream));
    if (!sym.empty())
        NBH_HIP(hipMemcpyAsync(e->symitems.p, sym.data(), sizeof(nbh::SymItem) * sym.size(),
                               hipMemcpyHostToDevice, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    e->nitems = (int)items.size();
    e->nreduce = (int)reds.size();
    e->nsym = (int)sym.size();
    e->nslots = nslots;
    e->maxtile = maxtile;
    e->list_ipl = ipl;
    const void* k =
        nbh::pick_ens_kernel(c->cfg.precision == NBH_MIXED, maxtile, e->any_mask[c->cfg.precision != 0], ipl);
    if (!k) return fail(NBH_ERR_ARG, "no ensemble kernel for tile %d, %d targets per lane", maxtile, ipl);
    int per_cu = 0;
    NBH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, nbh::kLanes, 0));
    e->grid = std::max(1, std::min(e->nitems, per_cu * c->cus));
    return NBH_OK;
}

int rebuild(nbh_ensemble* e)
{
    WorkList w;
    build_items(e, e->tile, e->split, e->ipl, {}, w);
    drop_graph(e);
    return upload_items(e, w, e->ipl);
}

// Record buffer q (0 or 1) of the scalar-load kernel.
void* rec_of(const nbh_ensemble* e, int q)
{
    return static_cast<char*>(e->rec.p) + (size_t)q * sizeof(nbh::JB64) * (size_t)e->total;
}

// Force evaluation of the current work list: src -> out.  prep: first the
// reference velocities (mixed) and the scalar-load kernel's records of src;
// without it they must already be there (written for the prediction by the
// previous step or the prologue).  nx (a step): the directed systems'
// targets are advanced into nx by the force kernels instead of written to
// out.
int enqueue_ens_force(nbh_ensemble* e, const nbh::State& src, const nbh::Forces& out, bool prep,
                      bool allow_events, const nbh::EnsNext* nx = nullptr)
{
    nbh_ctx* c = e->ctx;
    void* rec = e->nbig > 0 ? rec_of(e, e->parity) : nullptr;
    nbh::EnsArgs a;
    a.s = src;
    a.out = out;
    a.sys = static_cast<const nbh::EnsSys*>(e->sys.p);
    a.items = static_cast<const EnsItem*>(e->items.p);
    a.nitems = e->nitems;
    a.counter = static_cast<unsigned*>(e->counter.p);
    a.partial = e->partial.d();
    a.nx = nx ? *nx : nbh::EnsNext{};
    a.bodysys = static_cast<const int*>(e->bodysys.p);
    a.vref = e->vref.d();
    nbh::EventGuard ev;
    const bool timed = c->profiling && allow_events;
    if (timed) {
        int rc = ev.take(c);
        if (rc) return rc;
        NBH_HIP(hipEventRecord(ev.ep.a, c->stream));
    }
    const bool mixed = c->cfg.precision == NBH_MIXED;
    const bool mask = e->any_mask[mixed];
    if (prep) {
        e->pred_valid = false;  // the reference velocities and records now describe src
        if (mixed) {
            nbh::k_ens_vref<<<e->nsys, nbh::kLanes, 0, c->stream>>>(a.s, a.sys, e->nsys, e->vref.d());
            NBH_HIP(hipGetLastError());
        }
    }
    // with both launches present, the persistent one runs on the side
    // stream, so that its items fill the scalar-load launch's tail
    const bool fork = e->nbig > 0 && e->nitems > 0;
    hipStream_t ps = c->stream;
    if (fork) {
        NBH_HIP(hipEventRecord(e->ev_fork, c->stream));
        NBH_HIP(hipStreamWaitEvent(e->side, e->ev_fork, 0));
        ps = e->side;
    }
    nbh::k_zero_u32<<<1, 64, 0, ps>>>(a.counter);
    NBH_HIP(hipGetLastError());
    const void* k = nbh::pick_ens_kernel(mixed, e->maxtile, mask, e->list_ipl);
    if (!k) return fail(NBH_ERR_ARG, "no ensemble kernel for tile %d", e->maxtile);
    void* args[] = {&a};
    NBH_HIP(hipLaunchKernel(k, dim3(e->grid), dim3(nbh::kLanes), args, 0, ps));
    if (e->nreduce > 0) {
        nbh::k_ens_reduce<<<e->nreduce, nbh::kLanes, 0, ps>>>(static_cast<const EnsReduce*>(e->reduces.p),
                                                              e->nreduce, e->partial.d(), out, a.s, a.nx);
        NBH_HIP(hipGetLastError());
    }
    if (e->nbig > 0) {  // the large systems' source records, then their scalar-load kernel
        const int* list = static_cast<const int*>(e->biglist.p);
        if (prep) {
            if (mixed)
                nbh::k_ens_pack<true><<<e->nbigsys, nbh::kBigBlock, 0, c->stream>>>(a.s, a.sys, list, a.vref, rec);
            else
                nbh::k_ens_pack<false><<<e->nbigsys, nbh::kBigBlock, 0, c->stream>>>(a.s, a.sys, list, a.vref, rec);
            NBH_HIP(hipGetLastError());
        }
        const auto* big = static_cast<const EnsBig*>(e->bigitems.p);
        const dim3 g(e->nbig), b(nbh::kBigBlock);
        const nbh::EnsNext none{};
        const nbh::EnsNext& n = nx ? *nx : none;
        // <MIXED, MASK, ADVANCE>
        if (mixed && mask && nx)
            nbh::k_ens_big<true, true, true><<<g, b, 0, c->stream>>>(a.s, out, a.sys, big, rec, a.vref, n);
        else if (mixed && mask)
            nbh::k_ens_big<true, true, false><<<g, b, 0, c->stream>>>(a.s, out, a.sys, big, rec, a.vref, n);
        else if (mixed && nx)
            nbh::k_ens_big<true, false, true><<<g, b, 0, c->stream>>>(a.s, out, a.sys, big, rec, a.vref, n);
        else if (mixed)
            nbh::k_ens_big<true, false, false><<<g, b, 0, c->stream>>>(a.s, out, a.sys, big, rec, a.vref, n);
        else if (mask && nx)
            nbh::k_ens_big<false, true, true><<<g, b, 0, c->stream>>>(a.s, out, a.sys, big, rec, a.vref, n);
        else if (mask)
            nbh::k_ens_big<false, true, false><<<g, b, 0, c->stream>>>(a.s, out, a.sys, big, rec, a.vref, n);
        else if (nx)
            nbh::k_ens_big<false, false, true><<<g, b, 0, c->stream>>>(a.s, out, a.sys, big, rec, a.vref, n);
        else
            nbh::k_ens_big<false, false, false><<<g, b, 0, c->stream>>>(a.s, out, a.sys, big, rec, a.vref, n);
        NBH_HIP(hipGetLastError());
    }
    if (fork) {
        NBH_HIP(hipEventRecord(e->ev_join, e->side));
        NBH_HIP(hipStreamWaitEvent(c->stream, e->ev_join, 0));
    }
    if (e->nsym > 0)
        NBH_HIP(nbh::launch_ens_sym(c->cfg.precision, a.s, out, static_cast<const nbh::SymItem*>(e->symitems.p),
                                    e->nsym, mixed ? a.vref : nullptr, c->stream));
    if (timed) {
        NBH_HIP(hipEventRecord(ev.ep.b, c->stream));
        ev.pending();
    }
    return NBH_OK;
}

nbh::State current_state(const nbh_ensemble* e) { return nbh::state_of(e->state.d(), e->total); }

// Prediction buffer q (0 or 1) as a state (masses: the current state's).
nbh::State predicted_state(const nbh_ensemble* e, int q)
{
    const size_t n = e->total;
    double* p = e->pred.d() + (size_t)q * 6 * n;
    nbh::State s = current_state(e);
    s.x = p;
    s.y = p + n;
    s.z = p + 2 * n;
    s.vx = p + 3 * n;
    s.vy = p + 4 * n;
    s.vz = p + 5 * n;
    return s;
}

// Force buffer q (0 or 1).
nbh::Forces forces_buf(const nbh_ensemble* e, int q) { return nbh::forces_of(e->force[q].d(), e->total); }

// A step that starts from buffers `from` (prediction, forces at the current
// state) reads f0 there and puts the corrected state (the current state),
// its forces and the next prediction with its records into buffers `to`.
nbh::EnsNext next_of(const nbh_ensemble* e, int from, int to)
{
    nbh::EnsNext nx;
    const size_t n = e->total;
    double* b = e->state.d();
    double* p = e->pred.d() + (size_t)to * 6 * n;  // the layout of predicted_state
    for (int k = 0; k < 3; ++k) {
        nx.cp[k] = b + k * n;
        nx.cv[k] = b + (4 + k) * n;
        nx.np[k] = p + k * n;
        nx.nv[k] = p + (3 + k) * n;
    }
    nx.m = b + 3 * n;
    nx.f0 = forces_buf(e, from);
    nx.f1 = forces_buf(e, to);
    nx.dt = e->dt.d();
    nx.rec = e->nbig > 0 ? rec_of(e, to) : nullptr;
    return nx;
}

// First prediction of a run: the current state and its forces
// (force[parity]) into pred[parity] and its records.
int enqueue_ens_prologue(nbh_ensemble* e)
{
    nbh_ctx* c = e->ctx;
    const nbh::EnsNext nx = next_of(e, e->parity, e->parity);
    const int* bs = static_cast<const int*>(e->bodysys.p);
    const dim3 g((e->total + nbh::kAdvBlock - 1) / nbh::kAdvBlock), b(nbh::kAdvBlock);
    if (c->cfg.precision == NBH_MIXED)
        nbh::k_ens_prologue<true><<<g, b, 0, c->stream>>>(nx, bs, e->vref.d(), e->total);
    else
        nbh::k_ens_prologue<false><<<g, b, 0, c->stream>>>(nx, bs, e->vref.d(), e->total);
    NBH_HIP(hipGetLastError());
    return NBH_OK;
}

// One shared-dt Hermite step (src/nbody.cpp:224-289) with the prediction in
// pred[parity]: forces at the prediction into force[1 - parity], the force
// kernels correcting and predicting the targets of the directed systems (the
// scalar-load kernel, unsplit persistent items, the split reduction),
// k_ens_advance the symmetric systems' bodies; then the buffers swap roles
// (the reference swaps its force pointers the same way, src/nbody.cpp:263-288).
int enqueue_ens_step(nbh_ensemble* e, bool allow_events)
{
    nbh_ctx* c = e->ctx;
    const int p = e->parity;
    const nbh::EnsNext nx = next_of(e, p, 1 - p);
    const nbh::State pred = predicted_state(e, p);
    const nbh::Forces f1 = forces_buf(e, 1 - p);
    int rc = enqueue_ens_force(e, pred, f1, false, allow_events, &nx);
    if (rc) return rc;
    if (e->nflat > 0) {
        const dim3 g((e->total + nbh::kAdvBlock - 1) / nbh::kAdvBlock), b(nbh::kAdvBlock);
        nbh::k_ens_advance<<<g, b, 0, c->stream>>>(nx, pred, f1, static_cast<const int*>(e->bodysys.p), e->total);
        NBH_HIP(hipGetLastError());
    }
    e->parity = 1 - p;
    return NBH_OK;
}

}  // namespace

extern "C" {

int nbh_ens_create(nbh_ctx* ctx, int nsys, const int* counts, const double* eps2, nbh_ensemble** out)
{
    if (!out) return fail(NBH_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!ctx) return fail(NBH_ERR_ARG, "context is NULL");
    if (nsys < 1 || !counts || !eps2) return fail(NBH_ERR_ARG, "need nsys >= 1 and counts/eps2 arrays");
    long long total = 0;
    for (int k = 0; k < nsys; ++k) {
        if (counts[k] < 1) return fail(NBH_ERR_ARG, "system %d has %d bodies", k, counts[k]);
        if (!(eps2[k] >= 0.0) || !std::isfinite(eps2[k]))
            return fail(NBH_ERR_ARG, "system %d: eps2 must be finite and >= 0", k);
        total += counts[k];
    }
    if (total > INT_MAX / 8) return fail(NBH_ERR_ARG, "ensemble too large (%lld bodies)", total);
    NBH_HIP(hipSetDevice(ctx->cfg.device));
    auto* e = new (std::nothrow) nbh_ensemble();
    if (!e) return fail(NBH_ERR_ALLOC, "ensemble allocation failed");
    e->ctx = ctx;
    e->sym = ctx->cfg.pairing == NBH_PAIR_SYMMETRIC;
    {
        hipError_t he = hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming);
        if (he != hipSuccess) {
            nbh_ens_destroy(e);
            return fail(NBH_ERR_HIP, "ensemble stream/event creation failed: %s", hipGetErrorString(he));
        }
    }
    e->nsys = nsys;
    e->total = (int)total;
    e->cnt.assign(counts, counts + nsys);
    e->eps2.assign(eps2, eps2 + nsys);
    e->off.resize(nsys);
    std::vector<nbh::EnsSys> sys(nsys);
    int o = 0;
    for (int k = 0; k < nsys; ++k) {
        e->off[k] = o;
        sys[k] = {o, counts[k], eps2[k]};
        o += counts[k];
        for (int p = 0; p < 2; ++p) e->any_mask[p] |= nbh::self_mask(p, eps2[k]);
    }
    const size_t nb = sizeof(double) * (size_t)total;
    if (e->sys.reserve(sizeof(nbh::EnsSys) * nsys) != hipSuccess || e->state.reserve(7 * nb) != hipSuccess ||
        e->force[0].reserve(6 * nb) != hipSuccess || e->force[1].reserve(6 * nb) != hipSuccess ||
        e->pred.reserve(12 * nb) != hipSuccess ||
        e->counter.reserve(sizeof(unsigned)) != hipSuccess || e->dt.reserve(sizeof(double)) != hipSuccess ||
        e->vref.reserve(sizeof(double) * 3 * (size_t)nsys) != hipSuccess ||
        e->ekin.reserve(sizeof(double) * nsys) != hipSuccess || e->epot.reserve(sizeof(double) * nsys) != hipSuccess) {
        nbh_ens_destroy(e);
        return fail(NBH_ERR_ALLOC, "ensemble device allocation failed (%lld bodies)", total);
    }
    hipError_t he = hipMemcpy(e->sys.p, sys.data(), sizeof(nbh::EnsSys) * nsys, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        nbh_ens_destroy(e);
        return fail(NBH_ERR_HIP, "hipMemcpy failed: %s", hipGetErrorString(he));
    }
    int rc = rebuild(e);
    if (rc) {
        nbh_ens_destroy(e);
        return rc;
    }
    *out = e;
    return NBH_OK;
}

int nbh_ens_destroy(nbh_ensemble* e)
{
    if (!e) return NBH_OK;
    (void)hipSetDevice(e->ctx->cfg.device);
    (void)hipStreamSynchronize(e->ctx->stream);
    if (e->side) (void)hipStreamSynchronize(e->side);
    drop_graph(e);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    if (e->ev_join) (void)hipEventDestroy(e->ev_join);
    if (e->side) (void)hipStreamDestroy(e->side);
    for (DevBuf* b : {&e->sys, &e->state, &e->force[0], &e->force[1], &e->items, &e->reduces, &e->partial, &e->counter,
                      &e->dt, &e->ekin, &e->epot, &e->symitems, &e->vref, &e->bigitems, &e->biglist, &e->rec,
                      &e->bodysys, &e->pred})
        b->release();
    delete e;
    return NBH_OK;
}

int nbh_ens_upload(nbh_ensemble* e, const double* x, const double* y, const double* z, const double* m,
                   const double* vx, const double* vy, const double* vz)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    const double* ins[7] = {x, y, z, m, vx, vy, vz};
    for (auto p : ins)
        if (!p) return fail(NBH_ERR_ARG, "input array is NULL");
    nbh_ctx* c = e->ctx;
    NBH_HIP(hipSetDevice(c->cfg.device));
    const size_t nb = sizeof(double) * (size_t)e->total;
    for (int k = 0; k < 7; ++k)
        NBH_HIP(hipMemcpyAsync(e->state.d() + (size_t)k * e->total, ins[k], nb, hipMemcpyHostToDevice, c->stream));
    int rc = enqueue_ens_force(e, current_state(e), forces_buf(e, e->parity), true, false);
    if (rc) return rc;
    NBH_HIP(hipStreamSynchronize(c->stream));
    e->uploaded = true;
    return NBH_OK;
}

int nbh_ens_set_dt(nbh_ensemble* e, double dt)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(NBH_ERR_ARG, "dt must be finite and > 0, got %g", dt);
    NBH_HIP(hipSetDevice(e->ctx->cfg.device));
    NBH_HIP(hipMemcpy(e->dt.p, &dt, sizeof dt, hipMemcpyHostToDevice));
    e->dt_set = true;
    e->pred_valid = false;
    return NBH_OK;
}

int nbh_ens_step(nbh_ensemble* e, int nsteps, int use_graph)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    if (nsteps < 0) return fail(NBH_ERR_ARG, "nsteps must be >= 0, got %d", nsteps);
    if (!e->uploaded) return fail(NBH_ERR_STATE, "no state: call nbh_ens_upload first");
    if (!e->dt_set) return fail(NBH_ERR_STATE, "no time step: call nbh_ens_set_dt first");
    nbh_ctx* c = e->ctx;
    NBH_HIP(hipSetDevice(c->cfg.device));
    const auto t0 = std::chrono::steady_clock::now();
    if (nsteps > 0 && !e->pred_valid) {  // the run's first prediction, outside any graph
        int rc = enqueue_ens_prologue(e);
        if (rc) return rc;
        e->pred_valid = true;
    }
    int done = 0;
    if (use_graph && nsteps > 0) {
        const int chunk = std::min(nsteps, 16);
        // a captured sequence starts from the prediction buffer it was
        // captured at (an even number of steps ends there too)
        if (!e->gexec || e->g_steps != chunk || e->g_parity != e->parity) {
            drop_graph(e);
            const int p0 = e->parity;
            NBH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = NBH_OK;
            for (int k = 0; k < chunk && rc == NBH_OK; ++k) rc = enqueue_ens_step(e, false);
            e->parity = p0;  // nothing has run yet
            hipGraph_t g = nullptr;
            hipError_t he = hipStreamEndCapture(c->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (he != hipSuccess) return fail(NBH_ERR_HIP, "graph capture failed: %s", hipGetErrorString(he));
            he = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (he != hipSuccess) {
                e->gexec = nullptr;
                return fail(NBH_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(he));
            }
            e->g_steps = chunk;
            e->g_parity = p0;
        }
        while (nsteps - done >= e->g_steps && e->parity == e->g_parity) {
            NBH_HIP(hipGraphLaunch(e->gexec, c->stream));
            done += e->g_steps;
            e->parity ^= e->g_steps & 1;
        }
    }
    for (; done < nsteps; ++done) {
        int rc = enqueue_ens_step(e, true);
        if (rc) return rc;
    }
    NBH_HIP(hipStreamSynchronize(c->stream));
    const auto t1 = std::chrono::steady_clock::now();
    c->step_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    c->steps += nsteps;
    return nbh::settle_events(c);
}

int nbh_ens_download(nbh_ensemble* e, double* x, double* y, double* z, double* m, double* vx, double* vy,
                     double* vz, double* ax, double* ay, double* az, double* jx, double* jy, double* jz)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    if (!e->uploaded) return fail(NBH_ERR_STATE, "no state: call nbh_ens_upload first");
    nbh_ctx* c = e->ctx;
    NBH_HIP(hipSetDevice(c->cfg.device));
    const size_t nb = sizeof(double) * (size_t)e->total;
    double* so[7] = {x, y, z, m, vx, vy, vz};
    double* fo[6] = {ax, ay, az, jx, jy, jz};
    for (int k = 0; k < 7; ++k)
        if (so[k])
            NBH_HIP(hipMemcpyAsync(so[k], e->state.d() + (size_t)k * e->total, nb, hipMemcpyDeviceToHost, c->stream));
    for (int k = 0; k < 6; ++k)
        if (fo[k])
            NBH_HIP(hipMemcpyAsync(fo[k], e->force[e->parity].d() + (size_t)k * e->total, nb,
                                   hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    return NBH_OK;
}

int nbh_ens_energy(nbh_ensemble* e, double* kin, double* pot)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    if (!kin || !pot) return fail(NBH_ERR_ARG, "output arrays are NULL");
    if (!e->uploaded) return fail(NBH_ERR_STATE, "no state: call nbh_ens_upload first");
    nbh_ctx* c = e->ctx;
    NBH_HIP(hipSetDevice(c->cfg.device));
    nbh::k_ens_energy<<<e->nsys, nbh::kEnBlock, 0, c->stream>>>(nbh::state_of(e->state.d(), e->total),
                                                                static_cast<const nbh::EnsSys*>(e->sys.p),
                                                                e->ekin.d(), e->epot.d());
    NBH_HIP(hipGetLastError());
    NBH_HIP(hipMemcpyAsync(kin, e->ekin.p, sizeof(double) * e->nsys, hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipMemcpyAsync(pot, e->epot.p, sizeof(double) * e->nsys, hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    return NBH_OK;
}

int nbh_ens_set_tuning(nbh_ensemble* e, const int* tiles, const int* jsplits)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    if (!tiles || !jsplits) return fail(NBH_ERR_ARG, "tiles/jsplits are NULL");
    for (int b = 0; b < NBH_ENS_BUCKETS; ++b) {
        if (tiles[b] != 0 && tiles[b] != 64 && tiles[b] != 128 && tiles[b] != 256)
            return fail(NBH_ERR_ARG, "bucket %d: tile must be 0 (scalar-load kernel), 64, 128 or 256, got %d", b,
                        tiles[b]);
        if (jsplits[b] < 1) return fail(NBH_ERR_ARG, "bucket %d: jsplit must be >= 1, got %d", b, jsplits[b]);
    }
    NBH_HIP(hipSetDevice(e->ctx->cfg.device));
    std::copy(tiles, tiles + NBH_ENS_BUCKETS, e->tile);
    std::copy(jsplits, jsplits + NBH_ENS_BUCKETS, e->split);
    return rebuild(e);
}

int nbh_ens_autotune(nbh_ensemble* e, int* tiles, int* jsplits, double* ms_out)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    if (!e->uploaded) return fail(NBH_ERR_STATE, "no state: call nbh_ens_upload first");
    nbh_ctx* c = e->ctx;
    NBH_HIP(hipSetDevice(c->cfg.device));
    hipEvent_t a, b;
    NBH_HIP(hipEventCreate(&a));
    NBH_HIP(hipEventCreate(&b));
    double best_ms[NBH_ENS_BUCKETS] = {0, 0, 0, 0};
    int lds_tile[NBH_ENS_BUCKETS], lds_split[NBH_ENS_BUCKETS];  // each bucket's best LDS-tile shape
    std::copy(e->tile, e->tile + NBH_ENS_BUCKETS, lds_tile);
    std::copy(e->split, e->split + NBH_ENS_BUCKETS, lds_split);
    for (int b = 0; b < NBH_ENS_BUCKETS; ++b)
        if (lds_tile[b] == 0) lds_tile[b] = 256, lds_split[b] = 1;
    int rc = NBH_OK;
    const bool prof = c->profiling;
    c->profiling = false;
    // min over two timed force evaluations (after one untimed) of the uploaded list
    auto time_list = [&](double* out_ms) -> int {
        float t = 0.f;
        double mn = 1e300;
        for (int rep = 0; rep < 3; ++rep) {
            if (hipEventRecord(a, c->stream) != hipSuccess) return fail(NBH_ERR_HIP, "event record failed");
            // into the force buffer not holding the current forces
            int r = enqueue_ens_force(e, current_state(e), forces_buf(e, 1 - e->parity), true, false);
            if (r) return r;
            if (hipEventRecord(b, c->stream) != hipSuccess) return fail(NBH_ERR_HIP, "event record failed");
            if (hipEventSynchronize(b) != hipSuccess) return fail(NBH_ERR_HIP, "sync failed");
            if (hipEventElapsedTime(&t, a, b) != hipSuccess) return fail(NBH_ERR_HIP, "elapsed time failed");
            if (rep > 0) mn = std::min(mn, (double)t);  // first run warms up
        }
        *out_ms = mn;
        return (int)NBH_OK;
    };
    // tile and split per bucket, that bucket's systems alone
    for (int bk = 0; bk < NBH_ENS_BUCKETS && rc == NBH_OK; ++bk) {
        std::vector<char> use(e->nsys, 0);
        bool any = false;
        for (int k = 0; k < e->nsys; ++k)
            if (nbh::bucket_of(e->cnt[k]) == bk) use[k] = 1, any = true;
        if (!any) continue;
        double best = 1e300, best_lds = 1e300;
        int bt = e->tile[bk], bs = e->split[bk];
        for (int T : nbh::kTiles) {
            for (int S : nbh::kSplits) {
                if (T == 0 && S > 1) break;  // the scalar-load kernel takes whole systems
                int tl[NBH_ENS_BUCKETS], sp[NBH_ENS_BUCKETS];
                std::copy(e->tile, e->tile + NBH_ENS_BUCKETS, tl);
                std::copy(e->split, e->split + NBH_ENS_BUCKETS, sp);
                tl[bk] = T;
                sp[bk] = S;
                WorkList w;
                build_items(e, tl, sp, e->ipl, use, w);
                double ms = 0.0;
                if ((rc = upload_items(e, w, e->ipl)) || (rc = time_list(&ms))) break;
                if (ms < best) {
                    best = ms;
                    bt = T;
                    bs = S;
                }
                if (T > 0 && ms < best_lds) {
                    best_lds = ms;
                    lds_tile[bk] = T;
                    lds_split[bk] = S;
                }
            }
            if (rc) break;
        }
        e->tile[bk] = bt;
        e->split[bk] = bs;
        best_ms[bk] = best;
    }
    // then the targets per lane of the whole list
    if (rc == NBH_OK) {
        double best = 1e300;
        int bq = e->ipl;
        for (int Q : {1, 2}) {
            e->ipl = Q;
            double ms = 0.0;
            if ((rc = rebuild(e)) || (rc = time_list(&ms))) break;
            if (ms < best) {
                best = ms;
                bq = Q;
            }
        }
        e->ipl = bq;
    }
    // the buckets were timed alone; together, the scalar-load kernel is a
    // launch of its own beside the persistent one (on a second stream), so
    // the whole list is timed three ways -- as tuned, with every tile-0
    // bucket back on its best LDS tile, and with only the largest tile-0
    // bucket kept on the scalar-load kernel -- and the fastest is kept
    int last0 = -1;
    for (int b = 0; b < NBH_ENS_BUCKETS; ++b)
        if (e->tile[b] == 0) last0 = b;
    if (rc == NBH_OK && last0 >= 0) {
        int cand_t[3][NBH_ENS_BUCKETS], cand_s[3][NBH_ENS_BUCKETS];
        for (int v = 0; v < 3; ++v)
            for (int b = 0; b < NBH_ENS_BUCKETS; ++b) {
                const bool revert = e->tile[b] == 0 && (v == 1 || (v == 2 && b != last0));
                cand_t[v][b] = revert ? lds_tile[b] : e->tile[b];
                cand_s[v][b] = revert ? lds_split[b] : e->split[b];
            }
        double best = 1e300;
        int bv = 0;
        for (int v = 0; v < 3 && rc == NBH_OK; ++v) {
            std::copy(cand_t[v], cand_t[v] + NBH_ENS_BUCKETS, e->tile);
            std::copy(cand_s[v], cand_s[v] + NBH_ENS_BUCKETS, e->split);
            double ms = 0.0;
            if ((rc = rebuild(e)) || (rc = time_list(&ms))) break;
            if (ms < best) best = ms, bv = v;
        }
        std::copy(cand_t[bv], cand_t[bv] + NBH_ENS_BUCKETS, e->tile);
        std::copy(cand_s[bv], cand_s[bv] + NBH_ENS_BUCKETS, e->split);
    }
    c->profiling = prof;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    int rc2 = rebuild(e);
    if (rc) return rc;
    if (rc2) return rc2;
    if (tiles) std::copy(e->tile, e->tile + NBH_ENS_BUCKETS, tiles);
    if (jsplits) std::copy(e->split, e->split + NBH_ENS_BUCKETS, jsplits);
    if (ms_out) std::copy(best_ms, best_ms + NBH_ENS_BUCKETS, ms_out);
    return NBH_OK;
}

int nbh_ens_set_targets_per_lane(nbh_ensemble* e, int ipl)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    if (ipl != 1 && ipl != 2) return fail(NBH_ERR_ARG, "targets per lane must be 1 or 2, got %d", ipl);
    NBH_HIP(hipSetDevice(e->ctx->cfg.device));
    e->ipl = ipl;
    return rebuild(e);
}

int nbh_ens_tuning(nbh_ensemble* e, int* tiles, int* jsplits, int* ipl)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    if (tiles) std::copy(e->tile, e->tile + NBH_ENS_BUCKETS, tiles);
    if (jsplits) std::copy(e->split, e->split + NBH_ENS_BUCKETS, jsplits);
    if (ipl) *ipl = e->ipl;
    return NBH_OK;
}

int nbh_ens_info(nbh_ensemble* e, int* items, int* workgroups, int* max_tile)
{
    if (!e) return fail(NBH_ERR_ARG, "ensemble is NULL");
    if (items) *items = e->nitems;
    if (workgroups) *workgroups = e->grid;
    if (max_tile) *max_tile = e->maxtile;
    return NBH_OK;
}

}  // extern "C"
