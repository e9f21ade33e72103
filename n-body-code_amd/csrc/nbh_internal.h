// nbh_internal.h -- shared between the kernels (nbh_kernels.hip) and the
// C-ABI layer (nbh_api.hip).  Not installed; the public surface is
// include/nbody_hip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace nbh {

// Body state in HBM: seven SoA double[n] arrays (x, y, z, m, vx, vy, vz).
struct State {
    const double* x;
    const double* y;
    const double* z;
    const double* m;
    const double* vx;
    const double* vy;
    const double* vz;
};

// Six SoA output arrays (ax, ay, az, jx, jy, jz).
struct Forces {
    double* c[6];
};

// Whether a force evaluation at softening eps2 must drop the self pair
// explicitly (the MASK kernels).  With d = u = 0 the self pair adds
// m r^3 * 0 = 0 as long as r^3 = eps2^-3/2 and m r^3 are finite; below these
// thresholds they could overflow the pair arithmetic's type (FP32: eps2 <
// ~2e-26 for m = 1, and (float)eps2 rounds to 0 below 1.4e-45; FP64: eps2 <
// ~3e-206), and inf * 0 would put NaN into every target.  The thresholds
// keep m r^3 finite for masses up to 3e8 (FP32) and 1e38 (FP64).
inline bool self_mask(int precision, double eps2)
{
    return precision == 0 ? eps2 < 1e-180 : (float)eps2 < 1e-20f;
}

// Scratch of the reference-velocity reduction (doubles): partial sums of up
// to vref_parts() workgroups, then the reference velocity itself at
// vref_offset().  Mixed precision only.
constexpr int kVrefParts = 256;
constexpr size_t vref_work_doubles() { return 4 * (size_t)kVrefParts + 4; }
constexpr size_t vref_offset() { return 4 * (size_t)kVrefParts; }
// Centre-of-mass velocity of bodies 0..n-1 (sum m v / sum m; 0 when every
// mass is 0) into work[vref_offset() .. +3), summed in a fixed order
// (deterministic).  Graph-safe: kernels only.
hipError_t launch_vref(const State& s, int n, double* work, hipStream_t st);

// One force evaluation's launch shape.
struct Shape {
    int block;   // lanes per workgroup (= the j-tile length of the LDS kernels): 64, 128 or 256
    int ipl;     // targets per lane (1 or 2): block*ipl targets per workgroup
    int jsplit;  // workgroups along j; partial sums reduced in a second pass
    int jchunk;  // j-bodies per split (multiple of block)
    int sym_chunk;  // > 0: symmetric kernel, block pairs per workgroup
    int scalar;  // directed kernels: 1 = sources through scalar loads from packed records, 0 = LDS tiles
};

// Pair-kernel launch: writes forces for all n targets from all n sources
// (from the first nj sources when 0 < nj < n: the autotuner's short trial
// launches).  When shape.jsplit > 1, `partial` must hold 6*jsplit*n doubles
// and the split-reduction kernel is launched after the pair kernel, unless
// reduce_splits is false: then the partial planes (component c, split k at
// (c*jsplit + k)*n) are left for the caller to add in split order.  ev0/ev1
// (may be null) are recorded around the pair kernel only.
// rec: with shape.scalar, a device buffer of source_record_bytes(precision,
// n) bytes for the packed source records (unused, may be null, otherwise).
size_t source_record_bytes(int precision, int n);
// Packs sources 0..n-1 of s into rec (mixed: velocities less vref, the 3
// device doubles of the reference velocity).
hipError_t pack_source_records(int precision, const State& s, int n, const double* vref, void* rec,
                               hipStream_t st);
// vwork: mixed precision, vref_work_doubles() doubles of device memory
// for the reference velocity (computed first, from the n targets); unused
// (may be null) in FP64.
hipError_t launch_force(int precision, const State& s, int n, double eps2,
                        const Shape& shape, const Forces& out,
                        double* partial, void* rec, double* vwork, hipStream_t st, hipEvent_t ev0,
                        hipEvent_t ev1, int nj = 0, bool reduce_splits = true);

// Directed pair kernel restricted to the diagonal blocks: target i gets the
// sources of its own block [B*(i/B), B*(i/B)+B), self-pair excluded; the
// outputs are overwritten (the last block may be partial).  B % 256 == 0.
// vref: the reference velocity (3 doubles on the device; mixed only).
hipError_t launch_force_blockdiag(int precision, const State& s, int n, double eps2, int B,
                                  const Forces& out, const double* vref, hipStream_t st);

// Symmetric pair kernels (nbh_sym.hip): block-diagonal pass, then every
// off-diagonal block pair once with the reaction added to the source side.
// Usable when sym_supported(precision, n); ev0/ev1 (may be null) bracket
// both launches.
int sym_block(int precision);             // bodies per block
int sym_targets_per_lane(int precision);  // block = 64 * targets per lane * waves
bool sym_supported(int precision, int n);
// Entries in the source list of full block P of nr (the last entry is the
// partial block, if there is one).
__host__ __device__ inline int sym_list_len(int nr, int P, bool partial)
{
    return (nr - 1) / 2 + ((nr % 2 == 0 && P < nr / 2) ? 1 : 0) + (partial ? 1 : 0);
}
int sym_list_max(int precision, int n);  // longest source list of a full block
int sym_chunk(int precision, int n, int cus);
long sym_workgroups(int precision, int n, int chunk);
hipError_t launch_force_sym(int precision, const State& s, int n, double eps2,
                            int chunk, const Forces& out, const double* vref, hipStream_t st,
                            hipEvent_t ev0, hipEvent_t ev1);

// The ensemble's symmetric kernel cuts systems into blocks of its own size.
int ens_sym_block(int precision);
bool ens_sym_supported(int precision, int n);

// One work item of the list-driven symmetric kernel (the ensemble): full
// block P of the system occupying [off, off + n) against entries i0 .. i1 of
// its source list (blocks P+1 .. P+H mod nr, then the partial block if any),
// with the system's eps2.  The outputs must already hold the block-diagonal
// sums: the items add to them atomically.  sys indexes the per-system
// reference velocities vref[3 * sys ..] (mixed only).
struct SymItem {
    int off, n, P, i0, i1, sys;
    double eps2;
};
hipError_t launch_ens_sym(int precision, const State& s, const Forces& out,
                          const SymItem* items, int nitems, const double* vref, hipStream_t st);

// Hermite predictor in place (x, v) from forces f0 with dt read from *dt.
hipError_t launch_predict(int n, const double* dt, double* x, double* y,
                          double* z, double* vx, double* vy, double* vz,
                          const Forces& f0, hipStream_t st);

// Snap/crackle + corrector in place; then f0 <- f1.  With dt_next non-null
// the Aarseth candidate (eta) is min-reduced into *dt_next (as u64 bits).
hipError_t launch_correct(int n, const double* dt, double* x, double* y,
                          double* z, double* vx, double* vy, double* vz,
                          const Forces& f0, const Forces& f1,
                          unsigned long long* dt_next, double eta,
                          hipStream_t st);

// *dt = bits(*dt_next); *dt_next = +inf bits.
hipError_t launch_dt_advance(double* dt, unsigned long long* dt_next,
                             hipStream_t st);

// *out = bits(min_i eta0*|a_i|/|j_i|), *out pre-set by the caller.
hipError_t launch_initial_dt(int n, const Forces& f, double eta0,
                             unsigned long long* out, hipStream_t st);

// Fills *p with +inf bits (a kernel, so it is safe inside graph capture).
hipError_t launch_fill_inf(unsigned long long* p, hipStream_t st);

// Energy: out[0] = kinetic, out[1] = potential.  `work` holds
// 2*energy_blocks(n) doubles.
int energy_blocks(int n);
hipError_t launch_energy(const State& s, int n, double eps2, int compat,
                         double* work, double* out, hipStream_t st);

// Sets *flag to 1 if any of the 6n force values is not finite.
hipError_t launch_check_finite(int n, const Forces& f, int* flag,
                               hipStream_t st);

}  // namespace nbh
