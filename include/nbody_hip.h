/*
 * nbody_hip.h -- C ABI of the MI355X (gfx950) Hermite-4 force backend.
 *
 * Plain C types only: host pointers and sizes, an opaque context, int status
 * codes.  libnbody_hip.so exports exactly the functions declared here.  Each
 * entry point names the reference interface it replaces
 * (jlalmerol/N-Body-Code, file:line).
 *
 * Body state is struct-of-arrays, one double[n] per field, in the order of the
 * reference's force_calculation arguments (include/kernel.h:15-19):
 *   positions x, y, z; mass m (the reference's h_par_w); velocities vx, vy, vz.
 * Forces are acceleration ax, ay, az and jerk jx, jy, jz (h_acc_*, h_adot_*).
 *
 * Errors: every int-returning call returns NBH_OK (0) or an NBH_ERR_* code;
 * nbh_last_error() then describes the most recent failure on the calling
 * thread.  A context is single-threaded: calls on one context must not
 * overlap.  Calls that take host arrays are synchronous (outputs valid on
 * return), matching the reference's blocking force_calculation.
 */
#ifndef NBODY_HIP_H
#define NBODY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nbh_status {
    NBH_OK = 0,
    NBH_ERR_ARG = 1,       /* bad argument (n, pointer, config field) */
    NBH_ERR_HIP = 2,       /* HIP runtime error (no device, launch failure) */
    NBH_ERR_ALLOC = 3,     /* device or host allocation failed */
    NBH_ERR_STATE = 4,     /* call out of order (e.g. step before upload) */
    NBH_ERR_NONFINITE = 5  /* result holds inf/NaN (e.g. eps2 = 0, coincident bodies) */
};

enum nbh_precision {
    NBH_FP64 = 0,  /* plain FP64 pair arithmetic: the drop-in */
    NBH_MIXED = 1  /* FP64 x_j - x_i rounded to FP32; FP32 r^2, rsqrt, jerk;
                      FP32 sums over runs of sources, carried on in an FP32
                      two-sum (hi, lo) pair (directed pairing) or in FP64
                      (symmetric pairing: runs of at most 192 terms) */
};

/* How pairs are evaluated.  DIRECTED: every ordered pair (i <- j) once, sums
 * in a fixed order, bit-reproducible run to run.  SYMMETRIC: every unordered
 * pair once, applied to both bodies (Newton's third law, as the reference's
 * j < i loops at src/kernel.cpp:1126,1187), sums combined with FP64 atomic
 * adds, reproducible to FP64 rounding only.  SYMMETRIC cuts the bodies into
 * blocks of 768 and is used (in both precisions) when n >= 1536; smaller n
 * run DIRECTED (nbh_launch_pairing reports which). */
enum nbh_pairing {
    NBH_PAIR_DIRECTED = 0,
    NBH_PAIR_SYMMETRIC = 1
};

/* How the directed kernels bring in their source bodies (the same sums
 * either way, bit for bit).  LDS: tiles loaded by all 64 lanes at once and
 * read back as broadcasts; best for small n, where few waves are resident to
 * hide load latency.  SCALAR: sources packed once per evaluation into
 * records in HBM and read with scalar loads into SGPRs, with no LDS and more
 * waves resident; best for large n.  AUTO: the autotuner's choice for n's
 * size bucket, else SCALAR from n >= 65536 (mixed) / 262144 (FP64). */
enum nbh_sources {
    NBH_SRC_AUTO = 0,
    NBH_SRC_LDS = 1,
    NBH_SRC_SCALAR = 2
};

typedef struct nbh_ctx nbh_ctx;

typedef struct nbh_config {
    int device;     /* HIP device ordinal */
    int precision;  /* enum nbh_precision */
    double eps2;    /* Plummer softening squared; the reference hard-codes
                       EPS2 = 1e-7 (include/kernel.h:13) */
    int block;      /* i-bodies (= threads) per workgroup, and the j-tile
                       length with LDS sources: 64, 128 or 256; 0 = pick by
                       size */
    int jsplit;     /* j-range split across workgroups; 0 = pick by size */
    int ipl;        /* targets per lane, 1 or 2; 0 = pick by size */
    int pairing;    /* enum nbh_pairing; 0 = NBH_PAIR_DIRECTED.  block,
                       jsplit, ipl and sources shape the directed kernels;
                       the symmetric kernels have a fixed shape */
    int sources;    /* enum nbh_sources; 0 = NBH_SRC_AUTO */
} nbh_config;

/* Description of the last failure on this thread ("" if none). */
const char *nbh_last_error(void);

/* ABI version of this library (major*100 + minor). */
int nbh_version(void);

/* Opens a context on cfg->device with its own stream.  Replaces the
 * per-call device open/buffer/kernel set-up of the TT dispatch
 * (src/kernel.cpp:54-243): here it happens once. */
int nbh_create(const nbh_config *cfg, nbh_ctx **out);
int nbh_destroy(nbh_ctx *ctx);

/* Changes softening / precision / tiling of an existing context. */
int nbh_configure(nbh_ctx *ctx, const nbh_config *cfg);

/* Drop-in force evaluation: acceleration and jerk of all n bodies from all
 * n bodies, self-pair excluded.  Replaces force_calculation
 * (include/kernel.h:15-19; CPU FP64 src/kernel.cpp:1164-1224, _F32
 * src/kernel.cpp:1101-1163, TT dispatch src/kernel.cpp:47-398).  Host
 * arrays; outputs fully overwritten; synchronous.  kernel_ms (may be NULL)
 * receives the device time of the pair kernels, the analogue of
 * Sys::kernel_time (include/utils.h:42).  Returns NBH_ERR_NONFINITE if an
 * output is inf/NaN (outputs are still written). */
int nbh_force(nbh_ctx *ctx, int n,
              const double *x, const double *y, const double *z,
              const double *m, const double *vx, const double *vy,
              const double *vz,
              double *ax, double *ay, double *az,
              double *jx, double *jy, double *jz, double *kernel_ms);

/* Same evaluation on device-resident arrays (HBM pointers), enqueued on
 * `stream` (a hipStream_t; NULL = the context's own stream), asynchronous.
 * No counterpart in the reference (its arrays are host memory); it lets a
 * caller that already holds device buffers skip the copies.  With
 * NBH_PAIR_SYMMETRIC the outputs receive FP64 atomic adds, so they must be
 * ordinary device allocations (hipMalloc), not host or managed memory.
 * The context's j-split partials and packed source records are shared by
 * all of its calls: two evaluations that may run at the same time (on
 * different streams) need two contexts. */
int nbh_force_device(nbh_ctx *ctx, int n,
                     const double *x, const double *y, const double *z,
                     const double *m, const double *vx, const double *vy,
                     const double *vz,
                     double *ax, double *ay, double *az,
                     double *jx, double *jy, double *jz, void *stream);

/* nbh_force_device without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory during the call (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, semantics and kernels as nbh_force_device,
 * except that the evaluation's working memory (reference-velocity reduction,
 * packed source records, j-split partial sums) lives in a static device
 * array of 64 MiB that each GPU the library runs on holds from module load
 * to process end, instead of in context buffers; the call makes no
 * hipMalloc/hipFree or other allocation.  When the context's launch shape
 * for n needs more than 64 MiB, the call narrows it (fewer j-splits, then
 * LDS-tile sources, which need none of it); the result then equals
 * nbh_force_device's to FP64 rounding of the split sums rather than bit for
 * bit.  A j-split or source path forced by the config is kept, and a call
 * whose forced shape does not fit fails with NBH_ERR_ARG.  Every call on one
 * GPU shares that GPU's array: calls that may overlap on one GPU (other
 * streams, other contexts, other threads) race and must be serialised by
 * the caller.  Profiling events are not recorded. */
int nbh_force_device_noalloc(nbh_ctx *ctx, int n,
                             const double *x, const double *y, const double *z,
                             const double *m, const double *vx, const double *vy,
                             const double *vz,
                             double *ax, double *ay, double *az,
                             double *jx, double *jy, double *jz, void *stream);

/* ---- device-resident integrator (NBody, include/nbody.h:24-59) ------ */

/* Copies the state to HBM and evaluates the initial force
 * (NBody::cycle's first force_calculation, src/nbody.cpp:310). */
int nbh_upload(nbh_ctx *ctx, int n,
               const double *x, const double *y, const double *z,
               const double *m, const double *vx, const double *vy,
               const double *vz);

/* NBody::initialdt (src/nbody.cpp:131-150): min over bodies of
 * eta0*|a|/|jerk| from the resident forces. */
int nbh_initial_dt(nbh_ctx *ctx, double eta0, double *dt);

/* Time-step control.  adaptive = 0: every step uses dt.  adaptive = 1: the
 * first step uses dt and each later step the Aarseth value of
 * NBody::updatedt (src/nbody.cpp:198-222, parameter eta). */
int nbh_set_dt(nbh_ctx *ctx, double dt, int adaptive, double eta);
int nbh_get_dt(nbh_ctx *ctx, double *dt);

/* nsteps shared-step Hermite steps, NBody::timestepping
 * (src/nbody.cpp:224-289): predict (:152-163), force at the predicted state
 * (:232), snap/crackle + correct (:165-196), [updatedt (:261)], swap
 * (:263-288).  use_graph != 0 replays the step sequence from a captured
 * hipGraph.  Synchronous. */
int nbh_step(nbh_ctx *ctx, int nsteps, int use_graph);

/* Copies the resident state and current forces back; any pointer may be
 * NULL to skip that array. */
int nbh_download(nbh_ctx *ctx,
                 double *x, double *y, double *z, double *m,
                 double *vx, double *vy, double *vz,
                 double *ax, double *ay, double *az,
                 double *jx, double *jy, double *jz);

/* Energy of the resident state in FP64.  compat = 0: kinetic and the
 * physical softened potential -sum_{i<j} m_i m_j / sqrt(r^2 + eps2).
 * compat = 1: compute_kinetic / compute_potential exactly as the reference
 * defines them (src/utils.cpp:125-153: weight m_j*m_j, eps2 = 1e-7). */
int nbh_energy(nbh_ctx *ctx, int compat, double *kinetic, double *potential);

/* The same on host arrays (copied to the device for the evaluation). */
int nbh_energy_host(nbh_ctx *ctx, int n,
                    const double *x, const double *y, const double *z,
                    const double *m, const double *vx, const double *vy,
                    const double *vz, int compat,
                    double *kinetic, double *potential);

/* Per-kernel timing.  With profiling on, every pair-kernel launch is
 * bracketed by HIP events on the stream it runs on; nbh_timing sums them.
 * step_ms / steps cover nbh_step calls (host wall clock, synchronous). */
int nbh_set_profiling(nbh_ctx *ctx, int on);
int nbh_timing(nbh_ctx *ctx, double *force_ms, int64_t *force_launches,
               double *step_ms, int64_t *steps);
int nbh_reset_timing(nbh_ctx *ctx);

/* Launch shape actually used for n bodies: block (threads per workgroup),
 * jsplit, workgroups.  For the symmetric kernel, block x targets per lane is
 * its block of bodies and workgroups includes the block-diagonal pass. */
int nbh_launch_shape(nbh_ctx *ctx, int n, int *block, int *jsplit,
                     int *workgroups);
/* Targets per lane of that shape (directed: 1 or 2; symmetric: 3). */
int nbh_launch_ipl(nbh_ctx *ctx, int n, int *ipl);
/* Pairing (enum nbh_pairing) that a force evaluation of n bodies uses. */
int nbh_launch_pairing(nbh_ctx *ctx, int n, int *pairing);
/* Source path of that shape: NBH_SRC_LDS or NBH_SRC_SCALAR (directed), 0
 * for the symmetric kernel. */
int nbh_launch_sources(nbh_ctx *ctx, int n, int *sources);

/* Autotuner of the single-system force: times block {64,128,256} x targets
 * per lane {1,2} x j-split {1..32} x sources {LDS, SCALAR} (only the
 * config's, when it sets one) for n bodies on this context's device and
 * precision (a Plummer sphere, on a source range of max(min(n, 65536), n/16)
 * bodies so that a trial is short; when that range is truncated, the three
 * fastest shapes, and the fastest one with half and twice its j-split, are
 * timed again on all n sources) and keeps the fastest
 * for n's size bucket (floor(log2 n)) and precision.  Later evaluations in
 * that bucket use it unless the config forces block/jsplit/ipl (and its
 * source path unless the config sets one).  Directed
 * pairing only: the symmetric kernel has a fixed shape.  Outputs may be
 * NULL; ms is the winner's last timed launch. */
int nbh_autotune(nbh_ctx *ctx, int n, int *block, int *ipl, int *jsplit,
                 double *ms);

/* Device properties: compute units and wave size. */
int nbh_device_info(int device, int *compute_units, int *wave_size,
                    char *name, int name_len);

/* ---- batched ensemble of independent systems ------------------------- */
/* No counterpart in the reference (it integrates one system).  Thousands of
 * independent clusters of uneven size share one device state: the body
 * arrays are the systems' arrays concatenated in system order.  The force
 * evaluation is one persistent launch: one-wave workgroups take work items
 * (system, block of 64 or 128 targets, source range) from a device queue,
 * largest first; a size bucket's work items use the LDS tile length and
 * j-split chosen for that bucket, and all items one or two targets per lane
 * (by nbh_ens_autotune or nbh_ens_set_tuning / nbh_ens_set_targets_per_lane).
 * Precision and device come from the context.  When the context was created
 * with NBH_PAIR_SYMMETRIC, systems of at least two blocks (blocks of 256
 * bodies in mixed precision, 512 in FP64) are evaluated with symmetric
 * pairing instead: the persistent launch covers only their diagonal blocks,
 * and a second launch adds every block pair once (FP64 atomics, as in
 * nbh_force).  The step is the shared-dt Hermite step
 * of nbh_step applied to every system. */

typedef struct nbh_ensemble nbh_ensemble;

/* Size buckets (by bodies per system): n <= 64, <= 256, <= 1024, larger. */
#define NBH_ENS_BUCKETS 4

/* counts[k] bodies (>= 1) and softening eps2[k] (>= 0) of system k. */
int nbh_ens_create(nbh_ctx *ctx, int nsys, const int *counts,
                   const double *eps2, nbh_ensemble **out);
int nbh_ens_destroy(nbh_ensemble *ens);

/* Copies the concatenated state (sum(counts) bodies per array) to HBM and
 * evaluates the initial forces. */
int nbh_ens_upload(nbh_ensemble *ens, const double *x, const double *y,
                   const double *z, const double *m, const double *vx,
                   const double *vy, const double *vz);
int nbh_ens_set_dt(nbh_ensemble *ens, double dt);
/* nsteps Hermite steps of every system; use_graph != 0 replays a captured
 * hipGraph of the step sequence.  Synchronous. */
int nbh_ens_step(nbh_ensemble *ens, int nsteps, int use_graph);
/* Concatenated state and current forces; NULL skips an array. */
int nbh_ens_download(nbh_ensemble *ens, double *x, double *y, double *z,
                     double *m, double *vx, double *vy, double *vz,
                     double *ax, double *ay, double *az, double *jx,
                     double *jy, double *jz);
/* Per-system kinetic and softened potential energy (FP64), arrays [nsys]. */
int nbh_ens_energy(nbh_ensemble *ens, double *kinetic, double *potential);
/* Times the candidate (tile in {0,64,128,256}) x (jsplit in {1,2,4,8}) for
 * every populated bucket on the device and keeps the fastest, then one or
 * two targets per lane for the whole work list.  Outputs [NBH_ENS_BUCKETS]
 * each, may be NULL; ms = best force time of the bucket alone (0 for empty
 * buckets). */
int nbh_ens_autotune(nbh_ensemble *ens, int *tiles, int *jsplits, double *ms);
/* Sets the per-bucket tile (64, 128 or 256) and jsplit (>= 1) directly.
 * Tile 0 evaluates the bucket's systems (those not taken by symmetric
 * pairing) with the scalar-load kernel instead: one four-wave workgroup per
 * 256 targets of a system against all its sources, read from records packed
 * once per evaluation at wave-uniform indices (scalar loads, no LDS); its
 * jsplit is ignored.  Either path sums each target's sources in the same
 * order, so with jsplit 1 the forces are bitwise the same. */
int nbh_ens_set_tuning(nbh_ensemble *ens, const int *tiles, const int *jsplits);
/* Sets the targets per lane (1 or 2) of every work item. */
int nbh_ens_set_targets_per_lane(nbh_ensemble *ens, int ipl);
/* Current per-bucket tile and jsplit ([NBH_ENS_BUCKETS] each) and targets
 * per lane; any pointer may be NULL. */
int nbh_ens_tuning(nbh_ensemble *ens, int *tiles, int *jsplits, int *ipl);
/* Work items per force evaluation, persistent workgroups, largest tile. */
int nbh_ens_info(nbh_ensemble *ens, int *items, int *workgroups, int *max_tile);

/* ---- reverse-mode derivatives (FP64) ---------------------------------- */
/* No counterpart in the reference.  Arrays are blocked: a state is [7][n]
 * (x, y, z, m, vx, vy, vz blocks, the order of force_calculation's
 * arguments), forces and their cotangents are [6][n] (ax, ay, az, jx, jy,
 * jz).  The context must be FP64; its eps2 is the softening. */

/* Vector-Jacobian product of the force evaluation (the pair sum of
 * src/kernel.cpp:1164-1224): grad = d/d(state) of
 * sum_i cot_a,i . a_i + cot_j,i . jerk_i, as a [7][n] array. */
int nbh_force_vjp(nbh_ctx *ctx, int n, const double *state, const double *cot,
                  double *grad);

/* Adjoint of a fixed-dt Hermite trajectory (NBody::timestepping,
 * src/nbody.cpp:224-289, repeated).  The forward run keeps the carried
 * state (positions, velocities, forces) every `every` steps; the backward
 * sweep recomputes each segment from its checkpoint and runs the step
 * adjoints in reverse, then the adjoint of the initial force evaluation. */
typedef struct nbh_adjoint nbh_adjoint;
int nbh_adj_create(nbh_ctx *ctx, int n, int every, nbh_adjoint **out);
int nbh_adj_destroy(nbh_adjoint *adj);
/* state0 [7][n] -> state_final [7][n] (may be NULL) after nsteps steps. */
int nbh_adj_forward(nbh_adjoint *adj, const double *state0, double dt,
                    int nsteps, double *state_final);
/* cot_final [7][n]: cotangent of the final state (the m block adds to the
 * mass gradient directly) -> grad0 [7][n]: gradient w.r.t. state0. */
int nbh_adj_backward(nbh_adjoint *adj, const double *cot_final, double *grad0);

/* ---- initial conditions (host code, no device needed) --------------- */

/* Equal-mass Plummer sphere in Henon units (G = M = 1, E = -1/4):
 * Aarseth-Henon-Wielen radii by inverse CDF with the mass fraction
 * truncated at 0.999, velocities by von Neumann rejection, centre-of-mass
 * frame.  PRNG: std::mt19937_64(seed), u = (draw >> 11) * 2^-53.
 * No counterpart in the reference (its only generator is NBody::initialize,
 * below); used for the benchmark configurations. */
int nbh_plummer(int n, uint64_t seed,
                double *x, double *y, double *z, double *m,
                double *vx, double *vy, double *vz);

/* NBody::initialize (src/nbody.cpp:69-114): the reference CLI's initial
 * conditions (std::mt19937(42), uniform sphere of radius rmax, uniform
 * velocity cube, power-law masses, virial rescale), bit-identical. */
int nbh_reference_ic(int n, double vmax, double rmax, double mmin,
                     double mmax, double *x, double *y, double *z, double *m,
                     double *vx, double *vy, double *vz);

/* ---- body-state files (host code) ------------------------------------- */
/* No counterpart in the reference (it has no state I/O).  Little-endian
 * binary: a 64-byte header -- char magic[8] = "NBHSTATE", uint32 version = 1,
 * uint32 reserved = 0, int64 n, double time, int64 step, 24 zero bytes --
 * then x, y, z, m, vx, vy, vz as n doubles each. */
int nbh_state_save(const char *path, int n, const double *x, const double *y,
                   const double *z, const double *m, const double *vx,
                   const double *vy, const double *vz, double time,
                   int64_t step);
/* Reads the header: body count, time and step of the file. */
int nbh_state_info(const char *path, int *n, double *time, int64_t *step);
/* Reads the bodies; n must equal the file's body count. */
int nbh_state_load(const char *path, int n, double *x, double *y, double *z,
                   double *m, double *vx, double *vy, double *vz);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_HIP_H */
