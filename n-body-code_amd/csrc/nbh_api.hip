// nbh_api.hip -- the C ABI of include/nbody_hip.h: contexts, device
// buffers, the drop-in force call and the device-resident integrator.
//
// A context owns one HIP stream (default priority) on one device and keeps
// its device buffers across calls, growing them when n grows; the
// reference's TT dispatch instead opened the device and built buffers on
// every force_calculation (src/kernel.cpp:54-138, :397).
#include "nbh_ctx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace nbh {

namespace {
thread_local std::string g_err;
}

// Scratch of nbh_force_device_noalloc: the reference-velocity reduction,
// packed source records and j-split partial sums of one evaluation, in a
// file-scope device array instead of context buffers, so that the call
// allocates nothing.  Every GPU the library runs on holds its own copy,
// from module load to process end.
constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB
__device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

// Address of this device's copy of `scratch` (the current device must be
// `device`), looked up once per device.  The array's name is a host-side
// handle only; kernels and copies must be given this address.
int scratch_setup(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(NBH_ERR_ARG, "device %d beyond the scratch table (%d devices)", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(NBH_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                        hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return NBH_OK;
}
}  // namespace

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char* last_error() { return g_err.c_str(); }

State state_of(double* base, int64_t n)
{
    State s;
    s.x = base;
    s.y = base + n;
    s.z = base + 2 * n;
    s.m = base + 3 * n;
    s.vx = base + 4 * n;
    s.vy = base + 5 * n;
    s.vz = base + 6 * n;
    return s;
}

Forces forces_of(double* base, int64_t n)
{
    Forces f;
    for (int c = 0; c < 6; ++c) f.c[c] = base + c * n;
    return f;
}

int get_events(nbh_ctx* c, EventPair* ep)
{
    if (!c->ev_free.empty()) {
        *ep = c->ev_free.back();
        c->ev_free.pop_back();
        return NBH_OK;
    }
    NBH_HIP(hipEventCreate(&ep->a));
    const hipError_t e = hipEventCreate(&ep->b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(ep->a);
        return fail(NBH_ERR_HIP, "hipEventCreate failed: %s", hipGetErrorString(e));
    }
    return NBH_OK;
}

// Adds the elapsed time of every recorded pair (stream must be idle).
int settle_events(nbh_ctx* c)
{
    for (auto& ep : c->ev_pending) {
        float ms = 0.f;
        NBH_HIP(hipEventElapsedTime(&ms, ep.a, ep.b));
        c->force_ms += ms;
        c->force_launches += 1;
        c->ev_free.push_back(ep);
    }
    c->ev_pending.clear();
    return NBH_OK;
}

}  // namespace nbh

using nbh::DevBuf;
using nbh::EventPair;
using nbh::fail;
using nbh::forces_of;
using nbh::get_events;
using nbh::settle_events;
using nbh::state_of;

namespace {

int size_bucket(int n)
{
    int b = 0;
    while ((2 << b) <= n && b < 31) ++b;
    return b;
}

// Source path when neither the config nor the autotuner sets one: scalar
// loads once enough waves are resident to hide their latency (measured: the
// LDS kernels win below, e.g. 2.5x at n = 1024; the scalar-load ones above,
// by 12% (mixed) and 2.6% (FP64) at n = 2^20).
bool default_scalar(int precision, int n) { return n >= (precision == NBH_MIXED ? 65536 : 262144); }

// Shape with the given block / targets-per-lane / requested split / source
// path; the split is clamped to the source tiles and rounded to whole tiles
// per split.
nbh::Shape make_shape(int n, int block, int ipl, int split, bool scalar, int nj = 0)
{
    nbh::Shape s{};
    s.block = block;
    s.ipl = ipl;
    s.scalar = scalar ? 1 : 0;
    const int src = nj > 0 ? nj : n;
    const int jtiles = (src + block - 1) / block;
    split = std::max(1, std::min(split, jtiles));
    const int tiles_per = (jtiles + split - 1) / split;
    s.jchunk = tiles_per * block;
    s.jsplit = (src + s.jchunk - 1) / s.jchunk;
    return s;
}

nbh::Shape pick_shape(const nbh_ctx* c, int n)
{
    const nbh_config& f = c->cfg;
    if (f.pairing == NBH_PAIR_SYMMETRIC && nbh::sym_supported(f.precision, n)) {
        nbh::Shape s{};
        s.ipl = nbh::sym_targets_per_lane(f.precision);
        s.block = nbh::sym_block(f.precision) / s.ipl;  // threads per workgroup
        s.jsplit = 1;
        s.jchunk = n;
        s.sym_chunk = nbh::sym_chunk(f.precision, n, c->cus);
        return s;
    }
    const bool forced_src = f.sources != NBH_SRC_AUTO;
    if (f.block == 0 && f.jsplit == 0 && f.ipl == 0) {
        const auto& t = c->tuned[f.precision][size_bucket(n)];
        if (t.block)
            return make_shape(n, t.block, t.ipl, t.jsplit, forced_src ? f.sources == NBH_SRC_SCALAR : t.scalar != 0);
    }
    const bool scalar = forced_src ? f.sources == NBH_SRC_SCALAR : default_scalar(f.precision, n);
    const int block = f.block ? f.block : (n >= 131072 ? 256 : (n >= 16384 ? 128 : 64));
    const int ipl = f.ipl ? f.ipl : 1;
    int split = f.jsplit;
    if (split <= 0) {
        // enough workgroups that the last partial round of resident
        // workgroups is a small fraction of the launch (~128 per CU)
        const int iblocks = (n + block * ipl - 1) / (block * ipl);
        const int want = 128 * (c->cus > 0 ? c->cus : 256);
        split = std::min(64, (want + iblocks - 1) / iblocks);
    }
    return make_shape(n, block, ipl, split, scalar);
}

int check_cfg(const nbh_config* cfg)
{
    if (!cfg) return fail(NBH_ERR_ARG, "config is NULL");
    if (cfg->precision != NBH_FP64 && cfg->precision != NBH_MIXED)
        return fail(NBH_ERR_ARG, "precision must be NBH_FP64 or NBH_MIXED, got %d", cfg->precision);
    if (!(cfg->eps2 >= 0.0) || !std::isfinite(cfg->eps2))
        return fail(NBH_ERR_ARG, "eps2 must be finite and >= 0, got %g", cfg->eps2);
    if (cfg->block != 0 && cfg->block != 64 && cfg->block != 128 && cfg->block != 256)
        return fail(NBH_ERR_ARG, "block must be 0, 64, 128 or 256, got %d", cfg->block);
    if (cfg->jsplit < 0) return fail(NBH_ERR_ARG, "jsplit must be >= 0, got %d", cfg->jsplit);
    if (cfg->ipl < 0 || cfg->ipl > 2) return fail(NBH_ERR_ARG, "ipl must be 0, 1 or 2, got %d", cfg->ipl);
    if (cfg->pairing != NBH_PAIR_DIRECTED && cfg->pairing != NBH_PAIR_SYMMETRIC)
        return fail(NBH_ERR_ARG, "pairing must be NBH_PAIR_DIRECTED or NBH_PAIR_SYMMETRIC, got %d", cfg->pairing);
    if (cfg->sources != NBH_SRC_AUTO && cfg->sources != NBH_SRC_LDS && cfg->sources != NBH_SRC_SCALAR)
        return fail(NBH_ERR_ARG, "sources must be NBH_SRC_AUTO, NBH_SRC_LDS or NBH_SRC_SCALAR, got %d", cfg->sources);
    return NBH_OK;
}

void drop_graph(nbh_ctx* c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->g_steps = 0;
    c->g_n = -1;
}

// The split partials and packed source records that a force evaluation of
// shape sh needs; done before a graph capture too, so that no allocation
// happens while the stream is captured.
int reserve_force_buffers(nbh_ctx* c, const nbh::Shape& sh, int n)
{
    if (sh.jsplit > 1) NBH_HIP(c->partial.reserve(sizeof(double) * 6 * (size_t)sh.jsplit * n));
    if (sh.scalar) NBH_HIP(c->rec.reserve(nbh::source_record_bytes(c->cfg.precision, n)));
    return NBH_OK;
}

// One force evaluation on device arrays, enqueued on st.
int enqueue_force(nbh_ctx* c, const nbh::State& s, int n, const nbh::Forces& out,
                  hipStream_t st, bool allow_events)
{
    const nbh::Shape sh = pick_shape(c, n);
    if (int rc = reserve_force_buffers(c, sh, n)) return rc;
    double* part = sh.jsplit > 1 ? c->partial.d() : nullptr;
    EventPair ep;
    const bool timed = c->profiling && allow_events;
    if (timed) {
        int rc = get_events(c, &ep);
        if (rc) return rc;
    }
    hipError_t e = nbh::launch_force(c->cfg.precision, s, n, c->cfg.eps2, sh, out, part, c->rec.p, c->vwork.d(), st,
                                     timed ? ep.a : nullptr, timed ? ep.b : nullptr);
    if (timed) (e == hipSuccess ? c->ev_pending : c->ev_free).push_back(ep);
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "force launch failed: %s", hipGetErrorString(e));
    return NBH_OK;
}

bool bad_n(int n) { return n < 1; }

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Static-scratch bytes a force evaluation of shape sh needs: the
// reference-velocity reduction, then (directed kernels) the packed source
// records and the j-split partials, each 256-byte aligned.
size_t scratch_need(int precision, int n, const nbh::Shape& sh)
{
    size_t need = align256(sizeof(double) * nbh::vref_work_doubles());
    if (sh.sym_chunk > 0) return need;
    if (sh.scalar) need += align256(nbh::source_record_bytes(precision, n));
    if (sh.jsplit > 1) need += sizeof(double) * 6 * (size_t)sh.jsplit * n;
    return need;
}

// Narrows sh until it fits the static scratch: fewer j-splits, then LDS
// sources (which need no records; with one split the evaluation needs only
// the reduction's few KiB).  A j-split or source path that the config
// forces is not changed: if it does not fit, the call is rejected.
int fit_scratch(const nbh_ctx* c, int n, nbh::Shape& sh)
{
    const int prec = c->cfg.precision;
    if (scratch_need(prec, n, sh) <= nbh::kScratchBytes) return NBH_OK;
    const bool keep_split = c->cfg.jsplit > 0, keep_src = c->cfg.sources != NBH_SRC_AUTO;
    for (int scalar = sh.scalar; scalar >= (keep_src ? sh.scalar : 0); --scalar) {
        for (int split = sh.jsplit; split >= (keep_split ? sh.jsplit : 1); split = split > 1 ? split / 2 : 0) {
            const nbh::Shape t = make_shape(n, sh.block, sh.ipl, split, scalar != 0);
            if (scratch_need(prec, n, t) <= nbh::kScratchBytes) {
                sh = t;
                return NBH_OK;
            }
        }
    }
    return fail(NBH_ERR_ARG,
                "n=%d with the configured %s needs %zu bytes of scratch; nbh_force_device_noalloc has %zu "
                "(leave jsplit and sources automatic, or use nbh_force_device)",
                n, keep_split && keep_src ? "j-split and source path" : (keep_split ? "j-split" : "source path"),
                scratch_need(prec, n, sh), nbh::kScratchBytes);
}

}  // namespace

extern "C" {

const char* nbh_last_error(void) { return nbh::last_error(); }

int nbh_version(void) { return 104; }

int nbh_device_info(int device, int* compute_units, int* wave_size, char* name, int name_len)
{
    int count = 0;
    NBH_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(NBH_ERR_ARG, "device %d out of range (%d devices)", device, count);
    hipDeviceProp_t p;
    NBH_HIP(hipGetDeviceProperties(&p, device));
    if (compute_units) *compute_units = p.multiProcessorCount;
    if (wave_size) *wave_size = p.warpSize;
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s", p.gcnArchName);
    }
    return NBH_OK;
}

int nbh_create(const nbh_config* cfg, nbh_ctx** out)
{
    if (!out) return fail(NBH_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_cfg(cfg);
    if (rc) return rc;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(NBH_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (cfg->device < 0 || cfg->device >= count)
        return fail(NBH_ERR_ARG, "device %d out of range (%d devices)", cfg->device, count);
    NBH_HIP(hipSetDevice(cfg->device));
    hipDeviceProp_t p;
    NBH_HIP(hipGetDeviceProperties(&p, cfg->device));
    auto* c = new (std::nothrow) nbh_ctx();
    if (!c) return fail(NBH_ERR_ALLOC, "context allocation failed");
    c->cfg = *cfg;
    c->cus = p.multiProcessorCount;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(NBH_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    e = c->scal.reserve(sizeof(double) * 8);
    if (e == hipSuccess) e = c->flag.reserve(sizeof(int) * 2);
    if (e == hipSuccess) e = c->vwork.reserve(sizeof(double) * nbh::vref_work_doubles());
    if (e != hipSuccess) {
        nbh_destroy(c);
        return fail(NBH_ERR_ALLOC, "device allocation failed: %s", hipGetErrorString(e));
    }
    rc = nbh::scratch_setup(cfg->device, &c->scratch);
    if (rc) {
        nbh_destroy(c);
        return rc;
    }
    *out = c;
    return NBH_OK;
}

int nbh_destroy(nbh_ctx* c)
{
    if (!c) return NBH_OK;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (auto& ep : c->ev_pending) c->ev_free.push_back(ep);
    for (auto& ep : c->ev_free) {
        (void)hipEventDestroy(ep.a);
        (void)hipEventDestroy(ep.b);
    }
    for (DevBuf* b : {&c->in, &c->out, &c->partial, &c->rec, &c->flag, &c->scal, &c->state, &c->f0,
                      &c->f1, &c->ework, &c->vwork})
        b->release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return NBH_OK;
}

int nbh_configure(nbh_ctx* c, const nbh_config* cfg)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    int rc = check_cfg(cfg);
    if (rc) return rc;
    if (cfg->device != c->cfg.device)
        return fail(NBH_ERR_ARG, "cannot move a context to another device");
    c->cfg = *cfg;
    drop_graph(c);
    return NBH_OK;
}

int nbh_launch_shape(nbh_ctx* c, int n, int* block, int* jsplit, int* workgroups)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    const nbh::Shape s = pick_shape(c, n);
    if (block) *block = s.block;
    if (jsplit) *jsplit = s.jsplit;
    if (workgroups) {
        if (s.sym_chunk > 0) {
            // the symmetric grid, then the block-diagonal pass (256 targets per workgroup)
            *workgroups = (int)(nbh::sym_workgroups(c->cfg.precision, n, s.sym_chunk) + (n + 255) / 256);
        } else {
            *workgroups = ((n + s.block * s.ipl - 1) / (s.block * s.ipl)) * s.jsplit;
        }
    }
    return NBH_OK;
}

int nbh_launch_sources(nbh_ctx* c, int n, int* sources)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    const nbh::Shape s = pick_shape(c, n);
    if (sources) *sources = s.sym_chunk > 0 ? 0 : (s.scalar ? NBH_SRC_SCALAR : NBH_SRC_LDS);
    return NBH_OK;
}

int nbh_launch_pairing(nbh_ctx* c, int n, int* pairing)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    if (pairing) *pairing = pick_shape(c, n).sym_chunk > 0 ? NBH_PAIR_SYMMETRIC : NBH_PAIR_DIRECTED;
    return NBH_OK;
}

int nbh_launch_ipl(nbh_ctx* c, int n, int* ipl)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    if (ipl) *ipl = pick_shape(c, n).ipl;
    return NBH_OK;
}

int nbh_autotune(nbh_ctx* c, int n, int* block_out, int* ipl_out, int* jsplit_out, double* ms_out)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    NBH_HIP(hipSetDevice(c->cfg.device));
    const size_t nb = sizeof(double) * (size_t)n;
    std::vector<double> h((size_t)7 * n);
    double* p = h.data();
    if (nbh_plummer(n, 1, p, p + n, p + 2 * (size_t)n, p + 3 * (size_t)n, p + 4 * (size_t)n, p + 5 * (size_t)n,
                    p + 6 * (size_t)n) != NBH_OK)
        return fail(NBH_ERR_ARG, "autotune: cannot build a test system of %d bodies", n);
    if (c->in.reserve(7 * nb) != hipSuccess || c->out.reserve(6 * nb) != hipSuccess)
        return fail(NBH_ERR_ALLOC, "autotune: device allocation failed");
    NBH_HIP(hipMemcpyAsync(c->in.p, p, 7 * nb, hipMemcpyHostToDevice, c->stream));
    const nbh::State st = state_of(c->in.d(), n);
    const nbh::Forces fo = forces_of(c->out.d(), n);
    const int nj = n <= 65536 ? n : std::max(65536, n / 16);
    hipEvent_t e0, e1;
    NBH_HIP(hipEventCreate(&e0));
    NBH_HIP(hipEventCreate(&e1));
    struct Trial {
        double ms;
        int block, ipl, split, scalar;
    };
    std::vector<Trial> trials;
    int rc = NBH_OK;
    // one timed launch (after `reps - 1` untimed ones) of shape sh
    auto time_shape = [&](const nbh::Shape& sh, int reps, int trial_nj, double* out_ms) -> int {
        if (sh.jsplit > 1 && c->partial.reserve(sizeof(double) * 6 * (size_t)sh.jsplit * n) != hipSuccess)
            return fail(NBH_ERR_ALLOC, "autotune: partial-sum allocation failed");
        if (sh.scalar && c->rec.reserve(nbh::source_record_bytes(c->cfg.precision, n)) != hipSuccess)
            return fail(NBH_ERR_ALLOC, "autotune: source-record allocation failed");
        for (int rep = 0; rep < reps; ++rep) {
            hipError_t e = hipEventRecord(e0, c->stream);
            if (e == hipSuccess)
                e = nbh::launch_force(c->cfg.precision, st, n, c->cfg.eps2, sh, fo, c->partial.d(), c->rec.p,
                                      c->vwork.d(), c->stream, nullptr, nullptr, trial_nj);
            if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) return fail(NBH_ERR_HIP, "autotune trial failed: %s", hipGetErrorString(e));
            *out_ms = ms;
        }
        return (int)NBH_OK;
    };
    // stage 1: every shape on the truncated source range, with either source
    // path unless the config sets one
    const int src_lo = c->cfg.sources == NBH_SRC_SCALAR ? 1 : 0;
    const int src_hi = c->cfg.sources == NBH_SRC_LDS ? 0 : 1;
    for (int scalar = src_lo; scalar <= src_hi && rc == NBH_OK; ++scalar) {
        for (int block : {64, 128, 256}) {
            if (n >= 65536 && block == 64) continue;
            for (int ipl : {1, 2}) {
                int last = -1;
                for (int split : {1, 2, 4, 8, 16, 32}) {
                    const nbh::Shape sh = make_shape(n, block, ipl, split, scalar != 0, nj);
                    if (sh.jsplit == last) continue;
                    last = sh.jsplit;
                    double ms = 0.0;
                    if ((rc = time_shape(sh, 2, nj, &ms)) != NBH_OK) break;  // the first launch warms up
                    trials.push_back({ms, block, ipl, split, scalar});
                }
                if (rc) break;
            }
            if (rc) break;
        }
    }
    double best = 1e300;
    int bb = 0, bi = 0, bs = 0, bsc = 0;
    if (rc == NBH_OK) {
        std::sort(trials.begin(), trials.end(), [](const Trial& a, const Trial& b) { return a.ms < b.ms; });
        // stage 2 (truncated trials only): the three fastest again on all n
        // sources, since the truncation changes how the splits balance, and
        // the fastest one's shape with half and twice its split
        const bool full = nj < n;
        size_t top = full ? std::min<size_t>(3, trials.size()) : std::min<size_t>(1, trials.size());
        if (full && !trials.empty()) {
            const Trial b0 = trials[0];
            for (int sp : {b0.split / 2, b0.split * 2}) {
                if (sp < 1 || sp > 32) continue;
                bool have = false;
                for (size_t k = 0; k < top; ++k)
                    have |= trials[k].block == b0.block && trials[k].ipl == b0.ipl && trials[k].split == sp &&
                            trials[k].scalar == b0.scalar;
                if (!have) trials.insert(trials.begin() + top++, Trial{0.0, b0.block, b0.ipl, sp, b0.scalar});
            }
        }
        for (size_t k = 0; k < top && rc == NBH_OK; ++k) {
            double ms = trials[k].ms;
            if (full)
                rc = time_shape(make_shape(n, trials[k].block, trials[k].ipl, trials[k].split, trials[k].scalar != 0),
                                1, 0, &ms);
            if (rc == NBH_OK && ms < best) {
                best = ms;
                bb = trials[k].block;
                bi = trials[k].ipl;
                bs = trials[k].split;
                bsc = trials[k].scalar;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    auto& t = c->tuned[c->cfg.precision][size_bucket(n)];
    t.block = bb;
    t.ipl = bi;
    t.jsplit = bs;
    t.scalar = bsc;
    drop_graph(c);
    if (block_out) *block_out = bb;
    if (ipl_out) *ipl_out = bi;
    if (jsplit_out) *jsplit_out = pick_shape(c, n).jsplit;
    if (ms_out) *ms_out = best;
    return NBH_OK;
}

int nbh_force(nbh_ctx* c, int n, const double* x, const double* y, const double* z,
              const double* m, const double* vx, const double* vy, const double* vz,
              double* ax, double* ay, double* az, double* jx, double* jy, double* jz,
              double* kernel_ms)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    const double* ins[7] = {x, y, z, m, vx, vy, vz};
    double* outs[6] = {ax, ay, az, jx, jy, jz};
    for (auto p : ins)
        if (!p) return fail(NBH_ERR_ARG, "input array is NULL");
    for (auto p : outs)
        if (!p) return fail(NBH_ERR_ARG, "output array is NULL");
    NBH_HIP(hipSetDevice(c->cfg.device));
    const size_t nb = sizeof(double) * (size_t)n;
    if (c->in.reserve(7 * nb) != hipSuccess || c->out.reserve(6 * nb) != hipSuccess)
        return fail(NBH_ERR_ALLOC, "device allocation of %zu bytes failed", 13 * nb);
    for (int k = 0; k < 7; ++k)
        NBH_HIP(hipMemcpyAsync(c->in.d() + (size_t)k * n, ins[k], nb, hipMemcpyHostToDevice, c->stream));
    const nbh::State s = state_of(c->in.d(), n);
    const nbh::Forces f = forces_of(c->out.d(), n);
    const nbh::Shape sh = pick_shape(c, n);
    if (sh.jsplit > 1 && c->partial.reserve(sizeof(double) * 6 * (size_t)sh.jsplit * n) != hipSuccess)
        return fail(NBH_ERR_ALLOC, "device allocation of split partials failed");
    if (sh.scalar && c->rec.reserve(nbh::source_record_bytes(c->cfg.precision, n)) != hipSuccess)
        return fail(NBH_ERR_ALLOC, "device allocation of source records failed");
    double* part = sh.jsplit > 1 ? c->partial.d() : nullptr;
    NBH_HIP(hipMemsetAsync(c->flag.p, 0, sizeof(int), c->stream));
    // the host call times its own kernels regardless of the profiling switch;
    // the event pair goes back to the free list on every path below
    EventPair ep;
    int rc = get_events(c, &ep);
    if (rc) return rc;
    int flag = 0;
    float ms = 0.f;
    hipError_t e = nbh::launch_force(c->cfg.precision, s, n, c->cfg.eps2, sh, f, part, c->rec.p, c->vwork.d(),
                                     c->stream, ep.a, ep.b);
    const char* what = "force launch";
    if (e == hipSuccess) {
        what = "finite check";
        e = nbh::launch_check_finite(n, f, static_cast<int*>(c->flag.p), c->stream);
    }
    for (int k = 0; k < 6 && e == hipSuccess; ++k) {
        what = "result copy";
        e = hipMemcpyAsync(outs[k], c->out.d() + (size_t)k * n, nb, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, c->flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) {
        what = "stream synchronize";
        e = hipStreamSynchronize(c->stream);
    }
    if (e == hipSuccess) {
        what = "kernel timing";
        e = hipEventElapsedTime(&ms, ep.a, ep.b);
    }
    c->ev_free.push_back(ep);
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
    if (c->profiling) {
        c->force_ms += ms;
        c->force_launches += 1;
    }
    if (kernel_ms) *kernel_ms = ms;
    if (flag) return fail(NBH_ERR_NONFINITE, "force result holds inf/NaN (n=%d, eps2=%g)", n, c->cfg.eps2);
    return NBH_OK;
}

int nbh_force_device(nbh_ctx* c, int n, const double* x, const double* y, const double* z,
                     const double* m, const double* vx, const double* vy, const double* vz,
                     double* ax, double* ay, double* az, double* jx, double* jy, double* jz,
                     void* stream)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    const double* ins[7] = {x, y, z, m, vx, vy, vz};
    double* outs[6] = {ax, ay, az, jx, jy, jz};
    for (auto p : ins)
        if (!p) return fail(NBH_ERR_ARG, "input array is NULL");
    for (auto p : outs)
        if (!p) return fail(NBH_ERR_ARG, "output array is NULL");
    NBH_HIP(hipSetDevice(c->cfg.device));
    nbh::State s{x, y, z, m, vx, vy, vz};
    nbh::Forces f;
    for (int k = 0; k < 6; ++k) f.c[k] = outs[k];
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    // events are only timed on the context's own stream, which nbh_timing drains
    return enqueue_force(c, s, n, f, st, st == c->stream);
}

int nbh_force_device_noalloc(nbh_ctx* c, int n, const double* x, const double* y, const double* z,
                             const double* m, const double* vx, const double* vy, const double* vz,
                             double* ax, double* ay, double* az, double* jx, double* jy, double* jz,
                             void* stream)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    const double* ins[7] = {x, y, z, m, vx, vy, vz};
    double* outs[6] = {ax, ay, az, jx, jy, jz};
    for (auto p : ins)
        if (!p) return fail(NBH_ERR_ARG, "input array is NULL");
    for (auto p : outs)
        if (!p) return fail(NBH_ERR_ARG, "output array is NULL");
    if (!c->scratch) return fail(NBH_ERR_STATE, "context has no scratch address");
    NBH_HIP(hipSetDevice(c->cfg.device));
    nbh::Shape sh = pick_shape(c, n);
    if (int rc = fit_scratch(c, n, sh)) return rc;
    // layout: reduction | records | partials (every piece is fully written by
    // the evaluation before it is read, so nothing needs clearing)
    char* base = static_cast<char*>(c->scratch);
    size_t off = align256(sizeof(double) * nbh::vref_work_doubles());
    double* vwork = reinterpret_cast<double*>(base);
    void* rec = nullptr;
    if (sh.sym_chunk == 0 && sh.scalar) {
        rec = base + off;
        off += align256(nbh::source_record_bytes(c->cfg.precision, n));
    }
    double* part = (sh.sym_chunk == 0 && sh.jsplit > 1) ? reinterpret_cast<double*>(base + off) : nullptr;
    nbh::State s{x, y, z, m, vx, vy, vz};
    nbh::Forces f;
    for (int k = 0; k < 6; ++k) f.c[k] = outs[k];
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const hipError_t e = nbh::launch_force(c->cfg.precision, s, n, c->cfg.eps2, sh, f, part, rec, vwork, st,
                                           nullptr, nullptr);
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "force launch failed: %s", hipGetErrorString(e));
    return NBH_OK;
}

int nbh_upload(nbh_ctx* c, int n, const double* x, const double* y, const double* z,
               const double* m, const double* vx, const double* vy, const double* vz)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    const double* ins[7] = {x, y, z, m, vx, vy, vz};
    for (auto p : ins)
        if (!p) return fail(NBH_ERR_ARG, "input array is NULL");
    NBH_HIP(hipSetDevice(c->cfg.device));
    const size_t nb = sizeof(double) * (size_t)n;
    drop_graph(c);
    if (c->state.reserve(7 * nb) != hipSuccess || c->f0.reserve(6 * nb) != hipSuccess ||
        c->f1.reserve(6 * nb) != hipSuccess)
        return fail(NBH_ERR_ALLOC, "device allocation for n=%d failed", n);
    c->n = n;
    for (int k = 0; k < 7; ++k)
        NBH_HIP(hipMemcpyAsync(c->state.d() + (size_t)k * n, ins[k], nb, hipMemcpyHostToDevice, c->stream));
    int rc = enqueue_force(c, state_of(c->state.d(), n), n, forces_of(c->f0.d(), n), c->stream, false);
    if (rc) return rc;
    NBH_HIP(hipStreamSynchronize(c->stream));
    c->dt_set = false;
    return NBH_OK;
}

int nbh_initial_dt(nbh_ctx* c, double eta0, double* dt)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (!dt) return fail(NBH_ERR_ARG, "dt is NULL");
    if (c->n == 0) return fail(NBH_ERR_STATE, "no resident state: call nbh_upload first");
    NBH_HIP(hipSetDevice(c->cfg.device));
    auto* slot = reinterpret_cast<unsigned long long*>(c->scal.d() + 4);
    hipError_t e = nbh::launch_fill_inf(slot, c->stream);
    if (e == hipSuccess) e = nbh::launch_initial_dt(c->n, forces_of(c->f0.d(), c->n), eta0, slot, c->stream);
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "initial dt failed: %s", hipGetErrorString(e));
    NBH_HIP(hipMemcpyAsync(dt, slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    return NBH_OK;
}

int nbh_set_dt(nbh_ctx* c, double dt, int adaptive, double eta)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(NBH_ERR_ARG, "dt must be finite and > 0, got %g", dt);
    if (adaptive && !(eta > 0.0)) return fail(NBH_ERR_ARG, "eta must be > 0, got %g", eta);
    NBH_HIP(hipSetDevice(c->cfg.device));
    double v[2];
    v[0] = dt;
    const double inf = INFINITY;
    std::memcpy(&v[1], &inf, sizeof inf);
    NBH_HIP(hipMemcpyAsync(c->scal.d(), v, sizeof v, hipMemcpyHostToDevice, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    if ((adaptive != 0) != c->adaptive || eta != c->eta) drop_graph(c);
    c->adaptive = adaptive != 0;
    c->eta = eta;
    c->dt_set = true;
    return NBH_OK;
}

int nbh_get_dt(nbh_ctx* c, double* dt)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (!dt) return fail(NBH_ERR_ARG, "dt is NULL");
    NBH_HIP(hipSetDevice(c->cfg.device));
    NBH_HIP(hipMemcpyAsync(dt, c->scal.d(), sizeof(double), hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    return NBH_OK;
}

namespace {

// Enqueues one Hermite step of the resident state on c->stream.
int enqueue_step(nbh_ctx* c, bool allow_events)
{
    const int n = c->n;
    double* b = c->state.d();
    const nbh::Forces f0 = forces_of(c->f0.d(), n);
    const nbh::Forces f1 = forces_of(c->f1.d(), n);
    double* dt = c->scal.d();
    auto* dt_next = reinterpret_cast<unsigned long long*>(c->scal.d() + 1);
    double *x = b, *y = b + n, *z = b + 2 * (size_t)n;
    double *vx = b + 4 * (size_t)n, *vy = b + 5 * (size_t)n, *vz = b + 6 * (size_t)n;
    hipError_t e = nbh::launch_predict(n, dt, x, y, z, vx, vy, vz, f0, c->stream);
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "predict launch failed: %s", hipGetErrorString(e));
    int rc = enqueue_force(c, state_of(b, n), n, f1, c->stream, allow_events);
    if (rc) return rc;
    e = nbh::launch_correct(n, dt, x, y, z, vx, vy, vz, f0, f1, c->adaptive ? dt_next : nullptr,
                            c->eta, c->stream);
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "correct launch failed: %s", hipGetErrorString(e));
    if (c->adaptive) {
        e = nbh::launch_dt_advance(dt, dt_next, c->stream);
        if (e != hipSuccess) return fail(NBH_ERR_HIP, "dt advance failed: %s", hipGetErrorString(e));
    }
    return NBH_OK;
}

int build_graph(nbh_ctx* c, int steps)
{
    drop_graph(c);
    if (int rc = reserve_force_buffers(c, pick_shape(c, c->n), c->n)) return rc;
    NBH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = NBH_OK;
    for (int k = 0; k < steps && rc == NBH_OK; ++k) rc = enqueue_step(c, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        c->gexec = nullptr;
        return fail(NBH_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    }
    c->g_steps = steps;
    c->g_n = c->n;
    c->g_adaptive = c->adaptive;
    return NBH_OK;
}

}  // namespace

int nbh_step(nbh_ctx* c, int nsteps, int use_graph)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (nsteps < 0) return fail(NBH_ERR_ARG, "nsteps must be >= 0, got %d", nsteps);
    if (c->n == 0) return fail(NBH_ERR_STATE, "no resident state: call nbh_upload first");
    if (!c->dt_set) return fail(NBH_ERR_STATE, "no time step: call nbh_set_dt first");
    NBH_HIP(hipSetDevice(c->cfg.device));
    const auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    if (use_graph && nsteps > 0) {
        const int chunk = nsteps < 32 ? nsteps : 32;
        if (!c->gexec || c->g_steps != chunk || c->g_n != c->n || c->g_adaptive != c->adaptive) {
            int rc = build_graph(c, chunk);
            if (rc) return rc;
        }
        while (nsteps - done >= c->g_steps) {
            NBH_HIP(hipGraphLaunch(c->gexec, c->stream));
            done += c->g_steps;
        }
    }
    for (; done < nsteps; ++done) {
        int rc = enqueue_step(c, true);
        if (rc) return rc;
    }
    NBH_HIP(hipStreamSynchronize(c->stream));
    const auto t1 = std::chrono::steady_clock::now();
    c->step_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    c->steps += nsteps;
    return settle_events(c);
}

int nbh_download(nbh_ctx* c, double* x, double* y, double* z, double* m, double* vx, double* vy,
                 double* vz, double* ax, double* ay, double* az, double* jx, double* jy, double* jz)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (c->n == 0) return fail(NBH_ERR_STATE, "no resident state: call nbh_upload first");
    NBH_HIP(hipSetDevice(c->cfg.device));
    const int n = c->n;
    const size_t nb = sizeof(double) * (size_t)n;
    double* so[7] = {x, y, z, m, vx, vy, vz};
    double* fo[6] = {ax, ay, az, jx, jy, jz};
    for (int k = 0; k < 7; ++k)
        if (so[k])
            NBH_HIP(hipMemcpyAsync(so[k], c->state.d() + (size_t)k * n, nb, hipMemcpyDeviceToHost, c->stream));
    for (int k = 0; k < 6; ++k)
        if (fo[k])
            NBH_HIP(hipMemcpyAsync(fo[k], c->f0.d() + (size_t)k * n, nb, hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    return NBH_OK;
}

namespace {

int energy_of(nbh_ctx* c, const nbh::State& s, int n, int compat, double* kin, double* pot)
{
    const int nb = nbh::energy_blocks(n);
    if (c->ework.reserve(sizeof(double) * 2 * (size_t)nb) != hipSuccess)
        return fail(NBH_ERR_ALLOC, "device allocation for energy failed");
    double* out = c->scal.d() + 2;
    const double eps2 = compat ? 1.0e-7 : c->cfg.eps2;
    hipError_t e = nbh::launch_energy(s, n, eps2, compat, c->ework.d(), out, c->stream);
    if (e != hipSuccess) return fail(NBH_ERR_HIP, "energy launch failed: %s", hipGetErrorString(e));
    double v[2];
    NBH_HIP(hipMemcpyAsync(v, out, sizeof v, hipMemcpyDeviceToHost, c->stream));
    NBH_HIP(hipStreamSynchronize(c->stream));
    if (kin) *kin = v[0];
    if (pot) *pot = v[1];
    return NBH_OK;
}

}  // namespace

int nbh_energy(nbh_ctx* c, int compat, double* kin, double* pot)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (c->n == 0) return fail(NBH_ERR_STATE, "no resident state: call nbh_upload first");
    NBH_HIP(hipSetDevice(c->cfg.device));
    return energy_of(c, state_of(c->state.d(), c->n), c->n, compat, kin, pot);
}

int nbh_energy_host(nbh_ctx* c, int n, const double* x, const double* y, const double* z,
                    const double* m, const double* vx, const double* vy, const double* vz,
                    int compat, double* kin, double* pot)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    if (bad_n(n)) return fail(NBH_ERR_ARG, "n must be >= 1, got %d", n);
    const double* ins[7] = {x, y, z, m, vx, vy, vz};
    for (auto p : ins)
        if (!p) return fail(NBH_ERR_ARG, "input array is NULL");
    NBH_HIP(hipSetDevice(c->cfg.device));
    const size_t nb = sizeof(double) * (size_t)n;
    if (c->in.reserve(7 * nb) != hipSuccess) return fail(NBH_ERR_ALLOC, "device allocation failed");
    for (int k = 0; k < 7; ++k)
        NBH_HIP(hipMemcpyAsync(c->in.d() + (size_t)k * n, ins[k], nb, hipMemcpyHostToDevice, c->stream));
    return energy_of(c, state_of(c->in.d(), n), n, compat, kin, pot);
}

int nbh_set_profiling(nbh_ctx* c, int on)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    c->profiling = on != 0;
    return NBH_OK;
}

int nbh_timing(nbh_ctx* c, double* force_ms, int64_t* force_launches, double* step_ms, int64_t* steps)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    NBH_HIP(hipSetDevice(c->cfg.device));
    NBH_HIP(hipStreamSynchronize(c->stream));
    int rc = settle_events(c);
    if (rc) return rc;
    if (force_ms) *force_ms = c->force_ms;
    if (force_launches) *force_launches = c->force_launches;
    if (step_ms) *step_ms = c->step_ms;
    if (steps) *steps = c->steps;
    return NBH_OK;
}

int nbh_reset_timing(nbh_ctx* c)
{
    if (!c) return fail(NBH_ERR_ARG, "context is NULL");
    int rc = nbh_timing(c, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    c->force_ms = 0.0;
    c->force_launches = 0;
    c->step_ms = 0.0;
    c->steps = 0;
    return NBH_OK;
}

}  // extern "C"
