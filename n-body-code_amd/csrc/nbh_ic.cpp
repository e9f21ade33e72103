// nbh_ic.cpp -- initial conditions (host code; no device needed).
//
//   nbh_plummer       equal-mass Plummer sphere in Henon units, for the
//                     benchmark configurations (no reference counterpart).
//   nbh_reference_ic  NBody::initialize (src/nbody.cpp:69-114), the
//                     reference CLI's initial conditions, bit-identical:
//                     the same libstdc++ engine and distribution, the same
//                     arithmetic order, the same float rescale factor.
#include "../../include/nbody_hip.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

namespace nbh {
int fail(int code, const char* fmt, ...);  // nbh_api.hip: records the last error
}

namespace {

constexpr double kPi = 3.14159265358979323846;

// u in [0, 1) with 53 random bits.
inline double unit(std::mt19937_64& g) { return (double)(g() >> 11) * 0x1.0p-53; }

// Isotropic direction times length r.
inline void iso(std::mt19937_64& g, double r, double& x, double& y, double& z)
{
    const double ct = 1.0 - 2.0 * unit(g);
    const double st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
    const double ph = 2.0 * kPi * unit(g);
    x = r * st * std::cos(ph);
    y = r * st * std::sin(ph);
    z = r * ct;
}

// Kinetic energy in the reference's order (src/utils.cpp:125-135).
double kinetic_ref(int n, const double* w, const double* vx, const double* vy, const double* vz)
{
    double t = 0.0;
    for (int i = 0; i < n; i++) {
        const double a = vx[i] * vx[i];
        const double b = vy[i] * vy[i];
        const double c = vz[i] * vz[i];
        t += 0.5 * w[i] * (a + b + c);
    }
    return t;
}

// Potential as the reference defines it (src/utils.cpp:137-153): weight
// w_j*w_j, softening 1e-7, serial order.
double potential_ref(int n, const double* x, const double* y, const double* z, const double* w)
{
    double u = 0.0;
    for (int i = 0; i < n; i++) {
        double t = 0.0;
        for (int j = i + 1; j < n; j++) {
            const double dx = x[i] - x[j];
            const double dy = y[i] - y[j];
            const double dz = z[i] - z[j];
            const double r2 = dx * dx + dy * dy + dz * dz + 1.0e-7;
            t -= w[j] * w[j] * (1. / std::sqrt(r2));
        }
        u += t;
    }
    return u;
}

}  // namespace

extern "C" {

int nbh_plummer(int n, uint64_t seed, double* x, double* y, double* z, double* m, double* vx,
                double* vy, double* vz)
{
    if (n < 1 || !x || !y || !z || !m || !vx || !vy || !vz)
        return nbh::fail(NBH_ERR_ARG, "nbh_plummer: need n >= 1 and seven arrays (n=%d)", n);
    std::mt19937_64 g(seed);
    // Aarseth, Henon & Wielen (1974) with scale length 1, then the
    // Henon-unit scaling a = 3 pi / 16 (E = -1/4 for the untruncated model).
    const double a = 3.0 * kPi / 16.0;
    const double vs = 1.0 / std::sqrt(a);
    for (int i = 0; i < n; i++) {
        const double X = 0.999 * (1.0 - unit(g));  // mass fraction in (0, 0.999]
        const double r = 1.0 / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
        iso(g, r * a, x[i], y[i], z[i]);
        double q, f;
        do {  // von Neumann rejection on q^2 (1 - q^2)^(7/2), max < 0.1
            q = unit(g);
            f = 0.1 * unit(g);
        } while (f > q * q * std::pow(1.0 - q * q, 3.5));
        const double v = q * std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
        iso(g, v * vs, vx[i], vy[i], vz[i]);
        m[i] = 1.0 / n;
    }
    double c[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        c[0] += x[i];
        c[1] += y[i];
        c[2] += z[i];
        c[3] += vx[i];
        c[4] += vy[i];
        c[5] += vz[i];
    }
    for (double& v : c) v /= n;
    for (int i = 0; i < n; i++) {
        x[i] -= c[0];
        y[i] -= c[1];
        z[i] -= c[2];
        vx[i] -= c[3];
        vy[i] -= c[4];
        vz[i] -= c[5];
    }
    return NBH_OK;
}

int nbh_reference_ic(int n, double vmax, double rmax, double mmin, double mmax, double* x,
                     double* y, double* z, double* m, double* vx, double* vy, double* vz)
{
    if (n < 1 || !x || !y || !z || !m || !vx || !vy || !vz)
        return nbh::fail(NBH_ERR_ARG, "nbh_reference_ic: need n >= 1 and seven arrays (n=%d)", n);
    std::mt19937 engine(42);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    const double alpha = -2.35 + 1.0;  // Salpeter index + 1 (include/nbody.h:28)
    int i = 0;
    while (i < n) {
        const double xt = dist(engine) * rmax;
        const double yt = dist(engine) * rmax;
        const double zt = dist(engine) * rmax;
        const double r2 = xt * xt + yt * yt + zt * zt;
        if (r2 < rmax * rmax) {
            x[i] = xt;
            y[i] = yt;
            z[i] = zt;
            i++;
        }
    }
    const double lo = std::pow(mmin, alpha);
    const double hi = std::pow(mmax, alpha);
    const double ia = 1.0 / alpha;
    for (i = 0; i < n; i++) {
        vx[i] = dist(engine) * vmax;
        vy[i] = dist(engine) * vmax;
        vz[i] = dist(engine) * vmax;
        const double t = (hi - lo) * dist(engine) + lo;
        m[i] = std::pow(t, ia);
    }
    const double T = kinetic_ref(n, m, vx, vy, vz);
    const double U = potential_ref(n, x, y, z, m);
    const double Q = -2.0 * T / U;
    const float f = 1.0 / std::sqrt(Q);  // float, as src/nbody.cpp:104
    for (i = 0; i < n; i++) {
        vx[i] = vx[i] * f;
        vy[i] = vy[i] * f;
        vz[i] = vz[i] * f;
    }
    return NBH_OK;
}

namespace {

struct StateHeader {
    char magic[8];
    uint32_t version;
    uint32_t reserved;
    int64_t n;
    double time;
    int64_t step;
    unsigned char pad[24];
};
static_assert(sizeof(StateHeader) == 64, "state header is 64 bytes");

const char kMagic[8] = {'N', 'B', 'H', 'S', 'T', 'A', 'T', 'E'};

int read_header(FILE* f, const char* path, StateHeader* h)
{
    if (std::fread(h, sizeof *h, 1, f) != 1)
        return nbh::fail(NBH_ERR_ARG, "%s: too short for a state header", path);
    if (std::memcmp(h->magic, kMagic, 8) != 0) return nbh::fail(NBH_ERR_ARG, "%s: not a state file", path);
    if (h->version != 1) return nbh::fail(NBH_ERR_ARG, "%s: unsupported version %u", path, h->version);
    if (h->n < 1 || h->n > (1LL << 31) - 1) return nbh::fail(NBH_ERR_ARG, "%s: bad body count %lld", path, (long long)h->n);
    return NBH_OK;
}

}  // namespace

int nbh_state_save(const char* path, int n, const double* x, const double* y, const double* z, const double* m,
                   const double* vx, const double* vy, const double* vz, double time, int64_t step)
{
    const double* a[7] = {x, y, z, m, vx, vy, vz};
    if (!path || n < 1) return nbh::fail(NBH_ERR_ARG, "nbh_state_save: need a path and n >= 1");
    for (auto p : a)
        if (!p) return nbh::fail(NBH_ERR_ARG, "nbh_state_save: array is NULL");
    FILE* f = std::fopen(path, "wb");
    if (!f) return nbh::fail(NBH_ERR_ARG, "%s: cannot open for writing", path);
    StateHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = 1;
    h.n = n;
    h.time = time;
    h.step = step;
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
    for (int k = 0; k < 7 && ok; ++k) ok = std::fwrite(a[k], sizeof(double), (size_t)n, f) == (size_t)n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? NBH_OK : nbh::fail(NBH_ERR_ARG, "%s: write failed", path);
}

int nbh_state_info(const char* path, int* n, double* time, int64_t* step)
{
    if (!path) return nbh::fail(NBH_ERR_ARG, "nbh_state_info: path is NULL");
    FILE* f = std::fopen(path, "rb");
    if (!f) return nbh::fail(NBH_ERR_ARG, "%s: cannot open", path);
    StateHeader h;
    const int rc = read_header(f, path, &h);
    std::fclose(f);
    if (rc) return rc;
    if (n) *n = (int)h.n;
    if (time) *time = h.time;
    if (step) *step = h.step;
    return NBH_OK;
}

int nbh_state_load(const char* path, int n, double* x, double* y, double* z, double* m, double* vx, double* vy,
                   double* vz)
{
    double* a[7] = {x, y, z, m, vx, vy, vz};
    if (!path) return nbh::fail(NBH_ERR_ARG, "nbh_state_load: path is NULL");
    for (auto p : a)
        if (!p) return nbh::fail(NBH_ERR_ARG, "nbh_state_load: array is NULL");
    FILE* f = std::fopen(path, "rb");
    if (!f) return nbh::fail(NBH_ERR_ARG, "%s: cannot open", path);
    StateHeader h;
    int rc = read_header(f, path, &h);
    if (rc == NBH_OK && h.n != n)
        rc = nbh::fail(NBH_ERR_ARG, "%s holds %lld bodies, caller expects %d", path, (long long)h.n, n);
    for (int k = 0; k < 7 && rc == NBH_OK; ++k)
        if (std::fread(a[k], sizeof(double), (size_t)n, f) != (size_t)n)
            rc = nbh::fail(NBH_ERR_ARG, "%s: truncated (array %d of 7)", path, k);
    std::fclose(f);
    return rc;
}

}  // extern "C"
