// nbody.cpp -- NBody with the reference's interface and step order
// (src/nbody.cpp:14-415) over the HIP force backend.
//
// Drop-in mode keeps the reference's structure: the predictor, snap/crackle,
// corrector and step control run on the host in the reference's arithmetic
// order, and force_calculation evaluates forces on the GPU.  Resident mode
// uploads the state once and runs whole steps on the GPU (nbh_step).
#include "nbody.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <utility>

namespace {
constexpr double kEta0 = 1.0e-4;  // include/nbody.h:15
constexpr double kEta = 1.0e-2;   // include/nbody.h:16

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double* alloc(int n)
{
    auto* p = static_cast<double*>(std::calloc((size_t)n, sizeof(double)));
    if (!p) {
        std::fprintf(stderr, "nbody: host allocation of %d doubles failed\n", n);
        std::exit(1);
    }
    return p;
}

[[noreturn]] void die(const char* what)
{
    std::fprintf(stderr, "nbody: %s: %s\n", what, nbh_last_error());
    std::exit(1);
}
}  // namespace

NBody::NBody(int Nparticles, double v_max, double r_max, double m_min, double m_max)
    : N(Nparticles), vmax(v_max), rmax(r_max), mmin(m_min), mmax(m_max)
{
    alpha = n + 1.0;
    sys.N = N;
    sys.rank = 0;
    sys.size = 1;
    sys.start = 0;
    sys.end = N;
    double** arrays[] = {&h_par_x, &h_par_y, &h_par_z, &h_par_w, &h_vel_x, &h_vel_y,
                         &h_vel_z, &h_acc_x, &h_acc_y, &h_acc_z, &h_adot_x, &h_adot_y,
                         &h_adot_z, &h_acc1_x, &h_acc1_y, &h_acc1_z, &h_a1dot_x, &h_a1dot_y,
                         &h_a1dot_z, &h_a2dot_x, &h_a2dot_y, &h_a2dot_z, &h_a3dot_x, &h_a3dot_y,
                         &h_a3dot_z};
    for (double** a : arrays) *a = alloc(N);
}

NBody::~NBody()
{
    double* arrays[] = {h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z,
                        h_acc_x, h_acc_y, h_acc_z, h_adot_x, h_adot_y, h_adot_z, h_acc1_x,
                        h_acc1_y, h_acc1_z, h_a1dot_x, h_a1dot_y, h_a1dot_z, h_a2dot_x,
                        h_a2dot_y, h_a2dot_z, h_a3dot_x, h_a3dot_y, h_a3dot_z};
    for (double* a : arrays) std::free(a);
}

// src/nbody.cpp:69-114 (ic == "ref"), or a Plummer sphere.
void NBody::initialize()
{
    int rc;
    if (!load_path.empty())
        rc = nbh_state_load(load_path.c_str(), N, h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z);
    else if (ic == "plummer")
        rc = nbh_plummer(N, seed, h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z);
    else
        rc = nbh_reference_ic(N, vmax, rmax, mmin, mmax, h_par_x, h_par_y, h_par_z, h_par_w,
                              h_vel_x, h_vel_y, h_vel_z);
    if (rc != NBH_OK) die("initial conditions");
}

void NBody::initializewithzeros()
{
    for (int i = 0; i < N; i++) {
        h_par_x[i] = h_par_y[i] = h_par_z[i] = h_par_w[i] = 0.0;
        h_vel_x[i] = h_vel_y[i] = h_vel_z[i] = 0.0;
    }
}

// src/nbody.cpp:131-150
double NBody::initialdt()
{
    double best = std::numeric_limits<double>::max();
    for (int i = 0; i < N; i++) {
        const double an = std::sqrt(h_acc_x[i] * h_acc_x[i] + h_acc_y[i] * h_acc_y[i] +
                                    h_acc_z[i] * h_acc_z[i]);
        const double jn = std::sqrt(h_adot_x[i] * h_adot_x[i] + h_adot_y[i] * h_adot_y[i] +
                                    h_adot_z[i] * h_adot_z[i]);
        const double cur = kEta0 * an / jn;
        if (cur < best) best = cur;
    }
    return best;
}

// src/nbody.cpp:152-163
void NBody::prediction()
{
    const double h = dt;
    for (int i = 0; i < N; i++) {
        h_par_x[i] += h_vel_x[i] * h + 0.5 * h_acc_x[i] * h * h + 1.0 / 6.0 * h_adot_x[i] * h * h * h;
        h_par_y[i] += h_vel_y[i] * h + 0.5 * h_acc_y[i] * h * h + 1.0 / 6.0 * h_adot_y[i] * h * h * h;
        h_par_z[i] += h_vel_z[i] * h + 0.5 * h_acc_z[i] * h * h + 1.0 / 6.0 * h_adot_z[i] * h * h * h;
        h_vel_x[i] += h_acc_x[i] * h + 0.5 * h_adot_x[i] * h * h;
        h_vel_y[i] += h_acc_y[i] * h + 0.5 * h_adot_y[i] * h * h;
        h_vel_z[i] += h_acc_z[i] * h + 0.5 * h_adot_z[i] * h * h;
    }
}

// src/nbody.cpp:165-179
void NBody::snapandcrackle()
{
    const double* a0[3] = {h_acc_x, h_acc_y, h_acc_z};
    const double* j0[3] = {h_adot_x, h_adot_y, h_adot_z};
    const double* a1[3] = {h_acc1_x, h_acc1_y, h_acc1_z};
    const double* j1[3] = {h_a1dot_x, h_a1dot_y, h_a1dot_z};
    double* s[3] = {h_a2dot_x, h_a2dot_y, h_a2dot_z};
    double* c[3] = {h_a3dot_x, h_a3dot_y, h_a3dot_z};
    for (int i = 0; i < N; i++) {
        const double invdt2 = 1.0 / (dt * dt);
        const double invdt3 = 1.0 / (dt * dt * dt);
        for (int k = 0; k < 3; k++) {
            s[k][i] = 2.0 * invdt2 * (-3.0 * (a0[k][i] - a1[k][i]) - (2.0 * j0[k][i] + j1[k][i]) * dt);
            c[k][i] = 6.0 * invdt3 * (2.0 * (a0[k][i] - a1[k][i]) + (j0[k][i] + j1[k][i]) * dt);
        }
    }
}

// src/nbody.cpp:181-196
void NBody::correction()
{
    double* p[3] = {h_par_x, h_par_y, h_par_z};
    double* v[3] = {h_vel_x, h_vel_y, h_vel_z};
    const double* s[3] = {h_a2dot_x, h_a2dot_y, h_a2dot_z};
    const double* c[3] = {h_a3dot_x, h_a3dot_y, h_a3dot_z};
    for (int i = 0; i < N; i++) {
        const double dt3 = dt * dt * dt;
        const double dt4 = dt3 * dt;
        const double dt5 = dt4 * dt;
        for (int k = 0; k < 3; k++) p[k][i] += 1.0 / 24.0 * s[k][i] * dt4 + 1.0 / 120.0 * c[k][i] * dt5;
        for (int k = 0; k < 3; k++) v[k][i] += 1.0 / 6.0 * s[k][i] * dt3 + 1.0 / 24.0 * c[k][i] * dt4;
    }
}

// src/nbody.cpp:198-222 (uses the start-of-step a0, j0, before the swap)
double NBody::updatedt()
{
    double best = std::numeric_limits<double>::max();
    for (int i = 0; i < N; i++) {
        const double na2 = h_acc_x[i] * h_acc_x[i] + h_acc_y[i] * h_acc_y[i] + h_acc_z[i] * h_acc_z[i];
        const double na = std::sqrt(na2);
        const double nj2 = h_adot_x[i] * h_adot_x[i] + h_adot_y[i] * h_adot_y[i] + h_adot_z[i] * h_adot_z[i];
        const double nj = std::sqrt(nj2);
        const double ns2 = h_a2dot_x[i] * h_a2dot_x[i] + h_a2dot_y[i] * h_a2dot_y[i] + h_a2dot_z[i] * h_a2dot_z[i];
        const double ns = std::sqrt(ns2);
        const double nc2 = h_a3dot_x[i] * h_a3dot_x[i] + h_a3dot_y[i] * h_a3dot_y[i] + h_a3dot_z[i] * h_a3dot_z[i];
        const double nc = std::sqrt(nc2);
        const double num = na * ns + nj2;
        const double den = nj * nc + ns2;
        const double cur = std::sqrt(kEta * num / den);
        if (cur < best) best = cur;
    }
    return best;
}

// src/nbody.cpp:224-289
void NBody::timestepping()
{
    if (fixed_dt > 0.0) dt = fixed_dt;
    sim_time += dt;
    double t0 = now();
    prediction();
    sys.t_pred = now() - t0;

    force_calculation(h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z, h_acc1_x,
                      h_acc1_y, h_acc1_z, h_a1dot_x, h_a1dot_y, h_a1dot_z, sys);
    sys.t_eval = sys.kernel_time;
#if defined(_DEBUG)
    // the reference's debug build checks every evaluation (src/nbody.cpp:248-252)
    test(h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z, h_acc1_x, h_acc1_y, h_acc1_z,
         h_a1dot_x, h_a1dot_y, h_a1dot_z, N);
#endif

    snapandcrackle();
    t0 = now();
    correction();
    sys.t_corr = now() - t0;
    dt = updatedt();

    std::swap(h_acc_x, h_acc1_x);
    std::swap(h_adot_x, h_a1dot_x);
    std::swap(h_acc_y, h_acc1_y);
    std::swap(h_adot_y, h_a1dot_y);
    std::swap(h_acc_z, h_acc1_z);
    std::swap(h_adot_z, h_a1dot_z);
}

// src/nbody.cpp:291-370
void NBody::cycle(int ncycle)
{
    sys.t_evalv.assign(ncycle, 0.0);
    sys.t_predv.assign(ncycle, 0.0);
    sys.t_corrv.assign(ncycle, 0.0);
    initialize();
    if (resident) {
        cycle_resident(ncycle);
        return;
    }
    force_calculation(h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z, h_acc_x,
                      h_acc_y, h_acc_z, h_adot_x, h_adot_y, h_adot_z, sys);
    sys.t_eval_init = sys.kernel_time;
#if defined(_DEBUG)
    test(h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z, h_acc_x, h_acc_y, h_acc_z, h_adot_x,
         h_adot_y, h_adot_z, N);  // src/nbody.cpp:326-330
#endif
    dt = fixed_dt > 0.0 ? fixed_dt : initialdt();
    T = compute_kinetic(h_par_w, h_vel_x, h_vel_y, h_vel_z, N);
    U = compute_potential(h_par_x, h_par_y, h_par_z, h_par_w, N, compat_energy);
    sys.initE = T + U;
    for (int i = 0; i < ncycle; i++) {
        timestepping();
        T = compute_kinetic(h_par_w, h_vel_x, h_vel_y, h_vel_z, N);
        U = compute_potential(h_par_x, h_par_y, h_par_z, h_par_w, N, compat_energy);
        sys.finE = T + U;
        sys.Q = -2.0 * T / U;
        sys.t_evalv[i] = sys.t_eval;
        sys.t_predv[i] = sys.t_pred;
        sys.t_corrv[i] = sys.t_corr;
    }
    sys.Err = estimate_err(sys.finE, sys.initE);
}

// Whole steps on the GPU: predictor, force, corrector and step control
// fused into device kernels (EVAL = pair-kernel time, PRED/CORR = the rest
// of the step's wall time, split evenly).
void NBody::cycle_resident(int ncycle)
{
    nbh_ctx* c = force_backend_context();
    const double t0 = now();
    if (nbh_upload(c, N, h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z) != NBH_OK)
        die("nbh_upload");
    sys.t_eval_init = now() - t0;  // copy-in plus the initial force
    if (fixed_dt > 0.0) {
        dt = fixed_dt;
    } else if (nbh_initial_dt(c, kEta0, &dt) != NBH_OK) {
        die("nbh_initial_dt");
    }
    if (nbh_set_dt(c, dt, fixed_dt > 0.0 ? 0 : 1, kEta) != NBH_OK) die("nbh_set_dt");
    const int compat = compat_energy ? 1 : 0;
    if (nbh_energy(c, compat, &T, &U) != NBH_OK) die("nbh_energy");
    sys.initE = T + U;
    nbh_set_profiling(c, 1);
    double fms0 = 0.0, sms0 = 0.0;
    int64_t dummy;
    nbh_timing(c, &fms0, &dummy, &sms0, &dummy);
    for (int i = 0; i < ncycle; i++) {
        double h = 0.0;
        if (nbh_get_dt(c, &h) != NBH_OK) die("nbh_get_dt");
        sim_time += h;
        if (nbh_step(c, 1, use_graph ? 1 : 0) != NBH_OK) die("nbh_step");
        double fms = 0.0, sms = 0.0;
        nbh_timing(c, &fms, &dummy, &sms, &dummy);
        sys.t_evalv[i] = (fms - fms0) * 1e-3;
        const double rest = std::fmax(0.0, (sms - sms0) - (fms - fms0)) * 1e-3;
        sys.t_predv[i] = sys.t_corrv[i] = 0.5 * rest;
        fms0 = fms;
        sms0 = sms;
        if (nbh_energy(c, compat, &T, &U) != NBH_OK) die("nbh_energy");
        sys.finE = T + U;
        sys.Q = -2.0 * T / U;
    }
    nbh_set_profiling(c, 0);
    if (nbh_get_dt(c, &dt) != NBH_OK) die("nbh_get_dt");
    if (nbh_download(c, h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z, h_acc_x,
                     h_acc_y, h_acc_z, h_adot_x, h_adot_y, h_adot_z) != NBH_OK)
        die("nbh_download");
    sys.Err = estimate_err(sys.finE, sys.initE);
}
