// ref_driver.cpp -- extern "C" handles onto the compiled reference.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
// with the reference's own src/kernel.cpp, src/utils.cpp and src/nbody.cpp,
// where they lie under the reference checkout, into oracle/_ref/*.so.  It is
// used to pin the C restatement in oracle/oracle.c (bitwise) and to write
// the golden fixtures under tests/golden/.  Nothing in the product loads it.
//
// Two builds: the default (FP64) force of src/kernel.cpp:1164-1224 and the
// _F32 force of src/kernel.cpp:1101-1163.  EPS2 is the reference's fixed
// 1e-7 (include/kernel.h:13); every handle here therefore runs at that
// softening.
#include <cstring>
#include <mpi.h>
#include <omp.h>
#include "nbody.h"
#include "kernel.h"

namespace {
// One OpenMP thread: the reference's array reduction merges thread copies
// in an unspecified order, so only the single-thread result is reproducible.
void ensure_mpi()
{
    omp_set_num_threads(1);
    int inited = 0;
    MPI_Initialized(&inited);
    if (!inited) {
        int argc = 0;
        char** argv = nullptr;
        MPI_Init(&argc, &argv);
    }
}
}  // namespace

extern "C" {

double ref_invsqrt2d(double x) { return InvSqrt2D(x); }

float ref_invsqrt2(float x) { return InvSqrt2(x); }

double ref_eps2() { return EPS2; }

// force_calculation (include/kernel.h:15-19) on one rank, with nthreads
// OpenMP threads (1: the reproducible order; more: the CPU timing baseline).
void ref_force_threads(int n, int nthreads, double* x, double* y, double* z, double* w, double* vx,
                       double* vy, double* vz, double* ax, double* ay, double* az,
                       double* jx, double* jy, double* jz)
{
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
    Sys sys;
    sys.N = n;
    sys.rank = 0;
    sys.size = 1;
    sys.start = 0;
    sys.end = n;
    force_calculation(x, y, z, w, vx, vy, vz, ax, ay, az, jx, jy, jz, sys);
    omp_set_num_threads(1);
}

// force_calculation (include/kernel.h:15-19) on one rank.
void ref_force(int n, double* x, double* y, double* z, double* w, double* vx,
               double* vy, double* vz, double* ax, double* ay, double* az,
               double* jx, double* jy, double* jz)
{
    omp_set_num_threads(1);
    Sys sys;
    sys.N = n;
    sys.rank = 0;
    sys.size = 1;
    sys.start = 0;
    sys.end = n;
    force_calculation(x, y, z, w, vx, vy, vz, ax, ay, az, jx, jy, jz, sys);
}

double ref_kinetic(int n, double* w, double* vx, double* vy, double* vz)
{
    omp_set_num_threads(1);
    return compute_kinetic(w, vx, vy, vz, n);
}

double ref_potential(int n, double* x, double* y, double* z, double* w)
{
    omp_set_num_threads(1);
    return compute_potential(x, y, z, w, n);
}

// NBody::initialize (src/nbody.cpp:69-114) with the CLI's parameters.
int ref_initialize(int n, double vmax, double rmax, double mmin, double mmax,
                   double* x, double* y, double* z, double* w, double* vx,
                   double* vy, double* vz)
{
    ensure_mpi();
    NBody nb(n, vmax, rmax, mmin, mmax);
    nb.initialize();
    std::memcpy(x, nb.h_par_x, sizeof(double) * n);
    std::memcpy(y, nb.h_par_y, sizeof(double) * n);
    std::memcpy(z, nb.h_par_z, sizeof(double) * n);
    std::memcpy(w, nb.h_par_w, sizeof(double) * n);
    std::memcpy(vx, nb.h_vel_x, sizeof(double) * n);
    std::memcpy(vy, nb.h_vel_y, sizeof(double) * n);
    std::memcpy(vz, nb.h_vel_z, sizeof(double) * n);
    return 0;
}

// Runs the reference integrator from a given state: the initial force and
// dt of NBody::cycle (src/nbody.cpp:310-336), then nsteps calls of
// NBody::timestepping (src/nbody.cpp:224-289).  dt_fixed > 0 overrides the
// adaptive step before every step (updatedt still runs inside
// timestepping, its value is discarded); dt_fixed <= 0 keeps the
// reference's adaptive dt.  state: 7*n in (x,y,z,w,vx,vy,vz), overwritten
// with the final state.  force: 6*n final (a,jerk).  dts: nsteps+1 entries,
// the dt used by each step and the one after the last.  kin/pot: nsteps+1
// entries of compute_kinetic / compute_potential (src/utils.cpp:125-153).
int ref_run(int n, double* state, int nsteps, double dt_fixed, double* force,
            double* dts, double* kin, double* pot)
{
    ensure_mpi();
    NBody nb(n, 1.0, 1.0, 0.01, 1.0);
    double* s[7] = {nb.h_par_x, nb.h_par_y, nb.h_par_z, nb.h_par_w,
                    nb.h_vel_x, nb.h_vel_y, nb.h_vel_z};
    for (int k = 0; k < 7; k++) std::memcpy(s[k], state + (size_t)k * n, sizeof(double) * n);
    force_calculation(nb.h_par_x, nb.h_par_y, nb.h_par_z, nb.h_par_w, nb.h_vel_x,
                      nb.h_vel_y, nb.h_vel_z, nb.h_acc_x, nb.h_acc_y, nb.h_acc_z,
                      nb.h_adot_x, nb.h_adot_y, nb.h_adot_z, nb.sys);
    nb.dt = nb.initialdt();
    kin[0] = compute_kinetic(nb.h_par_w, nb.h_vel_x, nb.h_vel_y, nb.h_vel_z, n);
    pot[0] = compute_potential(nb.h_par_x, nb.h_par_y, nb.h_par_z, nb.h_par_w, n);
    for (int k = 0; k < nsteps; k++) {
        if (dt_fixed > 0.0) nb.dt = dt_fixed;
        dts[k] = nb.dt;
        nb.timestepping();
        kin[k + 1] = compute_kinetic(nb.h_par_w, nb.h_vel_x, nb.h_vel_y, nb.h_vel_z, n);
        pot[k + 1] = compute_potential(nb.h_par_x, nb.h_par_y, nb.h_par_z, nb.h_par_w, n);
    }
    dts[nsteps] = nb.dt;
    double* e[7] = {nb.h_par_x, nb.h_par_y, nb.h_par_z, nb.h_par_w,
                    nb.h_vel_x, nb.h_vel_y, nb.h_vel_z};
    for (int k = 0; k < 7; k++) std::memcpy(state + (size_t)k * n, e[k], sizeof(double) * n);
    double* f[6] = {nb.h_acc_x, nb.h_acc_y, nb.h_acc_z,
                    nb.h_adot_x, nb.h_adot_y, nb.h_adot_z};
    for (int k = 0; k < 6; k++) std::memcpy(force + (size_t)k * n, f[k], sizeof(double) * n);
    return 0;
}

}  // extern "C"
