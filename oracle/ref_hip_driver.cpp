// ref_hip_driver.cpp -- extern "C" handle onto the reference's integrator
// running on HIP forces (ref_hip_force.cpp).
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into
// oracle/_ref/libref_hip.so, loaded by tests/test_gpu_reference_integrator.py.
#include <cstring>
#include <mpi.h>
#include <omp.h>

#include "kernel.h"
#include "nbody.h"

extern "C" void refhip_select(int precision, int pairing);

extern "C" {

// The reference's NBody (src/nbody.cpp:14-66) from a given state: its
// initial force_calculation and initialdt (NBody::cycle, src/nbody.cpp:310-
// 336), then nsteps calls of NBody::timestepping (src/nbody.cpp:224-289)
// with dt fixed to dt_fixed before each (updatedt still runs inside,
// its value is discarded).  state: 7*n (x, y, z, w, vx, vy, vz), overwritten
// with the final state; force: 6*n, the final a and jerk.
int refhip_run(int n, int precision, int pairing, double* state, int nsteps, double dt_fixed, double* force)
{
    omp_set_num_threads(1);
    int inited = 0;
    MPI_Initialized(&inited);
    if (!inited) {
        int argc = 0;
        char** argv = nullptr;
        MPI_Init(&argc, &argv);
    }
    refhip_select(precision, pairing);
    NBody nb(n, 1.0, 1.0, 0.01, 1.0);
    double* s[7] = {nb.h_par_x, nb.h_par_y, nb.h_par_z, nb.h_par_w, nb.h_vel_x, nb.h_vel_y, nb.h_vel_z};
    for (int k = 0; k < 7; k++) std::memcpy(s[k], state + (size_t)k * n, sizeof(double) * n);
    force_calculation(nb.h_par_x, nb.h_par_y, nb.h_par_z, nb.h_par_w, nb.h_vel_x, nb.h_vel_y, nb.h_vel_z,
                      nb.h_acc_x, nb.h_acc_y, nb.h_acc_z, nb.h_adot_x, nb.h_adot_y, nb.h_adot_z, nb.sys);
    nb.dt = nb.initialdt();
    for (int k = 0; k < nsteps; k++) {
        nb.dt = dt_fixed;
        nb.timestepping();
    }
    double* e[7] = {nb.h_par_x, nb.h_par_y, nb.h_par_z, nb.h_par_w, nb.h_vel_x, nb.h_vel_y, nb.h_vel_z};
    for (int k = 0; k < 7; k++) std::memcpy(state + (size_t)k * n, e[k], sizeof(double) * n);
    double* f[6] = {nb.h_acc_x, nb.h_acc_y, nb.h_acc_z, nb.h_adot_x, nb.h_adot_y, nb.h_adot_z};
    for (int k = 0; k < 6; k++) std::memcpy(force + (size_t)k * n, f[k], sizeof(double) * n);
    return nb.sys.num_cores;
}

}  // extern "C"
