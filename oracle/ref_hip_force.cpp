// ref_hip_force.cpp -- the `_HIP` branch of force_calculation, built
// against the reference's own headers and integrator.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file with the
// reference's src/nbody.cpp and src/utils.cpp, where they lie under the
// reference checkout, in place of its src/kernel.cpp, and links
// libnbody_hip.so: the reference's NBody::cycle/timestepping
// (src/nbody.cpp:224-380) then run on HIP forces, exactly as a maintainer
// who adds the branch shown in INTEGRATION.md to src/kernel.cpp's #if chain
// (src/kernel.cpp:47,399,772,1101,1164) would run them.  The prototype and
// Sys come from the reference's include/kernel.h:15-19 and
// include/utils.h:31-53; the softening is its EPS2 (include/kernel.h:13).
#include <cstdio>
#include <mpi.h>

#include "kernel.h"
#include "nbody_hip.h"

namespace {
nbh_ctx* g_ctx = nullptr;
int g_precision = NBH_FP64;
int g_pairing = NBH_PAIR_DIRECTED;

[[noreturn]] void fatal(const char* what)
{
    std::fprintf(stderr, "%s: %s\n", what, nbh_last_error());
    MPI_Abort(MPI_COMM_WORLD, 1);
    __builtin_unreachable();
}
}  // namespace

// Precision and pairing of the context opened by the next first call (the
// test driver's switch; a build would fix them at compile time).
extern "C" void refhip_select(int precision, int pairing)
{
    if (g_ctx) nbh_destroy(g_ctx);
    g_ctx = nullptr;
    g_precision = precision;
    g_pairing = pairing;
}

void force_calculation(double* __restrict h_par_x, double* __restrict h_par_y, double* __restrict h_par_z,
                       double* __restrict h_par_w, double* __restrict h_vel_x, double* __restrict h_vel_y,
                       double* __restrict h_vel_z, double* __restrict h_acc_x, double* __restrict h_acc_y,
                       double* __restrict h_acc_z, double* __restrict h_adot_x, double* __restrict h_adot_y,
                       double* __restrict h_adot_z, Sys& sys)
{
    if (!g_ctx) {
        nbh_config cfg = {/*device*/ 0, g_precision, /*eps2*/ EPS2, /*block*/ 0, /*jsplit*/ 0,
                          /*ipl*/ 0, g_pairing, /*sources*/ NBH_SRC_AUTO};
        if (nbh_create(&cfg, &g_ctx) != NBH_OK) fatal("nbh_create");
    }
    double ms = 0.0;
    if (nbh_force(g_ctx, sys.N, h_par_x, h_par_y, h_par_z, h_par_w, h_vel_x, h_vel_y, h_vel_z, h_acc_x,
                  h_acc_y, h_acc_z, h_adot_x, h_adot_y, h_adot_z, &ms) != NBH_OK)
        fatal("force_calculation");
    sys.kernel_time = ms * 1e-3;  // the EVAL column, as the TT modes use it (src/nbody.cpp:234-235)
    if (nbh_device_info(0, &sys.num_cores, nullptr, nullptr, 0) != NBH_OK) fatal("nbh_device_info");
}
