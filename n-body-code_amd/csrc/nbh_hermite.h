// nbh_hermite.h -- the per-body Hermite predictor and corrector arithmetic,
// shared by the single-system step kernels (k_predict, k_correct in
// nbh_kernels.hip) and the ensemble's fused step kernel (k_ens_advance in
// nbh_ensemble.hip), so that both give the same bits.  Operations in the
// reference's order with no contraction, which its x86-64 build (no FMA)
// also performs.
#pragma once

#include <hip/hip_runtime.h>

namespace nbh {

// Taylor predictor, src/nbody.cpp:152-163: position p and velocity u of one
// axis advanced by dt from acceleration a and jerk j.
__device__ __forceinline__ void hermite_predict_axis(double dt, double a, double j, double& p, double& u)
{
#pragma clang fp contract(off)
    p = p + (u * dt + 0.5 * a * dt * dt + 1.0 / 6.0 * j * dt * dt * dt);
    u = u + (a * dt + 0.5 * j * dt * dt);
}

// Snap and crackle from the forces at both ends of the step, then the
// corrector on the predicted position p and velocity u of one axis,
// src/nbody.cpp:165-196.  invdt2 = 1 / (dt dt), invdt3 = 1 / (dt dt dt).
__device__ __forceinline__ void hermite_correct_axis(double dt, double invdt2, double invdt3, double a0, double j0,
                                                     double a1, double j1, double& p, double& u, double& a2,
                                                     double& a3)
{
#pragma clang fp contract(off)
    a2 = 2.0 * invdt2 * (-3.0 * (a0 - a1) - (2.0 * j0 + j1) * dt);
    a3 = 6.0 * invdt3 * (2.0 * (a0 - a1) + (j0 + j1) * dt);
    const double dt3 = dt * dt * dt;
    const double dt4 = dt3 * dt;
    const double dt5 = dt4 * dt;
    p = p + (1.0 / 24.0 * a2 * dt4 + 1.0 / 120.0 * a3 * dt5);
    u = u + (1.0 / 6.0 * a2 * dt3 + 1.0 / 24.0 * a3 * dt4);
}

}  // namespace nbh
