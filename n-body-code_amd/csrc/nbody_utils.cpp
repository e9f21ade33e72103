// nbody_utils.cpp -- the rest of the reference's host utility API
// (include/utils.h:29,55-63; src/utils.cpp:117-123,155-185,187-275) for code
// written against it: timer, InvSqrt2D / InvSqrt2, virial_ratio and the
// _DEBUG brute-force check test().  compute_kinetic, compute_potential and
// estimate_err live beside force_calculation (kernel_hip.cpp).
#include "nbody.hpp"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

// Wall-clock seconds since the epoch (src/utils.cpp:117-123 reads
// gettimeofday; the system clock is the same clock).
double timer()
{
    using namespace std::chrono;
    return duration<double>(system_clock::now().time_since_epoch()).count();
}

// 1/sqrt(x) by the published magic-constant method with modified Newton
// steps (Walczyk, Moroz, Ciesielski 2021), the reference's InvSqrt2D
// (src/utils.cpp:155-165): the same constant, coefficients and operation
// order, so results are bitwise the reference's; bit patterns are moved
// with memcpy rather than through pointer casts.
double InvSqrt2D(double x)
{
    const double halfx = 0.5 * x;
    int64_t i;
    std::memcpy(&i, &x, sizeof i);
    i = 0x5FE6ED2102DCBFDALL - (i >> 1);
    double y;
    std::memcpy(&y, &i, sizeof y);
    y *= 1.50087895511633457 - halfx * y * y;
    y *= 1.50000057967625766 - halfx * y * y;
    y *= 1.5000000000002520 - halfx * y * y;
    y *= 1.5000000000000000 - halfx * y * y;
    return y;
}

// FP32 form of the same method with two steps (src/utils.cpp:167-175).
float InvSqrt2(float x)
{
    const float halfx = 0.5f * x;
    int32_t i;
    std::memcpy(&i, &x, sizeof i);
    i = 0x5F376908 - (i >> 1);
    float y;
    std::memcpy(&y, &i, sizeof y);
    y *= 1.50087896f - halfx * y * y;
    y *= 1.50000057f - halfx * y * y;
    return y;
}

// Q = -2T/U with the reference's energy definitions (src/utils.cpp:177-181).
double virial_ratio(double* h_par_x, double* h_par_y, double* h_par_z, double* h_par_w, double* h_vel_x,
                    double* h_vel_y, double* h_vel_z, int N)
{
    const double T = compute_kinetic(h_par_w, h_vel_x, h_vel_y, h_vel_z, N);
    const double U = compute_potential(h_par_x, h_par_y, h_par_z, h_par_w, N, true);
    return -2.0 * T / U;
}

// The reference's _DEBUG check (src/utils.cpp:187-275): acceleration and
// jerk recomputed on the host by brute force in FP64 with 1/std::sqrt, every
// ordered pair, then the per-component relative differences of the given
// forces, printed for the first PRT = 5 bodies (include/utils.h:13) with the
// largest over all bodies (which the reference computes but discards).  The
// softening is the force backend's eps2 (the reference's is its fixed
// EPS2 = 1e-7, the backend's default).  A debugging aid: nothing feeds its
// values back into the run.
void test(double* h_par_x, double* h_par_y, double* h_par_z, double* h_par_w, double* h_vel_x,
          double* h_vel_y, double* h_vel_z, double* h_acc_x, double* h_acc_y, double* h_acc_z,
          double* h_adot_x, double* h_adot_y, double* h_adot_z, int N)
{
    constexpr int kPrint = 5;
    const double eps2 = force_backend_eps2();
    std::vector<double> f((size_t)6 * N, 0.0);
    for (int i = 0; i < N; i++) {
        double s[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < N; j++) {
            if (j == i) continue;
            const double d[3] = {h_par_x[j] - h_par_x[i], h_par_y[j] - h_par_y[i], h_par_z[j] - h_par_z[i]};
            const double u[3] = {h_vel_x[j] - h_vel_x[i], h_vel_y[j] - h_vel_y[i], h_vel_z[j] - h_vel_z[i]};
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps2;
            const double rinv = 1.0 / std::sqrt(r2);
            const double mr3 = h_par_w[j] * rinv * rinv * rinv;
            const double q = -3.0 * (d[0] * u[0] + d[1] * u[1] + d[2] * u[2]) * rinv * rinv;
            for (int c = 0; c < 3; c++) {
                s[c] += mr3 * d[c];
                s[3 + c] += mr3 * (u[c] + q * d[c]);
            }
        }
        for (int c = 0; c < 6; c++) f[(size_t)c * N + i] = s[c];
    }
    const double* got[6] = {h_acc_x, h_acc_y, h_acc_z, h_adot_x, h_adot_y, h_adot_z};
    const char* title[2] = {"acc", "adot"};
    for (int block = 0; block < 2; block++) {
        std::printf("Relative difference of %s from the host brute force:\n", title[block]);
        double worst[3] = {0, 0, 0};
        for (int i = 0; i < N; i++) {
            double rel[3];
            for (int c = 0; c < 3; c++) {
                const double ref = f[(size_t)(3 * block + c) * N + i];
                rel[c] = std::abs((ref - got[3 * block + c][i]) / ref);
                if (rel[c] > worst[c]) worst[c] = rel[c];
            }
            if (i < kPrint) std::printf("%.12g  %.12g  %.12g\n", rel[0], rel[1], rel[2]);
        }
        std::printf("max %s: %.12g  %.12g  %.12g\n", title[block], worst[0], worst[1], worst[2]);
    }
}
