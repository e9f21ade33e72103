// nbody.hpp -- host-side integrator with the reference's interface
// (include/nbody.h:24-59, include/utils.h:31-53, include/kernel.h:15-19),
// driving the HIP backend of include/nbody_hip.h.
//
// The body-state arrays, the method names and the step order are the
// reference's, so code written against NBody keeps working; the force
// evaluation behind force_calculation is the MI355X pair kernel, and MPI is
// dropped (one process, one GPU: sys.size == 1, start/end cover all bodies).
// A device-resident mode (NBody::resident) runs the whole step on the GPU.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "nbody_hip.h"

// Run metadata and timers, include/utils.h:31-53.
struct Sys {
    int N = 0;
    int start = 0;
    int end = 0;
    int rank = 0;
    int size = 1;
    int nthreads = 1;
    int thid = 0;
    double t_eval = 0.0;
    double t_pred = 0.0;
    double t_corr = 0.0;
    double t_eval_init = 0.0;
    double kernel_time = 0.0;
    double kernel_time_init = 0.0;
    std::vector<double> t_evalv;
    std::vector<double> t_predv;
    std::vector<double> t_corrv;
    double initE = 0.0;
    double finE = 0.0;
    double Q = 0.0;
    double Err = 0.0;
    int num_cores = 0;
};

// Options of the process-wide force backend behind force_calculation.
struct ForceBackend {
    int device = 0;
    int precision = NBH_FP64;
    double eps2 = 1.0e-7;  // the reference's EPS2 (include/kernel.h:13)
    int pairing = NBH_PAIR_DIRECTED;
    int sources = NBH_SRC_AUTO;  // how the directed kernels read their sources
};

// Sets (or changes) the backend; the context is opened on first use.
void force_backend_configure(const ForceBackend& fb);
// The backend's context (opened on demand); exits the process on failure.
nbh_ctx* force_backend_context();
void force_backend_shutdown();

// The drop-in boundary: same prototype as include/kernel.h:15-19.  Host
// arrays in, outputs fully overwritten, synchronous.  Sets sys.kernel_time
// (seconds of device time of the pair kernels) and sys.num_cores (CUs).
// On a backend error it prints the error and exits with status 1.
void force_calculation(double* __restrict h_par_x, double* __restrict h_par_y,
                       double* __restrict h_par_z, double* __restrict h_par_w,
                       double* __restrict h_vel_x, double* __restrict h_vel_y,
                       double* __restrict h_vel_z, double* __restrict h_acc_x,
                       double* __restrict h_acc_y, double* __restrict h_acc_z,
                       double* __restrict h_adot_x, double* __restrict h_adot_y,
                       double* __restrict h_adot_z, Sys& sys);

// Energies (src/utils.cpp:125-185), evaluated on the GPU.  compat = true
// reproduces the reference's potential (weight m_j^2, eps2 = 1e-7).
double compute_kinetic(double* h_par_w, double* h_vel_x, double* h_vel_y, double* h_vel_z, int N);
double compute_potential(double* h_par_x, double* h_par_y, double* h_par_z, double* h_par_w, int N,
                         bool compat = true);
double estimate_err(double En, double E0);

// The rest of the reference's host utilities (include/utils.h:29,55-63).
double timer();                 // wall-clock seconds
double InvSqrt2D(double x);     // the reference's fast 1/sqrt, bitwise (src/utils.cpp:155-165)
float InvSqrt2(float x);        // its FP32 form (src/utils.cpp:167-175)
double virial_ratio(double* h_par_x, double* h_par_y, double* h_par_z, double* h_par_w, double* h_vel_x,
                    double* h_vel_y, double* h_vel_z, int N);  // -2T/U, reference energies
// _DEBUG brute-force check of a force evaluation (src/utils.cpp:187-275).
void test(double* h_par_x, double* h_par_y, double* h_par_z, double* h_par_w, double* h_vel_x,
          double* h_vel_y, double* h_vel_z, double* h_acc_x, double* h_acc_y, double* h_acc_z,
          double* h_adot_x, double* h_adot_y, double* h_adot_z, int N);
// Softening of the force backend (the eps2 force_calculation uses).
double force_backend_eps2();

class NBody {
public:
    double alpha;
    double n = -2.35;  // Salpeter index (include/nbody.h:28)
    double T = 0.0, U = 0.0;
    int size = 1, rank = 0;
    Sys sys;

    double *h_par_x, *h_par_y, *h_par_z, *h_par_w;
    double *h_vel_x, *h_vel_y, *h_vel_z;
    double *h_acc_x, *h_acc_y, *h_acc_z;
    double *h_adot_x, *h_adot_y, *h_adot_z;
    double *h_acc1_x, *h_acc1_y, *h_acc1_z;
    double *h_a1dot_x, *h_a1dot_y, *h_a1dot_z;
    double *h_a2dot_x, *h_a2dot_y, *h_a2dot_z;
    double *h_a3dot_x, *h_a3dot_y, *h_a3dot_z;
    double dt = 0.0;

    // Extensions (no reference counterpart).
    double fixed_dt = 0.0;          // > 0: every step uses it; 0: adaptive (reference)
    std::string ic = "ref";         // "ref" (NBody::initialize) or "plummer"
    uint64_t seed = 1;              // Plummer seed
    bool compat_energy = true;      // reference energy definition
    bool resident = false;          // whole step on the GPU
    bool use_graph = false;         // resident: replay steps from a hipGraph
    std::string load_path;          // initial state from a state file (nbh_state_load)
    double sim_time = 0.0;          // simulated time advanced by cycle()

    NBody(int Nparticles, double v_max, double r_max, double m_min, double m_max);
    ~NBody();
    NBody(const NBody&) = delete;
    NBody& operator=(const NBody&) = delete;

    void initialize();
    void initializewithzeros();
    double initialdt();
    void prediction();
    void snapandcrackle();
    void correction();
    double updatedt();
    void timestepping();
    void cycle(int ncycle);

private:
    int N;
    double vmax, rmax, mmin, mmax;
    void cycle_resident(int ncycle);
};
