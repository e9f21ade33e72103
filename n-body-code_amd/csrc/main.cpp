// main.cpp -- nbody_app: the reference CLI (src/main.cpp:13-143) on the
// MI355X backend.  --N and --Ncycle behave as in the reference (positive
// integers; unknown or incomplete arguments print usage and exit 1); the
// other flags are extensions for the benchmark configurations.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>

#include "nbody.hpp"

namespace {
constexpr int kCol = 15;

const char* kUsage =
    "Usage: ./nbody_app [--N <value>] [--Ncycle <value>]\n"
    "  extensions: [--eps <softening>] [--dt <fixed step>] [--ic ref|plummer] [--seed <int>]\n"
    "              [--precision fp64|mixed] [--pairing directed|symmetric] [--sources auto|lds|scalar]\n"
    "              [--resident] [--graph]\n"
    "              [--energy compat|physical]\n"
    "              [--device <ordinal>] [--load <state file>] [--save <state file>]";

int positive_int(const std::string& name, const char* text)
{
    const int v = std::stoi(text);
    if (v <= 0) throw std::invalid_argument(name + " must be positive.");
    return v;
}

double wall()
{
    return std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch()).count();
}
}  // namespace

int main(int argc, char* argv[])
{
    int N = 16384, Ncycle = 10;
    ForceBackend fb;
    double fixed_dt = 0.0;
    std::string ic = "ref", energy = "compat", load_path, save_path;
    bool n_given = false;
    uint64_t seed = 1;
    bool resident = false, graph = false;

    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        const bool has = i + 1 < argc;
        try {
            if (arg == "--N" && has) {
                N = positive_int("N", argv[++i]);
                n_given = true;
            } else if (arg == "--Ncycle" && has) {
                Ncycle = positive_int("Ncycle", argv[++i]);
            } else if (arg == "--eps" && has) {
                const double e = std::stod(argv[++i]);
                if (!(e >= 0.0)) throw std::invalid_argument("eps must be >= 0.");
                fb.eps2 = e * e;
            } else if (arg == "--dt" && has) {
                fixed_dt = std::stod(argv[++i]);
                if (!(fixed_dt > 0.0)) throw std::invalid_argument("dt must be positive.");
            } else if (arg == "--ic" && has) {
                ic = argv[++i];
                if (ic != "ref" && ic != "plummer") throw std::invalid_argument("ic must be ref or plummer.");
            } else if (arg == "--seed" && has) {
                seed = std::stoull(argv[++i]);
            } else if (arg == "--precision" && has) {
                const std::string p = argv[++i];
                if (p == "fp64") fb.precision = NBH_FP64;
                else if (p == "mixed") fb.precision = NBH_MIXED;
                else throw std::invalid_argument("precision must be fp64 or mixed.");
            } else if (arg == "--pairing" && has) {
                const std::string p = argv[++i];
                if (p == "directed") fb.pairing = NBH_PAIR_DIRECTED;
                else if (p == "symmetric") fb.pairing = NBH_PAIR_SYMMETRIC;
                else throw std::invalid_argument("pairing must be directed or symmetric.");
            } else if (arg == "--sources" && has) {
                const std::string p = argv[++i];
                if (p == "auto") fb.sources = NBH_SRC_AUTO;
                else if (p == "lds") fb.sources = NBH_SRC_LDS;
                else if (p == "scalar") fb.sources = NBH_SRC_SCALAR;
                else throw std::invalid_argument("sources must be auto, lds or scalar.");
            } else if (arg == "--energy" && has) {
                energy = argv[++i];
                if (energy != "compat" && energy != "physical")
                    throw std::invalid_argument("energy must be compat or physical.");
            } else if (arg == "--device" && has) {
                fb.device = std::stoi(argv[++i]);
            } else if (arg == "--load" && has) {
                load_path = argv[++i];
            } else if (arg == "--save" && has) {
                save_path = argv[++i];
            } else if (arg == "--resident") {
                resident = true;
            } else if (arg == "--graph") {
                resident = graph = true;
            } else {
                std::cerr << "Unknown or incomplete argument: " << arg << std::endl;
                std::cerr << kUsage << std::endl;
                return 1;
            }
        } catch (const std::exception& e) {
            std::cerr << "Error parsing " << arg << ": " << e.what() << std::endl;
            return 1;
        }
    }

    if (!load_path.empty()) {
        int nf = 0;
        if (nbh_state_info(load_path.c_str(), &nf, nullptr, nullptr) != NBH_OK) {
            std::cerr << "Error reading --load: " << nbh_last_error() << std::endl;
            return 1;
        }
        if (n_given && nf != N) {
            std::cerr << "Error: --N " << N << " but " << load_path << " holds " << nf << " bodies" << std::endl;
            return 1;
        }
        N = nf;
    }
    std::cout << "Using N = " << N << ", Ncycle = " << Ncycle << std::endl;
    force_backend_configure(fb);

    const double start = wall();
    NBody nbody(N, 1.0, 1.0, 0.01, 1.0);  // vmax, rmax, mmin, mmax of src/main.cpp:73-76
    nbody.fixed_dt = fixed_dt;
    nbody.ic = ic;
    nbody.seed = seed;
    nbody.compat_energy = energy == "compat";
    nbody.resident = resident;
    nbody.use_graph = graph;
    nbody.load_path = load_path;
    nbody.cycle(Ncycle);
    const double total = wall() - start;
    if (!save_path.empty() &&
        nbh_state_save(save_path.c_str(), N, nbody.h_par_x, nbody.h_par_y, nbody.h_par_z, nbody.h_par_w,
                       nbody.h_vel_x, nbody.h_vel_y, nbody.h_vel_z, nbody.sim_time, Ncycle) != NBH_OK) {
        std::cerr << "Error writing --save: " << nbh_last_error() << std::endl;
        return 1;
    }

    std::cout << std::fixed << std::setprecision(9);
    std::cout << "Time stamp(start): " << start << std::endl << std::endl;
    std::cout << std::setw(kCol) << "N" << std::setw(kCol) << "NPES" << std::setw(kCol) << "PRED"
              << std::setw(kCol) << "EVAL" << std::setw(kCol) << "CORR" << std::setw(kCol)
              << "NCYCLE" << std::endl;
    std::cout << std::string(kCol * 6, '-') << std::endl;
    for (int i = 0; i < Ncycle; i++)
        std::cout << std::setw(kCol) << N << std::setw(kCol) << nbody.sys.size << std::setw(kCol)
                  << nbody.sys.t_predv[i] << std::setw(kCol) << nbody.sys.t_evalv[i]
                  << std::setw(kCol) << nbody.sys.t_corrv[i] << std::setw(kCol) << i << std::endl;
    std::cout << "\nInitial Evaluation Time: " << nbody.sys.t_eval_init << std::endl;
    std::cout << "Initial Energy: " << nbody.sys.initE << std::endl;
    std::cout << "Final Energy: " << nbody.sys.finE << std::endl;
    std::cout << "Q = " << nbody.sys.Q << std::endl;
    std::cout << "Err = " << nbody.sys.Err << std::endl;
    std::cout << "WALL time: " << total << " seconds" << std::endl;
    std::cout << "Total active cores: " << nbody.sys.num_cores << std::endl << std::endl;
    std::cout << "Time stamp (end): " << wall() << std::endl;
    force_backend_shutdown();
    return 0;
}
