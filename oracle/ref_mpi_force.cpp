// ref_mpi_force.cpp -- the reference's CPU force_calculation timed the way
// the reference runs it: one MPI rank per core, OMP_NUM_THREADS=1.
//
// BENCHMARK BASELINE ONLY.  oracle/Makefile links this file with the
// reference's own src/kernel.cpp and src/utils.cpp (FP64 and _F32 builds)
// into oracle/_ref/ref_mpi_force{64,32}; bench.py's cpu_baseline leg starts
// it under mpiexec.  Every rank holds all n bodies and takes the cyclic rows
// i = rank, rank + size, ... of the Newton's-third-law triangle
// (src/kernel.cpp:1185), as NBody sets up (src/nbody.cpp:14-33); the timed
// region is force_calculation alone, the maximum over ranks, like the
// reference's EVAL column (src/nbody.cpp:231-237, :346-353).  The partial
// forces are then summed with MPI_Allreduce (src/nbody.cpp:240-246) outside
// the timing, and rank 0 prints one line: n, ranks, the per-repetition
// times, and a checksum of the summed forces.
//
// usage: ref_mpi_force STATE_FILE N REPS   (STATE_FILE: 7*N raw doubles,
//        x y z m vx vy vz blocks)
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <mpi.h>
#include <omp.h>

#include "kernel.h"

int main(int argc, char** argv)
{
    MPI_Init(&argc, &argv);
    int rank = 0, size = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc != 4) {
        if (rank == 0) std::fprintf(stderr, "usage: %s STATE_FILE N REPS\n", argv[0]);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    const int n = std::atoi(argv[2]);
    const int reps = std::atoi(argv[3]);
    std::vector<double> s((size_t)7 * n), f((size_t)6 * n);
    FILE* fp = std::fopen(argv[1], "rb");
    if (!fp || std::fread(s.data(), sizeof(double), s.size(), fp) != s.size()) {
        std::fprintf(stderr, "rank %d: cannot read %d bodies from %s\n", rank, n, argv[1]);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    std::fclose(fp);
    Sys sys;
    sys.N = n;
    sys.rank = rank;
    sys.size = size;
    sys.start = 0;
    sys.end = n;
    double* p = s.data();
    double* q = f.data();
    std::vector<double> times;
    for (int r = 0; r < reps; ++r) {
        MPI_Barrier(MPI_COMM_WORLD);
        const double t0 = MPI_Wtime();
        force_calculation(p, p + n, p + 2 * (size_t)n, p + 3 * (size_t)n, p + 4 * (size_t)n, p + 5 * (size_t)n,
                          p + 6 * (size_t)n, q, q + n, q + 2 * (size_t)n, q + 3 * (size_t)n, q + 4 * (size_t)n,
                          q + 5 * (size_t)n, sys);
        double t = MPI_Wtime() - t0, tmax = 0.0;
        MPI_Reduce(&t, &tmax, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
        times.push_back(tmax);
        MPI_Allreduce(MPI_IN_PLACE, q, 6 * n, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    }
    if (rank == 0) {
        double sum = 0.0;
        for (double v : f) sum += v < 0 ? -v : v;
        std::printf("ref_mpi_force n=%d ranks=%d threads=%d times", n, size, omp_get_max_threads());
        for (double t : times) std::printf(" %.6f", t);
        std::printf(" abssum %.17g\n", sum);
    }
    MPI_Finalize();
    return 0;
}
