// ref_bench.cpp -- times the reference's own DslashNew on several MPI ranks.
//
// TEST INFRASTRUCTURE ONLY (the CPU baseline of bench.py).  Linked by
// oracle/Makefile with the reference's src/*.cpp, compiled where they lie,
// into oracle/_ref/ref_bench.  The reference runs one single-threaded MPI rank
// per core, each owning a sub-lattice subgs = L / grid (src/main.cpp:50-66,
// src/sub.sh:5-7), with the halo exchange inside DslashoffdNew.  This driver
// sets up that decomposition, fills the local source and links with seeded
// uniform numbers (the reference's own generator kind, src/main.cpp:80-84; the
// operator's cost does not depend on the values), calls the full-lattice
// DslashNew (src/invert.cpp:138-156) on every rank until about `seconds` of
// wall time have passed, and prints one JSON line on rank 0.
//
//   mpiexec -n N ref_bench Lx Ly Lz Lt  gx gy gz gt  seconds
#include <mpi.h>

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "invert.h"
#include "lattice_fermion.h"
#include "lattice_gauge.h"

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int rank = 0, size = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc != 10) {
        if (rank == 0)
            std::fprintf(stderr, "usage: ref_bench Lx Ly Lz Lt gx gy gz gt seconds\n");
        MPI_Finalize();
        return 2;
    }
    int site_vec[4], subgs[4], grid[4];
    long Vsub = 1, V = 1;
    int ranks = 1;
    for (int i = 0; i < 4; ++i) {
        site_vec[i] = std::atoi(argv[1 + i]);
        grid[i] = std::atoi(argv[5 + i]);
        if (grid[i] < 1 || site_vec[i] % grid[i] != 0) {
            if (rank == 0)
                std::fprintf(stderr, "ref_bench: extent %d not divisible by %d\n", site_vec[i],
                             grid[i]);
            MPI_Finalize();
            return 2;
        }
        subgs[i] = site_vec[i] / grid[i];
        Vsub *= subgs[i];
        V *= site_vec[i];
        ranks *= grid[i];
    }
    const double seconds = std::atof(argv[9]);
    if (ranks != size) {
        if (rank == 0)
            std::fprintf(stderr, "ref_bench: grid has %d ranks, MPI size is %d\n", ranks, size);
        MPI_Finalize();
        return 2;
    }
    srand(20211026 + rank);
    lattice_fermion src(subgs, site_vec), dest(subgs, site_vec);
    lattice_gauge U(subgs, site_vec);
    for (long i = 0; i < Vsub * 12; ++i)
        src.A[i] = std::complex<double>((double)rand() / RAND_MAX, (double)rand() / RAND_MAX);
    for (int mu = 0; mu < 4; ++mu)
        for (long i = 0; i < Vsub * 9; ++i)
            U.A[mu][i] = std::complex<double>((double)rand() / RAND_MAX - 0.5,
                                              (double)rand() / RAND_MAX - 0.5);
    const double mass = 1.0 / (2.0 * 0.135) - 4.0;

    DslashNew(src, dest, U, mass, false); // first touch of every buffer
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    DslashNew(src, dest, U, mass, false);
    MPI_Barrier(MPI_COMM_WORLD);
    double one = MPI_Wtime() - t0;
    MPI_Bcast(&one, 1, MPI_DOUBLE, 0, MPI_COMM_WORLD);
    int n = (int)(seconds / (one > 1e-6 ? one : 1e-6));
    if (n < 1)
        n = 1;
    MPI_Barrier(MPI_COMM_WORLD);
    t0 = MPI_Wtime();
    for (int k = 0; k < n; ++k)
        DslashNew(src, dest, U, mass, false);
    MPI_Barrier(MPI_COMM_WORLD);
    const double el = MPI_Wtime() - t0;
    if (rank == 0) {
        const double gflops = 1368.0 * (double)V * n / el / 1e9;
        std::printf("{\"value\": %.3f, \"unit\": \"GFLOP/s\", \"cores\": %d, \"kind\": "
                    "\"reference\", \"sample\": \"%d DslashNew applications (full-lattice M, "
                    "FP64) at %dx%dx%dx%d on %d MPI ranks of %dx%dx%dx%d sites in %.1f s\"}\n",
                    gflops, size, n, site_vec[0], site_vec[1], site_vec[2], site_vec[3], size,
                    subgs[0], subgs[1], subgs[2], subgs[3], el);
        std::fflush(stdout);
    }
    MPI_Finalize();
    return 0;
}
