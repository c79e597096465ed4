// api_selftest.cpp -- drives every operator entry point of the reference-style
// C++ API (host/include/*.h) on raw binary inputs and writes the outputs, so a
// test can compare them with the reference's golden vectors.
//
// usage: api_selftest DIR Lx Ly Lz Lt mass
//   reads  DIR/src.bin (complex<double>[V*12]) and DIR/gauge.bin (4 x complex<double>[V*9])
//   writes DIR/<name>.bin for each entry point listed below
// usage: api_selftest --time Lx Ly Lz Lt reps
//   times DslashNew on seeded inputs, uploading the links on every call and
//   with a gauge_resident guard, and prints both averages
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <chrono>

#include "dslash.h"
#include "dslash_new.h"
#include "gauge_resident.h"
#include "invert.h"

typedef std::complex<double> cd;

static bool read_all(const std::string &fn, void *p, size_t bytes)
{
    FILE *f = std::fopen(fn.c_str(), "rb");
    if (!f)
        return false;
    const bool ok = std::fread(p, 1, bytes, f) == bytes;
    std::fclose(f);
    return ok;
}

static void write_all(const std::string &fn, const void *p, size_t bytes)
{
    FILE *f = std::fopen(fn.c_str(), "wb");
    if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
        std::fprintf(stderr, "cannot write %s\n", fn.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

static int time_mode(char **argv)
{
    int L[4], S[4];
    long V = 1;
    for (int i = 0; i < 4; ++i) {
        L[i] = S[i] = std::atoi(argv[2 + i]);
        V *= L[i];
    }
    const int reps = std::atoi(argv[6]);
    lattice_fermion src(S, L), out(S, L);
    lattice_gauge U(S, L);
    srand(5);
    for (long k = 0; k < 12 * V; ++k)
        src.A[k] = cd((double)rand() / RAND_MAX, (double)rand() / RAND_MAX);
    for (int mu = 0; mu < 4; ++mu)
        for (long k = 0; k < 9 * V; ++k)
            U.A[mu][k] = cd((double)rand() / RAND_MAX - 0.5, (double)rand() / RAND_MAX - 0.5);
    auto run = [&]() {
        DslashNew(src, out, U, 0.1, false); // first call: context, fields
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < reps; ++k)
            DslashNew(src, out, U, 0.1, false);
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                   .count() /
               reps;
    };
    const double up = run();
    double kept;
    {
        gauge_resident keep(U);
        kept = run();
    }
    std::printf("api_selftest: DslashNew %dx%dx%dx%d, links uploaded per call %.3f ms, "
                "gauge_resident %.3f ms\n",
                L[0], L[1], L[2], L[3], up, kept);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 7 && std::string(argv[1]) == "--time")
        return time_mode(argv);
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s DIR Lx Ly Lz Lt mass\n", argv[0]);
        return 1;
    }
    const std::string dir = argv[1];
    int L[4], S[4];
    long V = 1;
    for (int i = 0; i < 4; ++i) {
        L[i] = S[i] = std::atoi(argv[2 + i]);
        V *= L[i];
    }
    const double mass = std::atof(argv[6]);
    lattice_fermion src(S, L), out(S, L);
    lattice_gauge U(S, L);
    std::vector<cd> g(4 * 9 * V);
    if (!read_all(dir + "/src.bin", src.A, sizeof(cd) * 12 * V) ||
        !read_all(dir + "/gauge.bin", g.data(), g.size() * sizeof(cd))) {
        std::fprintf(stderr, "cannot read inputs in %s\n", dir.c_str());
        return 1;
    }
    for (int mu = 0; mu < 4; ++mu)
        for (long k = 0; k < 9 * V; ++k)
            U.A[mu][k] = g[mu * 9 * V + k];
    const size_t bytes = sizeof(cd) * 12 * V;
    for (int dag = 0; dag < 2; ++dag) {
        for (int cb = 0; cb < 2; ++cb) {
            DslashoffdNew(src, out, U, dag != 0, cb);
            write_all(dir + "/hopnew_d" + std::to_string(dag) + "_cb" + std::to_string(cb) + ".bin",
                      out.A, bytes);
            Dslashoffd(src, out, U, dag != 0, cb);
            write_all(dir + "/hop_d" + std::to_string(dag) + "_cb" + std::to_string(cb) + ".bin",
                      out.A, bytes);
        }
        (dag ? DslashOENew : DslashEONew)(src, out, U, dag != 0);
        write_all(dir + "/eo_new_d" + std::to_string(dag) + ".bin", out.A, bytes);
        DslashNew(src, out, U, mass, dag != 0);
        write_all(dir + "/Mnew_d" + std::to_string(dag) + ".bin", out.A, bytes);
        Dslash(src, out, U, mass, dag != 0);
        write_all(dir + "/M_d" + std::to_string(dag) + ".bin", out.A, bytes);
    }
    DslashEEOONew(src, out, mass);
    write_all(dir + "/diag.bin", out.A, bytes);
    DslashEE(src, out, mass);
    write_all(dir + "/diag_ee.bin", out.A, bytes);
    DslashOO(src, out, mass);
    write_all(dir + "/diag_oo.bin", out.A, bytes);
    // the same operators with the links kept on the device by a guard
    {
        gauge_resident keep(U);
        for (int dag = 0; dag < 2; ++dag) {
            DslashNew(src, out, U, mass, dag != 0);
            write_all(dir + "/Mnew_res_d" + std::to_string(dag) + ".bin", out.A, bytes);
            DslashoffdNew(src, out, U, dag != 0, 1);
            write_all(dir + "/hopnew_res_d" + std::to_string(dag) + ".bin", out.A, bytes);
        }
    }
    // CGinvert on the same system, start vector from srand(11)
    srand(11);
    lattice_fermion x(S, L);
    CGinvert(src, x, U, mass, 1000, 1e-9);
    write_all(dir + "/cg_x.bin", x.A, bytes);
    // after the guard is gone the links are read from U again: all-zero links
    // make M = (4 + m) 1
    for (int mu = 0; mu < 4; ++mu)
        for (long k = 0; k < 9 * V; ++k)
            U.A[mu][k] = 0.0;
    DslashNew(src, out, U, mass, false);
    write_all(dir + "/Mnew_zero.bin", out.A, bytes);
    std::printf("api_selftest: done\n");
    return 0;
}
