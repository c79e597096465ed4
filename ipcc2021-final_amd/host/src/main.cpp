// main.cpp -- command-line driver with the reference's arguments
// (src/main.cpp:17-123): Mass gaugefile Lx Ly Lz Lt sx sy sz st [seed].
// One GPU holds the whole lattice, so the subgrid must equal the lattice.
#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "check.h"
#include "invert.h"
#include "load_gauge.h"

using namespace std;

static double now()
{
    return chrono::duration<double>(chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv)
{
    int site_vec[4] = {8, 8, 8, 8};
    int subgs[4] = {8, 8, 8, 8};
    double Mass = 0.05;
    string filename = "UNIT";
    int seedr = 0;
    if (argc == 11 || argc == 12) {
        Mass = stod(argv[1]);
        filename = string(argv[2]);
        for (int i = 0; i < 4; ++i) {
            site_vec[i] = stoi(string(argv[i + 3]));
            subgs[i] = stoi(string(argv[i + 7]));
        }
        if (argc == 12)
            seedr = atoi(argv[11]);
    } else if (argc != 1) {
        cerr << "./wilson_main Mass gaugefile  lattice_size   subgrid_size [seed]" << endl;
        cerr << "Example: ./wilson_main 0.005 UNIT  24 24 24 48  24 24 24 48" << endl;
        return 1;
    }
    const int MaxCG = 500;         // src/main.cpp:47
    const double Accuracy = 1.0e-9; // src/main.cpp:48
    long Vsub = 1;
    for (int i = 0; i < 4; ++i) {
        if (site_vec[i] != subgs[i]) {
            cerr << "Lattice partition size != 1: one GPU holds the whole lattice" << endl;
            return 1;
        }
        Vsub *= subgs[i];
    }
    cout << "  Mass  " << Mass << endl;
    cout << "  gauge file  " << filename << endl;
    printf("  lattice size  %d %d %d %d\n", site_vec[0], site_vec[1], site_vec[2], site_vec[3]);
    printf("  subgrid size  %d %d %d %d\n", subgs[0], subgs[1], subgs[2], subgs[3]);

    srand(0 + seedr); // rank 0 + seed (src/main.cpp:77-78)
    vector<double> source(Vsub * 24);
    for (auto &v : source)
        v = (double)rand() / RAND_MAX;
    const double start_t = now();
    lattice_fermion src(subgs, site_vec);
    for (long i = 0; i < Vsub * 12; ++i)
        src.A[i] = complex<double>(source[2 * i], source[2 * i + 1]);
    lattice_gauge U(subgs, site_vec);
    load_gauge(U.A, filename, subgs, site_vec);
    const double gauge_t = now();
    lattice_fermion dest(subgs, site_vec);
    CGinvert(src, dest, U, Mass, MaxCG, Accuracy);
    const double end_t = now();
    cout << "Source and Gauge, time: " << gauge_t - start_t << endl;
    cout << "CG invert, time: " << end_t - gauge_t << endl;
    cout << "Total time: " << end_t - start_t << endl;
    cout << "Result checking ..." << endl;
    cout << " [ Accuracy = " << Accuracy << " ]" << endl;
    check(src, dest, U, Mass);
    return 0;
}
