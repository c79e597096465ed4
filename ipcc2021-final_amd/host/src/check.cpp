// check.cpp -- the reference's post-solve report (src/check.cpp:18-147),
// operators on the GPU, reductions on the host.
#include "check.h"

#include <cmath>
#include <cstdio>
#include <iostream>

#include "invert.h"
#include "operator_mpi.h"

using std::complex;
using std::cout;
using std::endl;

int check(complex<double> *src_p, complex<double> *chi_p, complex<double> *gauge[4],
          const double mass, int *subgs, int *site_vec)
{
    lattice_fermion src(src_p, subgs, site_vec);
    lattice_fermion chi(chi_p, subgs, site_vec);
    lattice_gauge U(gauge, subgs, site_vec);
    check(src, chi, U, mass);
    return 0;
}

int check(lattice_fermion &src, lattice_fermion &chi, lattice_gauge &U, double mass)
{
    lattice_fermion Mx(chi.subgs, chi.site_vec), MMx(chi.subgs, chi.site_vec);
    lattice_fermion Mb(chi.subgs, chi.site_vec), Mdb(chi.subgs, chi.site_vec);
    lattice_fermion MMb(chi.subgs, chi.site_vec), tmp(chi.subgs, chi.site_vec);

    const double nrm_b = std::sqrt(norm_2(src));
    const double nrm_x = std::sqrt(norm_2(chi));
    Dslash(src, Mb, U, mass, false);
    Dslash(src, Mdb, U, mass, true);
    Dslash(Mb, MMb, U, mass, true);
    Dslash(chi, Mx, U, mass, false);
    Minus(Mx, src, tmp);
    const double rsd1 = std::sqrt(norm_2(tmp)); // |M x - b|
    Dslash(Mx, MMx, U, mass, true);
    Minus(MMx, Mdb, tmp);
    const double rsd = std::sqrt(norm_2(tmp)); // |M^dag M x - M^dag b|
    const complex<double> MbMb = vector_p(Mb, Mb), bMMb = vector_p(src, MMb);
    const complex<double> MxMx = vector_p(Mx, Mx), xMMx = vector_p(chi, MMx);
    cout << "===================================" << endl;
    cout << "|b| = " << nrm_b << "    |x| = " << nrm_x << endl;
    cout << "(Mb, Mb) = " << MbMb << "    (b, MMb) = " << bMMb << endl;
    cout << "(Mx, Mx) = " << MxMx << "    (x, MMx) = " << xMMx << endl;
    cout << "|M x -b| = " << rsd1 << endl;
    cout << "|M^dagger M x - Mdb| = " << rsd << endl;
    cout << "===================================" << endl;
    return 0;
}

void check_lattice_site(int *subgs, int *site_vec)
{
    printf("  lattice size  %d %d %d %d\n", site_vec[0], site_vec[1], site_vec[2], site_vec[3]);
    printf("  subgrid size  %d %d %d %d\n", subgs[0], subgs[1], subgs[2], subgs[3]);
}

void gamma5_lattcefermion(lattice_fermion &src, lattice_fermion &dest)
{
    for (int i = 0; i < src.size; i += 12)
        for (int k = 0; k < 12; ++k)
            dest.A[i + k] = (k < 6) ? src.A[i + k] : -src.A[i + k];
}

void check_gauge_unitary(complex<double> *U)
{
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) {
            complex<double> s = 0;
            for (int c = 0; c < 3; ++c)
                s += U[a * 3 + c] * std::conj(U[b * 3 + c]);
            cout << s << " ";
        }
        cout << endl;
    }
}

void check_plaq(lattice_gauge &U)
{
    int *L = U.subgs;
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    double sum = 0.0;
    for (long lex = 0; lex < V; ++lex) {
        int x[4];
        index2site((int)lex, x, L);
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = mu + 1; nu < 4; ++nu) {
                int xm[4] = {x[0], x[1], x[2], x[3]}, xn[4] = {x[0], x[1], x[2], x[3]};
                xm[mu] = (xm[mu] + 1) % L[mu];
                xn[nu] = (xn[nu] + 1) % L[nu];
                const complex<double> *a = U.A[mu] + 9 * site2eoindex(x, L);
                const complex<double> *b = U.A[nu] + 9 * site2eoindex(xm, L);
                const complex<double> *c = U.A[mu] + 9 * site2eoindex(xn, L);
                const complex<double> *d = U.A[nu] + 9 * site2eoindex(x, L);
                // tr(a b c^dag d^dag)
                complex<double> ab[9], abc[9], tr = 0;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        ab[i * 3 + j] = 0;
                        for (int k = 0; k < 3; ++k)
                            ab[i * 3 + j] += a[i * 3 + k] * b[k * 3 + j];
                    }
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        abc[i * 3 + j] = 0;
                        for (int k = 0; k < 3; ++k)
                            abc[i * 3 + j] += ab[i * 3 + k] * std::conj(c[j * 3 + k]);
                    }
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k)
                        tr += abc[i * 3 + k] * std::conj(d[i * 3 + k]);
                sum += tr.real() / 3.0;
            }
    }
    cout << "plaquette = " << sum / (6.0 * V) << endl;
}
