/**
 * check.h -- the reference's post-solve report (include/check.h:13-24,
 * src/check.cpp:18-110): |b|, |x|, (Mb,Mb) vs (b,M^dag M b), (Mx,Mx) vs
 * (x,M^dag M x), |Mx - b| and |M^dag M x - M^dag b|.
 */
#ifndef _CHECK_H_
#define _CHECK_H_

#include <complex>

#include "lattice_fermion.h"
#include "lattice_gauge.h"

int check(std::complex<double> *src_p, std::complex<double> *chi_p, std::complex<double> *gauge[4],
          const double mass, int *subgs, int *site_vec);

int check(lattice_fermion &src, lattice_fermion &chi, lattice_gauge &U, double mass);

// prints the lattice and subgrid extents
void check_lattice_site(int *subgs, int *site_vec);

// dest = gamma5 src, gamma5 = diag(1, 1, -1, -1) in the DeGrand-Rossi basis
void gamma5_lattcefermion(lattice_fermion &src, lattice_fermion &dest);

// prints U U^dagger for one 3x3 link
void check_gauge_unitary(std::complex<double> *U);

// prints the average plaquette Re tr(U_mu U_nu U_mu^dag U_nu^dag) / 3
void check_plaq(lattice_gauge &U);

#endif
