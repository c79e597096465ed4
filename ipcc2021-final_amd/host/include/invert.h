/**
 * invert.h -- the reference's solver and full-operator entry points
 * (include/invert.h:15-26), on the GPU through libwilson.so.
 *   CGinvert   CGNR on M^dagger M x = M^dagger b; x0 drawn from rand() like the
 *              reference (src/invert.cpp:42-44), absolute stop |r| < accuracy,
 *              loop f = 1 .. max-1; prints the reference's "CG: ..." lines
 *              (per-iteration lines unless compiled with -DVERBOSE_SIMPLE).
 *              Returns 0 like the reference.
 *   Dslash / DslashNew   dest = M src or M^dagger src over the full lattice.
 * Errors of the GPU layer (no device, allocation failure, a partitioned
 * subgrid) throw std::runtime_error; the reference has no error reporting.
 */
#ifndef LATTICE_invert_H
#define LATTICE_invert_H

#include <complex>

#include "dslash.h"
#include "dslash_new.h"
#include "lattice_fermion.h"
#include "lattice_gauge.h"

int CGinvert(std::complex<double> *src_p, std::complex<double> *dest_p,
             std::complex<double> *gauge[4], const double mass, const int max,
             const double accuracy, int *subgs, int *site_vec);

int CGinvert(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const double mass,
             const int max, const double accuracy);

void Dslash(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const double mass,
            const bool dagger);

void DslashNew(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const double mass,
               const bool dagger);

#endif
