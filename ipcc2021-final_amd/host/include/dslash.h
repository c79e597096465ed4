/**
 * dslash.h -- the reference's operator entry points (include/dslash.h:16-21),
 * computed on the GPU through libwilson.so.
 *   DslashEE / DslashOO   dest = (4+m) src on the even / odd half, other half 0
 *   Dslashoffd(cb)        dest[cb] = hopping term from src[1-cb], dest[1-cb] = 0
 *   DslashEO / DslashOE   Dslashoffd with cb = 0 / 1 (declared but never defined
 *                         in the reference, include/dslash.h:18-19)
 */
#ifndef LATTICE_DSLASH_H
#define LATTICE_DSLASH_H

#include "lattice_fermion.h"
#include "lattice_gauge.h"
#include "operator.h"
#include "operator_mpi.h"

void DslashEE(lattice_fermion &src, lattice_fermion &dest, const double mass);
void DslashOO(lattice_fermion &src, lattice_fermion &dest, const double mass);
void DslashEO(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag);
void DslashOE(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag);
void Dslashoffd(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag,
                int cb);
#endif
