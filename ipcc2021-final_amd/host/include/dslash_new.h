/**
 * dslash_new.h -- the reference's optimised operator entry points
 * (include/dslash_new.h:18-22), computed on the GPU through libwilson.so.
 *   DslashEEOONew          dest = (4+m) src over the full lattice
 *   DslashoffdNew(cb)      dest[cb] = hopping term from src[1-cb], dest[1-cb] = 0
 *   DslashEONew / OENew    DslashoffdNew with cb = 0 / 1 (declared but never
 *                          defined in the reference, include/dslash_new.h:19-20)
 */
#ifndef LATTICE_DSLASH_NEW_H
#define LATTICE_DSLASH_NEW_H

#include "lattice_fermion.h"
#include "lattice_gauge.h"
#include "operator.h"
#include "operator_mpi.h"

void DslashEEOONew(lattice_fermion &src, lattice_fermion &dest, const double mass);
void DslashEONew(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag);
void DslashOENew(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag);
void DslashoffdNew(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag,
                   int cb);
#endif
