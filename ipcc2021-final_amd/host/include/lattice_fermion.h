/**
 * lattice_fermion.h -- spinor field container with the reference's interface
 * (include/lattice_fermion.h:12-35): complex<double>[V*12] in the even-odd
 * order eo(x)*12 + spin*3 + colour.  Arithmetic runs on the host like the
 * reference's; the operators in dslash.h / dslash_new.h / invert.h move the
 * data to the GPU.
 */
#ifndef LATTICE_LATTICE_FERMION_H
#define LATTICE_LATTICE_FERMION_H

#include <complex>

class lattice_fermion {
  public:
    int *site_vec;
    std::complex<double> *A;
    int *subgs;
    int size; // complex entries: prod(subgs) * 12

    // non-owning view of an existing array (src/lattice_fermion.cpp:15-22)
    lattice_fermion(std::complex<double> *chroma_fermi, int *subgs1, int *site_vec1);
    // owning, 64-byte aligned, uninitialised (src/lattice_fermion.cpp:24-32)
    lattice_fermion(int *subgs1, int *site_vec1);
    ~lattice_fermion();
    // like the reference these mutate and return *this (src/lattice_fermion.cpp:51-65)
    lattice_fermion &operator-(const lattice_fermion &a);
    lattice_fermion &operator+(const lattice_fermion &a);
    void operator+=(const lattice_fermion &a);

    void clean();

  private:
    bool mem_flag;
};

void Minus(lattice_fermion &src1, lattice_fermion &src2, lattice_fermion &a);
void Plus(lattice_fermion &src1, lattice_fermion &src2, lattice_fermion &a);

#endif
