// lattice_fermion.cpp -- host container (src/lattice_fermion.cpp:15-162).
#include "lattice_fermion.h"

#include <cstdlib>
#include <new>

using std::complex;

lattice_fermion::lattice_fermion(complex<double> *chroma_fermi, int *subgs1, int *site_vec1)
    : site_vec(site_vec1), A(chroma_fermi), subgs(subgs1), mem_flag(false)
{
    size = subgs[0] * subgs[1] * subgs[2] * subgs[3] * 3 * 4;
}

lattice_fermion::lattice_fermion(int *subgs1, int *site_vec1)
    : site_vec(site_vec1), subgs(subgs1), mem_flag(true)
{
    size = subgs[0] * subgs[1] * subgs[2] * subgs[3] * 3 * 4;
    const size_t bytes = (sizeof(complex<double>) * (size_t)size + 63) / 64 * 64;
    A = static_cast<complex<double> *>(std::aligned_alloc(64, bytes));
    if (!A)
        throw std::bad_alloc();
}

lattice_fermion::~lattice_fermion()
{
    if (mem_flag)
        std::free(A);
    A = nullptr;
    subgs = nullptr;
    site_vec = nullptr;
}

void lattice_fermion::clean()
{
    for (int i = 0; i < size; i++)
        A[i] = 0;
}

lattice_fermion &lattice_fermion::operator-(const lattice_fermion &a)
{
    for (int i = 0; i < size; i++)
        A[i] -= a.A[i];
    return *this;
}

lattice_fermion &lattice_fermion::operator+(const lattice_fermion &a)
{
    for (int i = 0; i < size; i++)
        A[i] += a.A[i];
    return *this;
}

void lattice_fermion::operator+=(const lattice_fermion &a)
{
    for (int i = 0; i < size; i++)
        A[i] += a.A[i];
}

void Minus(lattice_fermion &src1, lattice_fermion &src2, lattice_fermion &a)
{
    for (int i = 0; i < src1.size; i++)
        a.A[i] = src1.A[i] - src2.A[i];
}

void Plus(lattice_fermion &src1, lattice_fermion &src2, lattice_fermion &a)
{
    for (int i = 0; i < src1.size; i++)
        a.A[i] = src1.A[i] + src2.A[i];
}
