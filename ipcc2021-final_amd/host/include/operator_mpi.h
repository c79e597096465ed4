/**
 * operator_mpi.h -- global reductions with the reference's names
 * (include/operator_mpi.h:15-79).  One GPU holds the whole lattice, so the
 * MPI_Reduce / MPI_Bcast steps of the reference are identities and the sums
 * run on the host over the caller's arrays.
 */
#ifndef LATTICE_OPERATOR_MPI_H
#define LATTICE_OPERATOR_MPI_H

#include <complex>

// sum |s_i|^2
template <typename T>
double norm_2(const T &s)
{
    double r = 0.0;
    for (int i = 0; i < s.size; i++)
        r += s.A[i].real() * s.A[i].real() + s.A[i].imag() * s.A[i].imag();
    return r;
}

// sum conj(r1_i) r2_i
template <typename T>
std::complex<double> vector_p(const T &r1, const T &r2)
{
    double re = 0.0, im = 0.0;
    for (int i = 0; i < r1.size; i++) {
        const std::complex<double> a = r1.A[i], b = r2.A[i];
        re += a.real() * b.real() + a.imag() * b.imag();
        im += a.real() * b.imag() - a.imag() * b.real();
    }
    return std::complex<double>(re, im);
}

#endif
