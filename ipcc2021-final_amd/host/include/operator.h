/**
 * operator.h -- lattice index helpers with the reference's semantics
 * (include/operator.h:11-75).
 */
#ifndef OPERATOR_H
#define OPERATOR_H

#include <complex>

static inline int local_site2(const int *coord, const int *latt_size)
{
    int order = 0;
    for (int mmu = 3; mmu >= 1; --mmu)
        order = latt_size[mmu - 1] * (coord[mmu] + order);
    return order + coord[0];
}

static inline int get_nodenum(const int *x, int *l, int nd)
{
    int n = 0;
    for (int i = nd - 1; i >= 0; i--)
        n = n * l[i] + x[i];
    return n;
}

// lex(x) = x0 + L0 (x1 + L1 (x2 + L2 x3)), coordinates taken modulo L
static inline int site2index(int *x, int *L)
{
    return x[0] % L[0] + L[0] * (x[1] % L[1] + L[1] * (x[2] % L[2] + L[2] * (x[3] % L[3])));
}

// eo(x) = lex(x)/2 + ((x0+x1+x2+x3) mod 2) * V/2
static inline int site2eoindex(int *x, int *L)
{
    const int Vh = L[0] * L[1] * L[2] * L[3] / 2;
    return site2index(x, L) / 2 + (x[0] + x[1] + x[2] + x[3]) % 2 * Vh;
}

static inline void index2site(int idx, int x[4], int *L)
{
    x[0] = idx % L[0];
    x[1] = (idx / L[0]) % L[1];
    x[2] = (idx / (L[0] * L[1])) % L[2];
    x[3] = (idx / (L[0] * L[1] * L[2])) % L[3];
}

// |v|^2 over the even and the odd half (include/operator.h:33-54, taken by
// reference here instead of by value)
template <typename T>
double norm_2_E(const T &s)
{
    double r = 0.0;
    for (int i = 0; i < s.size / 2; i++)
        r += std::norm(s.A[i]);
    return r;
}

template <typename T>
double norm_2_O(const T &s)
{
    double r = 0.0;
    for (int i = s.size / 2; i < s.size; i++)
        r += std::norm(s.A[i]);
    return r;
}

#endif
