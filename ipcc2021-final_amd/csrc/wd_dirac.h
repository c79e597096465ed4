// wd_dirac.h -- device building blocks of the Wilson hopping term.
//
// The spin projectors (1 -/+ f gamma_mu) in the DeGrand-Rossi basis act on
// two spin pairs per direction.  With h0 = psi0 + P0 psi_S0 and
// h1 = psi1 + P1 psi_S1 (P in {+1, -1, +i, -i}) the full 4-spinor
// (1 -/+ f gamma_mu) psi equals (h0, h1) on spins 0,1 and conj(P) h on spins
// S0, S1, so only the half spinor is multiplied by the link and the result is
// reconstructed in registers.  The codes per direction are those of the
// reference's scalar loops (src/dslash.cpp:541-553 x, :632-642 y, :725-735 z,
// :815-825 t for the forward hops, and :588-600, :678-688, :770-780, :862-872
// for the backward hops): backward = forward with P negated, dagger
// (f = -1) negates P once more.
#pragma once

#include <hip/hip_runtime.h>

template <typename T>
struct cpx {
    T r, i;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b)
{
    return {a.r + b.r, a.i + b.i};
}

// acc + a * b
template <typename T>
__device__ __forceinline__ cpx<T> cfma(cpx<T> acc, cpx<T> a, cpx<T> b)
{
    acc.r = fma(a.r, b.r, acc.r);
    acc.r = fma(-a.i, b.i, acc.r);
    acc.i = fma(a.r, b.i, acc.i);
    acc.i = fma(a.i, b.r, acc.i);
    return acc;
}

// acc + conj(a) * b
template <typename T>
__device__ __forceinline__ cpx<T> cfma_conj(cpx<T> acc, cpx<T> a, cpx<T> b)
{
    acc.r = fma(a.r, b.r, acc.r);
    acc.r = fma(a.i, b.i, acc.r);
    acc.i = fma(a.r, b.i, acc.i);
    acc.i = fma(-a.i, b.r, acc.i);
    return acc;
}

// Multiplication by a unit phase code: 0 -> +1, 1 -> -1, 2 -> +i, 3 -> -i.
template <int C, typename T>
__device__ __forceinline__ cpx<T> mulc(cpx<T> z)
{
    if constexpr (C == 0)
        return z;
    else if constexpr (C == 1)
        return {-z.r, -z.i};
    else if constexpr (C == 2)
        return {-z.i, z.r};
    else
        return {z.i, -z.r};
}

constexpr int conj_code(int c) { return c < 2 ? c : (c ^ 1); }

// Forward-hop projection pairs for f = +1: (S0, P0), (S1, P1).
template <int MU>
struct Proj;
template <>
struct Proj<0> { // x: h0 = psi0 - i psi3, h1 = psi1 - i psi2
    static constexpr int S0 = 3, P0 = 3, S1 = 2, P1 = 3;
};
template <>
struct Proj<1> { // y: h0 = psi0 + psi3, h1 = psi1 - psi2
    static constexpr int S0 = 3, P0 = 0, S1 = 2, P1 = 1;
};
template <>
struct Proj<2> { // z: h0 = psi0 - i psi2, h1 = psi1 + i psi3
    static constexpr int S0 = 2, P0 = 3, S1 = 3, P1 = 2;
};
template <>
struct Proj<3> { // t: h0 = psi0 - psi2, h1 = psi1 - psi3
    static constexpr int S0 = 2, P0 = 1, S1 = 3, P1 = 1;
};

template <int MU, bool FWD, bool DAG>
struct Dir {
    static constexpr int S0 = Proj<MU>::S0;
    static constexpr int S1 = Proj<MU>::S1;
    static constexpr int P0 = Proj<MU>::P0 ^ (FWD ? 0 : 1) ^ (DAG ? 1 : 0);
    static constexpr int P1 = Proj<MU>::P1 ^ (FWD ? 0 : 1) ^ (DAG ? 1 : 0);
    static constexpr int R0 = conj_code(P0);
    static constexpr int R1 = conj_code(P1);
};

// Projects psi (4 spins x 3 colours) to the half spinor of direction D.
template <class D, typename T>
__device__ __forceinline__ void project(const cpx<T> (&psi)[4][3], cpx<T> (&h)[2][3])
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        h[0][c] = cadd(psi[0][c], mulc<D::P0>(psi[D::S0][c]));
        h[1][c] = cadd(psi[1][c], mulc<D::P1>(psi[D::S1][c]));
    }
}

// r = W h where W = U (ADJ false) or U^dagger (ADJ true); U row-major 3x3.
template <bool ADJ, typename T>
__device__ __forceinline__ void su3_mul_half(const cpx<T> (&U)[3][3], const cpx<T> (&h)[2][3],
                                             cpx<T> (&r)[2][3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            cpx<T> acc = {T(0), T(0)};
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                if constexpr (ADJ)
                    acc = cfma_conj(acc, U[b][a], h[s][b]);
                else
                    acc = cfma(acc, U[a][b], h[s][b]);
            }
            r[s][a] = acc;
        }
    }
}

// acc += reconstruct(r) for direction D, scaled by the boundary sign.
template <class D, typename T>
__device__ __forceinline__ void reconstruct(cpx<T> (&acc)[4][3], const cpx<T> (&r)[2][3], T sgn)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const cpx<T> r0 = {sgn * r[0][a].r, sgn * r[0][a].i};
        const cpx<T> r1 = {sgn * r[1][a].r, sgn * r[1][a].i};
        acc[0][a] = cadd(acc[0][a], r0);
        acc[1][a] = cadd(acc[1][a], r1);
        acc[D::S0][a] = cadd(acc[D::S0][a], mulc<D::R0>(r0));
        acc[D::S1][a] = cadd(acc[D::S1][a], mulc<D::R1>(r1));
    }
}

// Third row of an SU(3) matrix from the first two: row2 = conj(row0 x row1).
template <typename T>
__device__ __forceinline__ void reconstruct_row2(cpx<T> (&U)[3][3])
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int a = (c + 1) % 3, b = (c + 2) % 3;
        // (row0 x row1)_c = row0_a row1_b - row0_b row1_a, each complex product
        // with explicit FMAs: the rounding is then the same in every kernel
        // that rebuilds row 2, whatever products the compiler would contract
        const cpx<T> p = U[0][a], q = U[1][b], s = U[0][b], t = U[1][a];
        const T xr = fma(p.r, q.r, -(p.i * q.i)), xi = fma(p.r, q.i, p.i * q.r);
        const T yr = fma(s.r, t.r, -(s.i * t.i)), yi = fma(s.r, t.i, s.i * t.r);
        U[2][c] = {xr - yr, yi - xi};
    }
}
