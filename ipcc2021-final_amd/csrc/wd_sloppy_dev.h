// wd_sloppy_dev.h -- device helpers shared by the FP32-arithmetic sloppy
// operators (wd_hop_sloppy.hip: R systems per thread; wd_hop_sloppy_loop.hip:
// every system of a site in one thread): 32-bit lane offsets into the
// [chunks][stride] spinor arrays, F16 site loads and stores with the per-site
// scale, FP32 link loads (row 2 rebuilt for 12-real formats), the projection
// of an F16 neighbour into a half spinor, the reconstruction of a direction's
// half spinor into the 4-spinor accumulator, and the workgroup reduction of
// the per-system partials.
#pragma once

#include "wd_fields.h"
#include "wd_internal.h"

namespace wdsl {

using namespace wdf;

// Byte offset of 16-byte chunk q of site j in a [chunks][stride] array; the
// launcher checks that every field fits 32-bit offsets, so loads use a
// uniform 64-bit base plus a 32-bit lane offset.
__device__ __forceinline__ unsigned off16(int q, int stride, unsigned j)
{
    return ((unsigned)q * (unsigned)stride + j) * 16u;
}

// The 12 packed complex of an F16 site (component k = 3 * spin + colour).
__device__ __forceinline__ void load16(const char *base, int stride, unsigned j, unsigned (&w)[12])
{
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const u4v c = *reinterpret_cast<const u4v *>(base + off16(q, stride, j));
        w[4 * q] = c.x;
        w[4 * q + 1] = c.y;
        w[4 * q + 2] = c.z;
        w[4 * q + 3] = c.w;
    }
}

__device__ __forceinline__ void load32(const char *base, int stride, unsigned j, cpx<float> (&v)[4][3])
{
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const f4v c = *reinterpret_cast<const f4v *>(base + off16(q, stride, j));
        v[(2 * q) / 3][(2 * q) % 3] = {c.x, c.y};
        v[(2 * q + 1) / 3][(2 * q + 1) % 3] = {c.z, c.w};
    }
}

// a + mulc<P>(b) from two packed complex, one v_fma_mix_f32 per real.  `one`
// is 1.0f passed in from the host: with a literal 1 the compiler would fold
// the fma into widen + add (three instructions instead of one).
template <int P>
__device__ __forceinline__ cpx<float> proj16(unsigned wa, unsigned wb, float one)
{
    if constexpr (P == 0)
        return {fmaf(re16(wb), one, re16(wa)), fmaf(im16(wb), one, im16(wa))};
    else if constexpr (P == 1)
        return {fmaf(re16(wb), -one, re16(wa)), fmaf(im16(wb), -one, im16(wa))};
    else if constexpr (P == 2) // +i b = (-b.i, b.r)
        return {fmaf(im16(wb), -one, re16(wa)), fmaf(re16(wb), one, im16(wa))};
    else // -i b = (b.i, -b.r)
        return {fmaf(im16(wb), one, re16(wa)), fmaf(re16(wb), -one, im16(wa))};
}

// One link with 32-bit lane offsets (row 2 rebuilt for 12-real formats).  NT:
// non-temporal loads, right when every link is read once per launch.
template <int GF, bool NT>
__device__ __forceinline__ void load_link32(const char *base, int stride, unsigned j,
                                            cpx<float> (&U)[3][3])
{
    if constexpr (GF == G_F32_18) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const f2v c = ldg<NT>(reinterpret_cast<const f2v *>(
                base + ((unsigned)k * (unsigned)stride + j) * 8u));
            U[k / 3][k % 3] = {c.x, c.y};
        }
    } else if constexpr (GF == G_F32_12) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const f4v c = ldg<NT>(reinterpret_cast<const f4v *>(base + off16(q, stride, j)));
            U[(2 * q) / 3][(2 * q) % 3] = {c.x, c.y};
            U[(2 * q + 1) / 3][(2 * q + 1) % 3] = {c.z, c.w};
        }
        reconstruct_row2(U);
    } else {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const u2v c = ldg<NT>(reinterpret_cast<const u2v *>(
                base + ((unsigned)q * (unsigned)stride + j) * 8u));
            U[(2 * q) / 3][(2 * q) % 3] = {re16(c.x), im16(c.x)};
            U[(2 * q + 1) / 3][(2 * q + 1) % 3] = {re16(c.y), im16(c.y)};
        }
        reconstruct_row2(U);
    }
}

__device__ __forceinline__ float load_scale(const float *base, unsigned j)
{
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + j * 4u);
}

// acc += f * mulc<C>(z)
template <int C>
__device__ __forceinline__ void acc_mulc(cpx<float> &acc, cpx<float> z, float f)
{
    if constexpr (C == 0) {
        acc.r = fmaf(f, z.r, acc.r);
        acc.i = fmaf(f, z.i, acc.i);
    } else if constexpr (C == 1) {
        acc.r = fmaf(-f, z.r, acc.r);
        acc.i = fmaf(-f, z.i, acc.i);
    } else if constexpr (C == 2) {
        acc.r = fmaf(-f, z.i, acc.r);
        acc.i = fmaf(f, z.r, acc.i);
    } else {
        acc.r = fmaf(f, z.i, acc.r);
        acc.i = fmaf(-f, z.r, acc.i);
    }
}

// acc += f * reconstruct(r) for direction D; f carries -1/2, the boundary
// phase and the F16 site scale, so the reconstruction is 24 FMAs.
template <class D>
__device__ __forceinline__ void reconstruct_f(cpx<float> (&acc)[4][3], const cpx<float> (&r)[2][3],
                                              float f)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        acc_mulc<0>(acc[0][c], r[0][c], f);
        acc_mulc<0>(acc[1][c], r[1][c], f);
        acc_mulc<D::R0>(acc[D::S0][c], r[0][c], f);
        acc_mulc<D::R1>(acc[D::S1][c], r[1][c], f);
    }
}

// Block-wide sum of (sr, si) stored by thread 0 to dst; safe to call repeatedly.
__device__ __forceinline__ void block_sum_store(double sr, double si, double2 *dst)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    __shared__ double lr[kBlock / 64], li[kBlock / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lr[w] = sr;
        li[w] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, ti = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) {
            tr += lr[k];
            ti += li[k];
        }
        *dst = make_double2(tr, ti);
    }
    __syncthreads();
}

// The F16 form of a site: words = values / max |component| (packed halves),
// returns the scale (that max).  Below 1e-30 the site is stored as zero (1/m
// would overflow for denormal m, turning zeros into NaN); NaN inputs still
// propagate.
__device__ __forceinline__ float pack16(const cpx<float> (&v)[4][3], unsigned (&w)[12])
{
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 12; ++k)
        m = fmaxf(m, fmaxf(fabsf(v[k / 3][k % 3].r), fabsf(v[k / 3][k % 3].i)));
    const bool tiny = m < 1.0e-30f;
    const float inv = tiny ? 0.f : 1.0f / m;
#pragma unroll
    for (int k = 0; k < 12; ++k)
        w[k] = f2h_scaled(v[k / 3][k % 3].r, v[k / 3][k % 3].i, inv);
    return tiny ? 0.f : m;
}

// Stores the F16 words and scale of site i.  Outputs are streamed
// (non-temporal) so they do not evict from L2 the input spinors that
// neighbouring workgroups still read.
__device__ __forceinline__ void store16w(char *base, float *scale, int stride, unsigned i,
                                         const unsigned (&w)[12], float s)
{
    __builtin_nontemporal_store(s, reinterpret_cast<float *>(reinterpret_cast<char *>(scale) + i * 4u));
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        u4v c;
        c.x = w[4 * q];
        c.y = w[4 * q + 1];
        c.z = w[4 * q + 2];
        c.w = w[4 * q + 3];
        __builtin_nontemporal_store(c, reinterpret_cast<u4v *>(base + off16(q, stride, i)));
    }
}

// Stores an F16 site and returns the packed words and the scale as stored.
__device__ __forceinline__ float store16(char *base, float *scale, int stride, unsigned i,
                                         const cpx<float> (&v)[4][3], unsigned (&w)[12])
{
    const float s = pack16(v, w);
    store16w(base, scale, stride, i, w, s);
    return s;
}

__device__ __forceinline__ void store32(char *base, int stride, unsigned i, const cpx<float> (&v)[4][3])
{
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        f4v c;
        c.x = v[(2 * q) / 3][(2 * q) % 3].r;
        c.y = v[(2 * q) / 3][(2 * q) % 3].i;
        c.z = v[(2 * q + 1) / 3][(2 * q + 1) % 3].r;
        c.w = v[(2 * q + 1) / 3][(2 * q + 1) % 3].i;
        __builtin_nontemporal_store(c, reinterpret_cast<f4v *>(base + off16(q, stride, i)));
    }
}

} // namespace wdsl
