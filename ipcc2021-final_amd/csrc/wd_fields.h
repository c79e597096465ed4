// wd_fields.h -- device access to every spinor and link storage format.
//
// Spinor fields, per right-hand side and parity, 16-byte chunks, site fastest:
//   F64  double2 in 64-site tiles [stride/64][12][64] (f64_idx)
//   F32  float4  [ 6][stride]   chunk = 2 complex (components 2j, 2j+1)
//   F16  uint4   [ 3][stride]   chunk = 4 complex as half2, values v / s with a
//        float scale s [stride] = max |real component| of the site
// Gauge fields, per direction and parity:
//   F64 R18 / R12  double2 in 64-site tiles [stride/64][9 or 6][64]
//   F32 R18 float2  [9][stride]   F32 R12 float4  [3][stride]
//   F16 R12 uint2   [3][stride]   (chunk = 2 complex as half2; |U_ab| <= 1
//        for unitary links, so no scale)
// 12-real links hold rows 0 and 1; row 2 = conj(row0 x row1) is rebuilt in
// registers (exact for SU(3); the 12-real storage of refs/utils_nersc_lat.cpp:60-68).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "wd_dirac.h"
#include "wd_internal.h"

namespace wdf {

typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

// FP64 fields (spinors and links) are tiled by 64 sites: component k of site i
// of a parity block sits at (i / 64) * (64 * NC) + k * 64 + i % 64 for NC
// components per site, so a wavefront's access to one component of 64
// consecutive sites is one 1 KiB run and all NC components of those sites are
// one contiguous NC KiB block (fewer DRAM pages per wavefront than NC
// separate arrays; measured +2 % of HBM peak for the FP64 hop on MI355X).
__host__ __device__ __forceinline__ size_t f64_idx(int k, unsigned i, int NC)
{
    return (size_t)(i >> 6) * (64u * NC) + (size_t)k * 64u + (i & 63u);
}

template <bool NT, typename V>
__device__ __forceinline__ V ldg(const V *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <bool NT, typename V>
__device__ __forceinline__ void stg(V *p, V v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// ---- half helpers ------------------------------------------------------------
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

// Real and imaginary part of a complex packed as half2, widened.  Used as
// operands of fmaf, these become the f16 sources of v_fma_mix_f32 (op_sel
// picks the half), so no separate conversion instructions are issued.
__device__ __forceinline__ float re16(unsigned w) { return (float)__builtin_bit_cast(h2v, w).x; }
__device__ __forceinline__ float im16(unsigned w) { return (float)__builtin_bit_cast(h2v, w).y; }

// |v|^2 of an F16 site from its packed words and scale: the f16 squares are
// exact in FP32, the scale is applied once in FP64.
__device__ __forceinline__ double norm2_packed(const unsigned (&w)[12], float s)
{
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        t = fmaf(re16(w[k]), re16(w[k]), t);
        t = fmaf(im16(w[k]), im16(w[k]), t);
    }
    return (double)t * ((double)s * (double)s);
}
__device__ __forceinline__ cpx<float> h2f(unsigned w, float s)
{
    __half2 h = *reinterpret_cast<__half2 *>(&w);
    const float2 f = __half22float2(h);
    return {f.x * s, f.y * s};
}

__device__ __forceinline__ unsigned f2h(float re, float im)
{
    __half2 h = __floats2half2_rn(re, im);
    return *reinterpret_cast<unsigned *>(&h);
}

// f2h(re * s, im * s) with the products rounded to FP32 first, in every
// kernel: the compiler may otherwise fuse a product and its F16 conversion
// into one mixed-precision FMA (one rounding instead of two) in some kernels
// and not in others, and two kernels packing the same values would disagree.
__device__ __forceinline__ unsigned f2h_scaled(float re, float im, float s)
{
    float a = re * s, b = im * s;
    asm volatile("" : "+v"(a), "+v"(b));
    return f2h(a, b);
}

// ---- spinor sites ---------------------------------------------------------------
// base points at the parity block of one rhs; scale at that parity's scales (F16).
template <int FMT>
struct Spinor;

template <>
struct Spinor<WD_F64> {
    using T = double;
    template <typename R, bool NT = false>
    __device__ static __forceinline__ void load(const void *base, const float *, int stride,
                                                unsigned i, cpx<R> (&v)[4][3])
    {
        const double2 *b = static_cast<const double2 *>(base);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const d2v c = ldg<NT>(reinterpret_cast<const d2v *>(b + f64_idx(k, i, 12)));
            v[k / 3][k % 3] = {(R)c.x, (R)c.y};
        }
    }
    template <typename R, bool NT = false>
    __device__ static __forceinline__ void store(void *base, float *, int stride, unsigned i,
                                                 const cpx<R> (&v)[4][3])
    {
        double2 *b = static_cast<double2 *>(base);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            d2v c;
            c.x = (double)v[k / 3][k % 3].r;
            c.y = (double)v[k / 3][k % 3].i;
            stg<NT>(reinterpret_cast<d2v *>(b + f64_idx(k, i, 12)), c);
        }
    }
};

template <>
struct Spinor<WD_F32> {
    using T = float;
    template <typename R, bool NT = false>
    __device__ static __forceinline__ void load(const void *base, const float *, int stride,
                                                unsigned i, cpx<R> (&v)[4][3])
    {
        const float4 *b = static_cast<const float4 *>(base);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const f4v c = ldg<NT>(reinterpret_cast<const f4v *>(b + (size_t)j * (unsigned)stride + i));
            v[(2 * j) / 3][(2 * j) % 3] = {(R)c.x, (R)c.y};
            v[(2 * j + 1) / 3][(2 * j + 1) % 3] = {(R)c.z, (R)c.w};
        }
    }
    template <typename R, bool NT = false>
    __device__ static __forceinline__ void store(void *base, float *, int stride, unsigned i,
                                                 const cpx<R> (&v)[4][3])
    {
        float4 *b = static_cast<float4 *>(base);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            f4v c;
            c.x = (float)v[(2 * j) / 3][(2 * j) % 3].r;
            c.y = (float)v[(2 * j) / 3][(2 * j) % 3].i;
            c.z = (float)v[(2 * j + 1) / 3][(2 * j + 1) % 3].r;
            c.w = (float)v[(2 * j + 1) / 3][(2 * j + 1) % 3].i;
            stg<NT>(reinterpret_cast<f4v *>(b + (size_t)j * (unsigned)stride + i), c);
        }
    }
};

template <>
struct Spinor<WD_F16> {
    using T = float;
    template <typename R, bool NT = false>
    __device__ static __forceinline__ void load(const void *base, const float *scale, int stride,
                                                unsigned i, cpx<R> (&v)[4][3])
    {
        const uint4 *b = static_cast<const uint4 *>(base);
        const float s = ldg<NT>(scale + i);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const u4v c = ldg<NT>(reinterpret_cast<const u4v *>(b + (size_t)j * (unsigned)stride + i));
            const unsigned w[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const cpx<float> f = h2f(w[q], s);
                v[(4 * j + q) / 3][(4 * j + q) % 3] = {(R)f.r, (R)f.i};
            }
        }
    }
    template <typename R, bool NT = false>
    __device__ static __forceinline__ void store(void *base, float *scale, int stride, unsigned i,
                                                 const cpx<R> (&v)[4][3])
    {
        unsigned w[12];
        (void)store_packed<R, NT>(base, scale, stride, i, v, w);
    }
    // As store; also returns the packed words and the scale as stored, so a
    // caller can reduce over the rounded values without reading them back.
    template <typename R, bool NT = false>
    __device__ static __forceinline__ float store_packed(void *base, float *scale, int stride,
                                                         unsigned i, const cpx<R> (&v)[4][3],
                                                         unsigned (&wout)[12])
    {
        uint4 *b = static_cast<uint4 *>(base);
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < 12; ++k)
            m = fmaxf(m, fmaxf(fabsf((float)v[k / 3][k % 3].r), fabsf((float)v[k / 3][k % 3].i)));
        // a site whose largest component is below 1e-30 is stored as zero: 1/m
        // would overflow for denormal m (far from a point source the fields
        // decay to ~1e-38), turning the zeros into NaN.  NaN inputs still
        // propagate, so breakdowns stay visible.
        const bool tiny = m < 1.0e-30f;
        const float inv = tiny ? 0.f : 1.0f / m;
        const float sc = tiny ? 0.f : m;
        stg<NT>(scale + i, sc);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const cpx<R> z = v[(4 * j + q) / 3][(4 * j + q) % 3];
                wout[4 * j + q] = f2h_scaled((float)z.r, (float)z.i, inv);
            }
            u4v c;
            c.x = wout[4 * j];
            c.y = wout[4 * j + 1];
            c.z = wout[4 * j + 2];
            c.w = wout[4 * j + 3];
            stg<NT>(reinterpret_cast<u4v *>(b + (size_t)j * (unsigned)stride + i), c);
        }
        return sc;
    }
};

// chunks per parity block and bytes per chunk of a spinor format
__host__ __device__ constexpr int spinor_chunks(int fmt) { return fmt == WD_F64 ? 12 : fmt == WD_F32 ? 6 : 3; }

// ---- links ----------------------------------------------------------------------
// GF: 0 = F64 R18, 1 = F64 R12, 2 = F32 R18, 3 = F32 R12, 4 = F16 R12
enum GaugeFmt { G_F64_18 = 0, G_F64_12 = 1, G_F32_18 = 2, G_F32_12 = 3, G_F16_12 = 4 };

__host__ __device__ constexpr int gauge_fmt(int prec, int recon)
{
    return prec == WD_F64 ? (recon == WD_RECON_18 ? G_F64_18 : G_F64_12)
         : prec == WD_F32 ? (recon == WD_RECON_18 ? G_F32_18 : G_F32_12)
                          : G_F16_12;
}

// chunks per (mu, parity) block
__host__ __device__ constexpr int gauge_chunks(int gf)
{
    return gf == G_F64_18 ? 9 : gf == G_F64_12 ? 6 : gf == G_F32_18 ? 9 : 3;
}
__host__ __device__ constexpr int gauge_chunk_bytes(int gf)
{
    return gf == G_F64_18 || gf == G_F64_12 ? 16 : gf == G_F32_18 ? 8 : gf == G_F32_12 ? 16 : 8;
}

// Loads one link (rebuilding row 2 for 12-real formats).  NT selects
// non-temporal loads: right when every link is read once per launch, wrong
// when several right-hand-side groups re-read it from L2.
template <int GF, bool NT, typename R>
__device__ __forceinline__ void load_link(const void *base, int stride, unsigned j,
                                          cpx<R> (&U)[3][3])
{
    if constexpr (GF == G_F64_18 || GF == G_F64_12) {
        const double2 *b = static_cast<const double2 *>(base);
        constexpr int n = GF == G_F64_18 ? 9 : 6;
#pragma unroll
        for (int k = 0; k < n; ++k) {
            const d2v c = ldg<NT>(reinterpret_cast<const d2v *>(b + f64_idx(k, j, n)));
            U[k / 3][k % 3] = {(R)c.x, (R)c.y};
        }
        if constexpr (GF == G_F64_12)
            reconstruct_row2(U);
    } else if constexpr (GF == G_F32_18) {
        const float2 *b = static_cast<const float2 *>(base);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const f2v c = ldg<NT>(reinterpret_cast<const f2v *>(b + (size_t)k * (unsigned)stride + j));
            U[k / 3][k % 3] = {(R)c.x, (R)c.y};
        }
    } else if constexpr (GF == G_F32_12) {
        const float4 *b = static_cast<const float4 *>(base);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const f4v c = ldg<NT>(reinterpret_cast<const f4v *>(b + (size_t)q * (unsigned)stride + j));
            U[(2 * q) / 3][(2 * q) % 3] = {(R)c.x, (R)c.y};
            U[(2 * q + 1) / 3][(2 * q + 1) % 3] = {(R)c.z, (R)c.w};
        }
        reconstruct_row2(U);
    } else {
        const uint2 *b = static_cast<const uint2 *>(base);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const u2v c = ldg<NT>(reinterpret_cast<const u2v *>(b + (size_t)q * (unsigned)stride + j));
            const cpx<float> a = h2f(c.x, 1.0f), d = h2f(c.y, 1.0f);
            U[(2 * q) / 3][(2 * q) % 3] = {(R)a.r, (R)a.i};
            U[(2 * q + 1) / 3][(2 * q + 1) % 3] = {(R)d.r, (R)d.i};
        }
        reconstruct_row2(U);
    }
}

} // namespace wdf
