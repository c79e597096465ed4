// wd_blas.h -- device reduction helpers and BLAS-1 launchers shared by the
// library's translation units.
#pragma once

#include "wd_internal.h"

// Sums (sr, si) over the workgroup (kBlock threads: wavefront shuffles, then
// LDS across wavefronts in a fixed order); thread 0 stores the result.
__device__ __forceinline__ void block_reduce_store(double sr, double si, double2 *dst)
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
    // lr/li are one array per kernel: a second call must not overwrite them
    // before thread 0 has read this call's values
    __syncthreads();
}

// Sum of `count` partials by one 1024-thread workgroup, fixed order.  The
// result is valid in thread 0.
__device__ __forceinline__ double2 sum_partials_1024(const double2 *__restrict__ partials,
                                                     int count)
{
    double sr = 0.0, si = 0.0;
    for (int k = threadIdx.x; k < count; k += 1024) {
        const double2 v = partials[k];
        sr += v.x;
        si += v.y;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    __shared__ double lr[16], li[16];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lr[w] = sr;
        li[w] = si;
    }
    __syncthreads();
    double2 out = make_double2(0.0, 0.0);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 16; ++k) {
            out.x += lr[k];
            out.y += li[k];
        }
    }
    __syncthreads(); // as in block_reduce_store: safe to call repeatedly
    return out;
}

int wd_launch_scale(wd_ctx *ctx, double2 *out, const double2 *in, double a, size_t n);
int wd_launch_sub(wd_ctx *ctx, double2 *out, const double2 *a, const double2 *b, size_t n);
// result (device double2) = sum conj(a) b, or sum |a|^2 when b == nullptr.
int wd_launch_dot(wd_ctx *ctx, const double2 *a, const double2 *b, size_t n, double2 *result);
int wd_red_grid(size_t n);
int wd_elem_grid(size_t n);
int wd_hop_grid(const wd_ctx *ctx, int nparity);
