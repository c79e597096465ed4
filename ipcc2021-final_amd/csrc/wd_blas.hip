// wd_blas.hip -- BLAS-1 and reduction kernels on FP64 fields.
//
// Fields are flat arrays of double2 (complex) of length n; padding sites are
// zero and stay zero under every operation here.  Reductions are two-stage and
// deterministic: kRedBlocks workgroups each reduce a fixed grid-stride slice
// (wavefront shuffle, then LDS across the 4 wavefronts) into one partial, and
// a single-workgroup kernel sums the partials in a fixed tree order.  Reruns
// are therefore bitwise reproducible.
#include "wd_blas.h"

namespace {

typedef double d2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ d2v ldv(const double2 *p, size_t i)
{
    return *reinterpret_cast<const d2v *>(p + i);
}
__device__ __forceinline__ void stv(double2 *p, size_t i, d2v v)
{
    *reinterpret_cast<d2v *>(p + i) = v;
}

__global__ __launch_bounds__(kBlock) void k_scale(double2 *__restrict__ out,
                                                  const double2 *__restrict__ in, double a,
                                                  size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBlock)
        stv(out, i, a * ldv(in, i));
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(uint4 *__restrict__ p, size_t n16)
{
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n16;
         i += (size_t)gridDim.x * kBlock)
        p[i] = make_uint4(0, 0, 0, 0);
}

__global__ __launch_bounds__(kBlock) void k_sub(double2 *__restrict__ out,
                                                const double2 *__restrict__ a,
                                                const double2 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBlock)
        stv(out, i, ldv(a, i) - ldv(b, i));
}

// partial of sum conj(a) b (CONJ_DOT) or sum |a|^2 (NORM)
template <bool NORM>
__global__ __launch_bounds__(kBlock) void k_dot_partial(const double2 *__restrict__ a,
                                                        const double2 *__restrict__ b, size_t n,
                                                        double2 *__restrict__ partials)
{
    double sr = 0.0, si = 0.0;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBlock) {
        const d2v x = ldv(a, i);
        if constexpr (NORM) {
            sr = fma(x.x, x.x, sr);
            sr = fma(x.y, x.y, sr);
        } else {
            const d2v y = ldv(b, i);
            sr = fma(x.x, y.x, sr);
            sr = fma(x.y, y.y, sr);
            si = fma(x.x, y.y, si);
            si = fma(-x.y, y.x, si);
        }
    }
    block_reduce_store(sr, si, partials + blockIdx.x);
}

__global__ __launch_bounds__(1024) void k_sum_partials(const double2 *__restrict__ partials,
                                                       int count, double2 *__restrict__ out)
{
    double2 s = sum_partials_1024(partials, count);
    if (threadIdx.x == 0)
        *out = s;
}

int red_grid(size_t n)
{
    const size_t want = (n + kBlock - 1) / kBlock;
    return (int)(want < (size_t)kRedBlocks ? want : (size_t)kRedBlocks);
}

int launch_grid(size_t n)
{
    const size_t want = (n + kBlock - 1) / kBlock;
    const size_t cap = 8192;
    return (int)(want < cap ? (want > 0 ? want : 1) : cap);
}

} // namespace

int wd_launch_zero(wd_ctx *ctx, void *p, size_t bytes)
{
    if (bytes == 0)
        return WD_OK;
    if (bytes % 16 != 0)
        return wd_set_error(ctx, WD_EARG, "zero: size not a multiple of 16 bytes");
    const size_t n16 = bytes / 16;
    hipLaunchKernelGGL(k_fill_zero, dim3(launch_grid(n16)), dim3(kBlock), 0, ctx->stream,
                       (uint4 *)p, n16);
    return wd_check_launch(ctx, "k_fill_zero");
}

int wd_launch_scale(wd_ctx *ctx, double2 *out, const double2 *in, double a, size_t n)
{
    hipLaunchKernelGGL(k_scale, dim3(launch_grid(n)), dim3(kBlock), 0, ctx->stream, out, in, a, n);
    return wd_check_launch(ctx, "k_scale");
}

int wd_launch_sub(wd_ctx *ctx, double2 *out, const double2 *a, const double2 *b, size_t n)
{
    hipLaunchKernelGGL(k_sub, dim3(launch_grid(n)), dim3(kBlock), 0, ctx->stream, out, a, b, n);
    return wd_check_launch(ctx, "k_sub");
}

int wd_launch_dot(wd_ctx *ctx, const double2 *a, const double2 *b, size_t n, double2 *result)
{
    const int g = red_grid(n);
    double2 *partials = (double2 *)ctx->partials;
    if (b == nullptr)
        hipLaunchKernelGGL(k_dot_partial<true>, dim3(g), dim3(kBlock), 0, ctx->stream, a, a, n,
                           partials);
    else
        hipLaunchKernelGGL(k_dot_partial<false>, dim3(g), dim3(kBlock), 0, ctx->stream, a, b, n,
                           partials);
    int rc = wd_check_launch(ctx, "k_dot_partial");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, ctx->stream, partials, g, result);
    return wd_check_launch(ctx, "k_sum_partials");
}

int wd_red_grid(size_t n) { return red_grid(n); }
int wd_elem_grid(size_t n) { return launch_grid(n); }
