// wd_convert.hip -- layout and precision conversion between the reference's
// AoS even-odd arrays and the device formats, and between device formats.
#include <hip/hip_runtime.h>

#include "wd_fields.h"
#include "wd_internal.h"

using namespace wdf;

namespace {

// reference spinor complex<double>[V*12] (eo*12 + k) -> device format FMT
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_spinor_from_ref(char *__restrict__ dst,
                                                            float *__restrict__ scale,
                                                            const double2 *__restrict__ src,
                                                            int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    cpx<double> v[4][3];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const double2 c = src[(size_t)e * 12 + k];
        v[k / 3][k % 3] = {c.x, c.y};
    }
    const size_t pbytes = (size_t)spinor_chunks(FMT) * 16 * stride;
    Spinor<FMT>::store(dst + p * pbytes, scale ? scale + (size_t)p * stride : nullptr, stride,
                       (unsigned)i, v);
}

template <int FMT>
__global__ __launch_bounds__(kBlock) void k_spinor_to_ref(double2 *__restrict__ dst,
                                                          const char *__restrict__ src,
                                                          const float *__restrict__ scale,
                                                          int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    cpx<double> v[4][3];
    const size_t pbytes = (size_t)spinor_chunks(FMT) * 16 * stride;
    Spinor<FMT>::load(src + p * pbytes, scale ? scale + (size_t)p * stride : nullptr, stride,
                      (unsigned)i, v);
#pragma unroll
    for (int k = 0; k < 12; ++k)
        dst[(size_t)e * 12 + k] = make_double2(v[k / 3][k % 3].r, v[k / 3][k % 3].i);
}

// device format FI -> FO over both parities of one rhs (padding sites included)
template <int FI, int FO>
__global__ __launch_bounds__(kBlock) void k_spinor_convert(char *__restrict__ dst,
                                                           float *__restrict__ dscale,
                                                           const char *__restrict__ src,
                                                           const float *__restrict__ sscale,
                                                           int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * stride)
        return;
    const int p = e / stride, i = e - p * stride;
    cpx<double> v[4][3];
    Spinor<FI>::load(src + (size_t)p * spinor_chunks(FI) * 16 * stride,
                     sscale ? sscale + (size_t)p * stride : nullptr, stride, (unsigned)i, v);
    Spinor<FO>::store(dst + (size_t)p * spinor_chunks(FO) * 16 * stride,
                      dscale ? dscale + (size_t)p * stride : nullptr, stride, (unsigned)i, v);
}

// reference links A[mu] complex<double>[V*9] -> device gauge format GF
template <int GF>
__global__ __launch_bounds__(kBlock) void k_gauge_from_ref(char *__restrict__ dst,
                                                           const double2 *__restrict__ src,
                                                           int mu, int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    const size_t block = (size_t)gauge_chunks(GF) * gauge_chunk_bytes(GF) * stride;
    char *base = dst + (size_t)(mu * 2 + p) * block;
    const double2 *s = src + (size_t)e * 9;
    if constexpr (GF == G_F64_18 || GF == G_F64_12) {
        double2 *b = reinterpret_cast<double2 *>(base);
        for (int k = 0; k < gauge_chunks(GF); ++k)
            b[f64_idx(k, (unsigned)i, gauge_chunks(GF))] = s[k];
    } else if constexpr (GF == G_F32_18) {
        float2 *b = reinterpret_cast<float2 *>(base);
        for (int k = 0; k < 9; ++k)
            b[(size_t)k * stride + i] = make_float2((float)s[k].x, (float)s[k].y);
    } else if constexpr (GF == G_F32_12) {
        float4 *b = reinterpret_cast<float4 *>(base);
        for (int q = 0; q < 3; ++q)
            b[(size_t)q * stride + i] = make_float4((float)s[2 * q].x, (float)s[2 * q].y,
                                                    (float)s[2 * q + 1].x, (float)s[2 * q + 1].y);
    } else {
        uint2 *b = reinterpret_cast<uint2 *>(base);
        for (int q = 0; q < 3; ++q)
            b[(size_t)q * stride + i] =
                make_uint2(f2h((float)s[2 * q].x, (float)s[2 * q].y),
                           f2h((float)s[2 * q + 1].x, (float)s[2 * q + 1].y));
    }
}

template <int FMT>
void launch_from_ref(wd_ctx *ctx, void *dst, float *scale, const double2 *src)
{
    const int grid = (int)((2 * ctx->Vh + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_spinor_from_ref<FMT>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                       (char *)dst, scale, src, (int)ctx->Vh, (int)ctx->stride);
}

template <int FMT>
void launch_to_ref(wd_ctx *ctx, double2 *dst, const void *src, const float *scale)
{
    const int grid = (int)((2 * ctx->Vh + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_spinor_to_ref<FMT>, dim3(grid), dim3(kBlock), 0, ctx->stream, dst,
                       (const char *)src, scale, (int)ctx->Vh, (int)ctx->stride);
}

template <int FI, int FO>
void launch_convert(wd_ctx *ctx, void *dst, float *dscale, const void *src, const float *sscale)
{
    const int grid = (int)((2 * ctx->stride + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((k_spinor_convert<FI, FO>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                       (char *)dst, dscale, (const char *)src, sscale, (int)ctx->stride);
}

} // namespace

void *wd_rhs_ptr(const wd_field *f, int rhs)
{
    return (char *)f->data + (size_t)rhs * f->rhs_elems * prec_bytes_complex(f->prec);
}

float *wd_rhs_scale(const wd_field *f, int rhs)
{
    return f->scale ? f->scale + (size_t)rhs * 2 * f->ctx->stride : nullptr;
}

int wd_launch_spinor_from_ref(wd_ctx *ctx, wd_field *f, int rhs, const double2 *src)
{
    void *d = wd_rhs_ptr(f, rhs);
    float *s = wd_rhs_scale(f, rhs);
    if (f->prec == WD_F64)
        launch_from_ref<WD_F64>(ctx, d, s, src);
    else if (f->prec == WD_F32)
        launch_from_ref<WD_F32>(ctx, d, s, src);
    else
        launch_from_ref<WD_F16>(ctx, d, s, src);
    return wd_check_launch(ctx, "k_spinor_from_ref");
}

int wd_launch_spinor_to_ref(wd_ctx *ctx, const wd_field *f, int rhs, double2 *dst)
{
    const void *d = wd_rhs_ptr(f, rhs);
    const float *s = wd_rhs_scale(f, rhs);
    if (f->prec == WD_F64)
        launch_to_ref<WD_F64>(ctx, dst, d, s);
    else if (f->prec == WD_F32)
        launch_to_ref<WD_F32>(ctx, dst, d, s);
    else
        launch_to_ref<WD_F16>(ctx, dst, d, s);
    return wd_check_launch(ctx, "k_spinor_to_ref");
}

int wd_launch_gauge_from_ref(wd_ctx *ctx, wd_gauge *U, int mu, const double2 *src)
{
    const int grid = (int)((2 * ctx->Vh + kBlock - 1) / kBlock);
    const int gf = gauge_fmt(U->prec, U->recon);
    char *d = (char *)U->data;
    const int Vh = (int)ctx->Vh, st = (int)ctx->stride;
    switch (gf) {
    case G_F64_18:
        hipLaunchKernelGGL(k_gauge_from_ref<G_F64_18>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           d, src, mu, Vh, st);
        break;
    case G_F64_12:
        hipLaunchKernelGGL(k_gauge_from_ref<G_F64_12>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           d, src, mu, Vh, st);
        break;
    case G_F32_18:
        hipLaunchKernelGGL(k_gauge_from_ref<G_F32_18>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           d, src, mu, Vh, st);
        break;
    case G_F32_12:
        hipLaunchKernelGGL(k_gauge_from_ref<G_F32_12>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           d, src, mu, Vh, st);
        break;
    default:
        hipLaunchKernelGGL(k_gauge_from_ref<G_F16_12>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           d, src, mu, Vh, st);
        break;
    }
    return wd_check_launch(ctx, "k_gauge_from_ref");
}

// dst[rd] = src[rs], converting between any two spinor formats
int wd_convert_rhs(wd_ctx *ctx, wd_field *dst, int rd, const wd_field *src, int rs)
{
    void *d = wd_rhs_ptr(dst, rd);
    const void *s = wd_rhs_ptr(src, rs);
    float *ds = wd_rhs_scale(dst, rd);
    const float *ss = wd_rhs_scale(src, rs);
    if (dst->prec == src->prec) {
        WD_HIP_CHECK(ctx, hipMemcpyAsync(d, s, src->rhs_elems * prec_bytes_complex(src->prec),
                                         hipMemcpyDeviceToDevice, ctx->stream));
        if (ds)
            WD_HIP_CHECK(ctx, hipMemcpyAsync(ds, ss, sizeof(float) * 2 * ctx->stride,
                                             hipMemcpyDeviceToDevice, ctx->stream));
        return WD_OK;
    }
    const int a = src->prec, b = dst->prec;
    if (a == WD_F64 && b == WD_F32)
        launch_convert<WD_F64, WD_F32>(ctx, d, ds, s, ss);
    else if (a == WD_F64 && b == WD_F16)
        launch_convert<WD_F64, WD_F16>(ctx, d, ds, s, ss);
    else if (a == WD_F32 && b == WD_F64)
        launch_convert<WD_F32, WD_F64>(ctx, d, ds, s, ss);
    else if (a == WD_F32 && b == WD_F16)
        launch_convert<WD_F32, WD_F16>(ctx, d, ds, s, ss);
    else if (a == WD_F16 && b == WD_F64)
        launch_convert<WD_F16, WD_F64>(ctx, d, ds, s, ss);
    else
        launch_convert<WD_F16, WD_F32>(ctx, d, ds, s, ss);
    return wd_check_launch(ctx, "k_spinor_convert");
}

extern "C" int wd_field_copy(wd_ctx *ctx, wd_field *dst, const wd_field *src)
{
    if (!ctx || !dst || !src)
        return WD_EARG;
    WD_ARG_CHECK(ctx, dst != src, "copy: dst and src must be distinct");
    WD_ARG_CHECK(ctx, dst->nrhs == src->nrhs, "copy: nrhs mismatch");
    WD_ARG_CHECK(ctx, dst->ctx == ctx && src->ctx == ctx, "copy: field from another context");
    for (int r = 0; r < src->nrhs; ++r) {
        int rc = wd_convert_rhs(ctx, dst, r, src, r);
        if (rc)
            return rc;
    }
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}
