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

// One link (9 complex, row-major) stored at site i of a (mu, parity) block of
// device gauge format GF (12-real formats keep rows 0 and 1).
template <int GF>
__device__ __forceinline__ void store_link(char *base, int stride, int i, const double2 (&s)[9])
{
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
    double2 s[9];
    for (int k = 0; k < 9; ++k)
        s[k] = src[(size_t)e * 9 + k];
    store_link<GF>(dst + (size_t)(mu * 2 + p) * block, stride, i, s);
}

// ---- temporal gauge (WD_RECON_12T) --------------------------------------------
// Omega(x, y, z, t) = U_t(t = 0) U_t(t = 1) ... U_t(t - 1) along each t column
// (Omega = 1 on slice 0), so that Omega(x) U_t(x) Omega(x + t)^dag = 1 off the
// last slice.  One thread per column walks t; FP64 products.  omega: complex
// [parity][9][stride]; ut: the t links in ref layout ([parity][site][9]).
__global__ __launch_bounds__(kBlock) void k_tg_omega(double2 *__restrict__ omega,
                                                     const double2 *__restrict__ ut, Geom g)
{
    const int Lx = 2 * g.Lxh;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= Lx * g.Ly * g.Lz)
        return;
    const int x = c % Lx, y = (c / Lx) % g.Ly, z = c / (Lx * g.Ly);
    cpx<double> W[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            W[a][b] = {a == b ? 1.0 : 0.0, 0.0};
    for (int t = 0; t < g.Lt; ++t) {
        const int p = (x + y + z + t) & 1;
        const int i = (x >> 1) + g.Lxh * (y + g.Ly * (z + g.Lz * t));
        double2 *o = omega + (size_t)p * 9 * g.stride;
        for (int k = 0; k < 9; ++k)
            o[(size_t)k * g.stride + i] = make_double2(W[k / 3][k % 3].r, W[k / 3][k % 3].i);
        if (t + 1 == g.Lt)
            break;
        const double2 *u = ut + ((size_t)p * g.Vh + i) * 9;
        cpx<double> N[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                cpx<double> acc = {0.0, 0.0};
                for (int k = 0; k < 3; ++k)
                    acc = cfma(acc, W[a][k], cpx<double>{u[3 * k + b].x, u[3 * k + b].y});
                N[a][b] = acc;
            }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                W[a][b] = N[a][b];
    }
}

// U'_mu(x) = Omega(x) U_mu(x) Omega(x + mu)^dag in FP64, stored in format GF
template <int GF>
__global__ __launch_bounds__(kBlock) void k_tg_links(char *__restrict__ dst,
                                                     const double2 *__restrict__ src,
                                                     const double2 *__restrict__ omega, int mu,
                                                     Geom g)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * g.Vh)
        return;
    const int p = e >= g.Vh, i = e - p * g.Vh;
    int rr = i / g.Lxh;
    const int xh = i - rr * g.Lxh;
    const int y = rr % g.Ly;
    rr /= g.Ly;
    const int z = rr % g.Lz;
    const int t = rr / g.Lz;
    const int x = 2 * xh + ((p + y + z + t) & 1);
    int j; // x + mu, of the other parity
    if (mu == 0) {
        const int xf = (x + 1 == 2 * g.Lxh) ? 0 : x + 1;
        j = i - xh + (xf >> 1);
    } else if (mu == 1) {
        j = (y + 1 == g.Ly) ? i - (g.Ly - 1) * g.Lxh : i + g.Lxh;
    } else if (mu == 2) {
        const int s = g.Lxh * g.Ly;
        j = (z + 1 == g.Lz) ? i - (g.Lz - 1) * s : i + s;
    } else {
        const int s = g.Lxh * g.Ly * g.Lz;
        j = (t + 1 == g.Lt) ? i - (g.Lt - 1) * s : i + s;
    }
    const double2 *oa = omega + (size_t)p * 9 * g.stride;
    const double2 *ob = omega + (size_t)(1 - p) * 9 * g.stride;
    cpx<double> A[3][3], B[3][3], Uu[3][3];
    for (int k = 0; k < 9; ++k) {
        const double2 va = oa[(size_t)k * g.stride + i], vb = ob[(size_t)k * g.stride + j];
        const double2 vu = src[(size_t)e * 9 + k];
        A[k / 3][k % 3] = {va.x, va.y};
        B[k / 3][k % 3] = {vb.x, vb.y};
        Uu[k / 3][k % 3] = {vu.x, vu.y};
    }
    double2 s[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            // (A U B^dag)_ab = sum_c (A U)_ac conj(B_bc)
            cpx<double> acc = {0.0, 0.0};
            for (int c = 0; c < 3; ++c) {
                cpx<double> au = {0.0, 0.0};
                for (int k = 0; k < 3; ++k)
                    au = cfma(au, A[a][k], Uu[k][c]);
                acc = cfma_conj(acc, B[b][c], au);
            }
            s[3 * a + b] = make_double2(acc.r, acc.i);
        }
    const size_t block = (size_t)gauge_chunks(GF) * gauge_chunk_bytes(GF) * g.stride;
    store_link<GF>(dst + (size_t)(mu * 2 + p) * block, g.stride, i, s);
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

int wd_launch_tg_omega(wd_ctx *ctx, wd_gauge *U, const double2 *src)
{
    const Geom &g = ctx->geom;
    const int cols = 2 * g.Lxh * g.Ly * g.Lz;
    hipLaunchKernelGGL(k_tg_omega, dim3((cols + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream,
                       U->omega, src, g);
    return wd_check_launch(ctx, "k_tg_omega");
}

int wd_launch_tg_links(wd_ctx *ctx, wd_gauge *U, int mu, const double2 *src)
{
    const int grid = (int)((2 * ctx->Vh + kBlock - 1) / kBlock);
    char *d = (char *)U->data;
    if (gauge_fmt(U->prec, U->recon) == G_F32_12)
        hipLaunchKernelGGL(k_tg_links<G_F32_12>, dim3(grid), dim3(kBlock), 0, ctx->stream, d, src,
                           (const double2 *)U->omega, mu, ctx->geom);
    else
        hipLaunchKernelGGL(k_tg_links<G_F16_12>, dim3(grid), dim3(kBlock), 0, ctx->stream, d, src,
                           (const double2 *)U->omega, mu, ctx->geom);
    return wd_check_launch(ctx, "k_tg_links");
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
