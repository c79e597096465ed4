// wd_core.hip -- context, fields, gauge storage, layout conversion and the
// operator / reduction entry points of include/wilson.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "wd_blas.h"
#include "wd_internal.h"

// ---- errors -------------------------------------------------------------------
int wd_set_error(wd_ctx *ctx, int code, const std::string &msg)
{
    if (ctx)
        ctx->last_error = msg;
    return code;
}

int wd_check_launch(wd_ctx *ctx, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return wd_set_error(ctx, WD_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return WD_OK;
}

extern "C" const char *wd_strerror(int code)
{
    switch (code) {
    case WD_OK:
        return "ok";
    case WD_EARG:
        return "invalid argument";
    case WD_EHIP:
        return "HIP runtime error";
    case WD_ENOMEM:
        return "out of memory";
    case WD_EBREAKDOWN:
        return "solver breakdown";
    case WD_ENOTCONV:
        return "not converged";
    default:
        return "unknown status";
    }
}

extern "C" const char *wd_last_error(const wd_ctx *ctx)
{
    return ctx ? ctx->last_error.c_str() : "no context";
}

// ---- layout conversion kernels ---------------------------------------------------
namespace {

// ref spinor (complex<double>[V*12], eo*12 + k) -> device SoA of precision T
__global__ __launch_bounds__(kBlock) void k_spinor_from_ref_f64(double2 *__restrict__ dst,
                                                                const double2 *__restrict__ src,
                                                                int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x; // eo site index in [0, 2*Vh)
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    for (int k = 0; k < 12; ++k)
        dst[(size_t)(p * 12 + k) * stride + i] = src[(size_t)e * 12 + k];
}

__global__ __launch_bounds__(kBlock) void k_spinor_to_ref_f64(double2 *__restrict__ dst,
                                                              const double2 *__restrict__ src,
                                                              int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    for (int k = 0; k < 12; ++k)
        dst[(size_t)e * 12 + k] = src[(size_t)(p * 12 + k) * stride + i];
}

__global__ __launch_bounds__(kBlock) void k_spinor_from_ref_f32(float2 *__restrict__ dst,
                                                                const double2 *__restrict__ src,
                                                                int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    for (int k = 0; k < 12; ++k) {
        const double2 v = src[(size_t)e * 12 + k];
        dst[(size_t)(p * 12 + k) * stride + i] = make_float2((float)v.x, (float)v.y);
    }
}

__global__ __launch_bounds__(kBlock) void k_spinor_to_ref_f32(double2 *__restrict__ dst,
                                                              const float2 *__restrict__ src,
                                                              int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    for (int k = 0; k < 12; ++k) {
        const float2 v = src[(size_t)(p * 12 + k) * stride + i];
        dst[(size_t)e * 12 + k] = make_double2(v.x, v.y);
    }
}

// F16 storage: per-site scale s = max |component|, values stored as v / s.
__global__ __launch_bounds__(kBlock) void k_spinor_from_ref_f16(__half2 *__restrict__ dst,
                                                                float *__restrict__ scale,
                                                                const double2 *__restrict__ src,
                                                                int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    double m = 0.0;
    for (int k = 0; k < 12; ++k) {
        const double2 v = src[(size_t)e * 12 + k];
        m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
    }
    const float s = (float)m;
    const float inv = s > 0.f ? 1.0f / s : 0.f;
    scale[(size_t)p * stride + i] = s;
    for (int k = 0; k < 12; ++k) {
        const double2 v = src[(size_t)e * 12 + k];
        dst[(size_t)(p * 12 + k) * stride + i] =
            __floats2half2_rn((float)v.x * inv, (float)v.y * inv);
    }
}

__global__ __launch_bounds__(kBlock) void k_spinor_to_ref_f16(double2 *__restrict__ dst,
                                                              const __half2 *__restrict__ src,
                                                              const float *__restrict__ scale,
                                                              int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    const float s = scale[(size_t)p * stride + i];
    for (int k = 0; k < 12; ++k) {
        const float2 v = __half22float2(src[(size_t)(p * 12 + k) * stride + i]);
        dst[(size_t)e * 12 + k] = make_double2((double)v.x * s, (double)v.y * s);
    }
}

// ref gauge A[mu] (complex<double>[V*9]) -> device [mu][parity][nrow*3][stride]
template <typename T2>
__device__ __forceinline__ T2 cvt2(double2 v);
template <>
__device__ __forceinline__ double2 cvt2<double2>(double2 v)
{
    return v;
}
template <>
__device__ __forceinline__ float2 cvt2<float2>(double2 v)
{
    return make_float2((float)v.x, (float)v.y);
}
template <>
__device__ __forceinline__ __half2 cvt2<__half2>(double2 v)
{
    return __floats2half2_rn((float)v.x, (float)v.y);
}

template <typename T2>
__global__ __launch_bounds__(kBlock) void k_gauge_from_ref(T2 *__restrict__ dst,
                                                           const double2 *__restrict__ src,
                                                           int mu, int nrow, int Vh, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * Vh)
        return;
    const int p = e >= Vh, i = e - p * Vh;
    T2 *base = dst + (size_t)((mu * 2 + p) * nrow * 3) * stride;
    for (int k = 0; k < nrow * 3; ++k)
        base[(size_t)k * stride + i] = cvt2<T2>(src[(size_t)e * 9 + k]);
}

} // namespace

int wd_ensure_staging(wd_ctx *ctx, size_t bytes)
{
    if (ctx->staging_bytes >= bytes)
        return WD_OK;
    if (ctx->staging) {
        WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        WD_HIP_CHECK(ctx, hipFree(ctx->staging));
        ctx->staging = nullptr;
        ctx->staging_bytes = 0;
    }
    if (hipMalloc(&ctx->staging, bytes) != hipSuccess) {
        ctx->staging = nullptr;
        (void)hipGetLastError();
        return wd_set_error(ctx, WD_ENOMEM, "staging buffer allocation failed");
    }
    ctx->staging_bytes = bytes;
    return WD_OK;
}

// ---- context ----------------------------------------------------------------------
extern "C" int wd_create(wd_ctx **out, const int L[4], const int bc[4], int device)
{
    if (!out || !L)
        return WD_EARG;
    *out = nullptr;
    for (int mu = 0; mu < 4; ++mu)
        if (L[mu] < 2 || (L[mu] & 1))
            return WD_EARG;
    if (bc)
        for (int mu = 0; mu < 4; ++mu)
            if (bc[mu] != 1 && bc[mu] != -1)
                return WD_EARG;
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    const long Vh = V / 2;
    const long stride = (Vh + 63) / 64 * 64;
    // 32-bit index arithmetic inside kernels: 24 complex components per site
    if (stride * 24 >= (1L << 31))
        return WD_EARG;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return WD_EHIP;
    }
    wd_ctx *c = new (std::nothrow) wd_ctx();
    if (!c)
        return WD_ENOMEM;
    for (int mu = 0; mu < 4; ++mu) {
        c->L[mu] = L[mu];
        c->bc[mu] = bc ? bc[mu] : 1;
        c->geom.bc[mu] = (double)c->bc[mu];
    }
    c->device = device;
    c->V = V;
    c->Vh = Vh;
    c->stride = stride;
    c->geom.Lxh = L[0] / 2;
    c->geom.Ly = L[1];
    c->geom.Lz = L[2];
    c->geom.Lt = L[3];
    c->geom.Vh = (int)Vh;
    c->geom.stride = (int)stride;
    c->xcd_remap = 1;
    // partials: enough for a both-parity hop launch and the reduction grid
    const int hop_blocks = (int)(2 * ((Vh + kBlock - 1) / kBlock));
    c->max_partials = hop_blocks > kRedBlocks ? hop_blocks : kRedBlocks;
    c->dscalars_bytes = 1 << 16;
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&c->partials, sizeof(double2) * c->max_partials) == hipSuccess;
    ok = ok && hipMalloc(&c->dscalars, c->dscalars_bytes) == hipSuccess;
    ok = ok && hipHostMalloc(&c->hpinned, 4096, hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        wd_destroy(c);
        return WD_ENOMEM;
    }
    *out = c;
    return WD_OK;
}

extern "C" void wd_destroy(wd_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    if (c->partials)
        (void)hipFree(c->partials);
    if (c->dscalars)
        (void)hipFree(c->dscalars);
    if (c->staging)
        (void)hipFree(c->staging);
    if (c->hpinned)
        (void)hipHostFree(c->hpinned);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int wd_sync(wd_ctx *ctx)
{
    if (!ctx)
        return WD_EARG;
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

extern "C" long wd_volume(const wd_ctx *ctx) { return ctx ? ctx->V : 0; }

// ---- gauge --------------------------------------------------------------------------
extern "C" int wd_gauge_create(wd_ctx *ctx, wd_gauge **U, int prec, int recon)
{
    if (!ctx || !U)
        return WD_EARG;
    *U = nullptr;
    WD_ARG_CHECK(ctx, prec == WD_F64 || prec == WD_F32 || prec == WD_F16, "gauge: bad precision");
    WD_ARG_CHECK(ctx, recon == WD_RECON_18 || recon == WD_RECON_12, "gauge: bad recon");
    wd_gauge *g = new (std::nothrow) wd_gauge();
    if (!g)
        return WD_ENOMEM;
    g->ctx = ctx;
    g->prec = prec;
    g->recon = recon;
    const int ncomp = recon == WD_RECON_18 ? 9 : 6;
    g->bytes = (size_t)4 * 2 * ncomp * ctx->stride * prec_bytes_complex(prec);
    if (hipMalloc(&g->data, g->bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete g;
        return wd_set_error(ctx, WD_ENOMEM, "gauge allocation failed");
    }
    int rc = wd_launch_zero(ctx, g->data, g->bytes);
    if (rc) {
        wd_gauge_destroy(g);
        return rc;
    }
    *U = g;
    return WD_OK;
}

extern "C" void wd_gauge_destroy(wd_gauge *U)
{
    if (!U)
        return;
    (void)hipStreamSynchronize(U->ctx->stream);
    (void)hipFree(U->data);
    delete U;
}

extern "C" int wd_gauge_upload_ref(wd_ctx *ctx, wd_gauge *U, const double *const A[4])
{
    if (!ctx || !U || !A)
        return WD_EARG;
    const size_t bytes = sizeof(double2) * 9 * ctx->V;
    int rc = wd_ensure_staging(ctx, bytes);
    if (rc)
        return rc;
    const int nrow = U->recon == WD_RECON_18 ? 3 : 2;
    const int grid = (int)((2 * ctx->Vh + kBlock - 1) / kBlock);
    for (int mu = 0; mu < 4; ++mu) {
        WD_ARG_CHECK(ctx, A[mu] != nullptr, "gauge upload: null direction array");
        WD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->staging, A[mu], bytes, hipMemcpyHostToDevice,
                                         ctx->stream));
        const double2 *src = (const double2 *)ctx->staging;
        if (U->prec == WD_F64)
            hipLaunchKernelGGL(k_gauge_from_ref<double2>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                               (double2 *)U->data, src, mu, nrow, (int)ctx->Vh, (int)ctx->stride);
        else if (U->prec == WD_F32)
            hipLaunchKernelGGL(k_gauge_from_ref<float2>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                               (float2 *)U->data, src, mu, nrow, (int)ctx->Vh, (int)ctx->stride);
        else
            hipLaunchKernelGGL(k_gauge_from_ref<__half2>, dim3(grid), dim3(kBlock), 0,
                               ctx->stream, (__half2 *)U->data, src, mu, nrow, (int)ctx->Vh,
                               (int)ctx->stride);
        rc = wd_check_launch(ctx, "k_gauge_from_ref");
        if (rc)
            return rc;
    }
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

// ---- fields -------------------------------------------------------------------------
extern "C" int wd_field_create(wd_ctx *ctx, wd_field **f, int prec, int nrhs)
{
    if (!ctx || !f)
        return WD_EARG;
    *f = nullptr;
    WD_ARG_CHECK(ctx, prec == WD_F64 || prec == WD_F32 || prec == WD_F16, "field: bad precision");
    WD_ARG_CHECK(ctx, nrhs >= 1, "field: nrhs < 1");
    wd_field *x = new (std::nothrow) wd_field();
    if (!x)
        return WD_ENOMEM;
    x->ctx = ctx;
    x->prec = prec;
    x->nrhs = nrhs;
    x->rhs_elems = (size_t)24 * ctx->stride;
    x->bytes = x->rhs_elems * nrhs * prec_bytes_complex(prec);
    if (hipMalloc(&x->data, x->bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete x;
        return wd_set_error(ctx, WD_ENOMEM, "field allocation failed");
    }
    if (prec == WD_F16) {
        const size_t sb = sizeof(float) * 2 * ctx->stride * nrhs;
        if (hipMalloc((void **)&x->scale, sb) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipFree(x->data);
            delete x;
            return wd_set_error(ctx, WD_ENOMEM, "field scale allocation failed");
        }
    }
    int rc = wd_field_zero(ctx, x);
    if (rc) {
        wd_field_destroy(x);
        return rc;
    }
    *f = x;
    return WD_OK;
}

extern "C" void wd_field_destroy(wd_field *f)
{
    if (!f)
        return;
    (void)hipStreamSynchronize(f->ctx->stream);
    (void)hipFree(f->data);
    if (f->scale)
        (void)hipFree(f->scale);
    delete f;
}

extern "C" int wd_field_zero(wd_ctx *ctx, wd_field *f)
{
    if (!ctx || !f)
        return WD_EARG;
    int rc = wd_launch_zero(ctx, f->data, f->bytes);
    if (!rc && f->scale)
        rc = wd_launch_zero(ctx, f->scale, sizeof(float) * 2 * ctx->stride * f->nrhs);
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

static void *rhs_ptr(const wd_field *f, int rhs)
{
    return (char *)f->data + (size_t)rhs * f->rhs_elems * prec_bytes_complex(f->prec);
}

static float *rhs_scale(const wd_field *f, int rhs)
{
    return f->scale ? f->scale + (size_t)rhs * 2 * f->ctx->stride : nullptr;
}

extern "C" int wd_field_upload_ref(wd_ctx *ctx, wd_field *f, const double *A, int rhs)
{
    if (!ctx || !f || !A)
        return WD_EARG;
    WD_ARG_CHECK(ctx, rhs >= 0 && rhs < f->nrhs, "field upload: rhs out of range");
    const size_t bytes = sizeof(double2) * 12 * ctx->V;
    int rc = wd_ensure_staging(ctx, bytes);
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx,
                 hipMemcpyAsync(ctx->staging, A, bytes, hipMemcpyHostToDevice, ctx->stream));
    const int grid = (int)((2 * ctx->Vh + kBlock - 1) / kBlock);
    const double2 *src = (const double2 *)ctx->staging;
    const int Vh = (int)ctx->Vh, st = (int)ctx->stride;
    if (f->prec == WD_F64)
        hipLaunchKernelGGL(k_spinor_from_ref_f64, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           (double2 *)rhs_ptr(f, rhs), src, Vh, st);
    else if (f->prec == WD_F32)
        hipLaunchKernelGGL(k_spinor_from_ref_f32, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           (float2 *)rhs_ptr(f, rhs), src, Vh, st);
    else
        hipLaunchKernelGGL(k_spinor_from_ref_f16, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           (__half2 *)rhs_ptr(f, rhs), rhs_scale(f, rhs), src, Vh, st);
    rc = wd_check_launch(ctx, "k_spinor_from_ref");
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

extern "C" int wd_field_download_ref(wd_ctx *ctx, const wd_field *f, double *A, int rhs)
{
    if (!ctx || !f || !A)
        return WD_EARG;
    WD_ARG_CHECK(ctx, rhs >= 0 && rhs < f->nrhs, "field download: rhs out of range");
    const size_t bytes = sizeof(double2) * 12 * ctx->V;
    int rc = wd_ensure_staging(ctx, bytes);
    if (rc)
        return rc;
    const int grid = (int)((2 * ctx->Vh + kBlock - 1) / kBlock);
    double2 *dst = (double2 *)ctx->staging;
    const int Vh = (int)ctx->Vh, st = (int)ctx->stride;
    if (f->prec == WD_F64)
        hipLaunchKernelGGL(k_spinor_to_ref_f64, dim3(grid), dim3(kBlock), 0, ctx->stream, dst,
                           (const double2 *)rhs_ptr(f, rhs), Vh, st);
    else if (f->prec == WD_F32)
        hipLaunchKernelGGL(k_spinor_to_ref_f32, dim3(grid), dim3(kBlock), 0, ctx->stream, dst,
                           (const float2 *)rhs_ptr(f, rhs), Vh, st);
    else
        hipLaunchKernelGGL(k_spinor_to_ref_f16, dim3(grid), dim3(kBlock), 0, ctx->stream, dst,
                           (const __half2 *)rhs_ptr(f, rhs), rhs_scale(f, rhs), Vh, st);
    rc = wd_check_launch(ctx, "k_spinor_to_ref");
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipMemcpyAsync(A, ctx->staging, bytes, hipMemcpyDeviceToHost, ctx->stream));
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

// ---- operator entry points (F64) ------------------------------------------------------
static int check_f64_op(wd_ctx *ctx, const wd_field *out, const wd_field *in, const wd_gauge *U)
{
    WD_ARG_CHECK(ctx, out && in, "operator: null field");
    WD_ARG_CHECK(ctx, out != in, "operator: out and in must be distinct fields");
    WD_ARG_CHECK(ctx, out->ctx == ctx && in->ctx == ctx, "operator: field from another context");
    WD_ARG_CHECK(ctx, out->prec == WD_F64 && in->prec == WD_F64,
                 "operator: F64 fields required");
    WD_ARG_CHECK(ctx, out->nrhs == 1 && in->nrhs == 1, "operator: single-rhs fields required");
    if (U) {
        WD_ARG_CHECK(ctx, U->ctx == ctx, "operator: gauge from another context");
        WD_ARG_CHECK(ctx, U->prec == WD_F64 && U->recon == WD_RECON_18,
                     "operator: F64 18-real gauge required");
    }
    return WD_OK;
}

extern "C" int wd_hop(wd_ctx *ctx, wd_field *out, const wd_field *in, const wd_gauge *U,
                      int dagger, int cb)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, U != nullptr, "hop: null gauge");
    int rc = check_f64_op(ctx, out, in, U);
    if (rc)
        return rc;
    WD_ARG_CHECK(ctx, cb == 0 || cb == 1, "hop: cb must be 0 or 1");
    double2 *o = (double2 *)out->data;
    // the other parity is zeroed, as DslashoffdNew's memset does
    rc = wd_launch_zero(ctx, o + (size_t)(1 - cb) * 12 * ctx->stride,
                        sizeof(double2) * 12 * ctx->stride);
    if (rc)
        return rc;
    rc = wd_launch_hop_f64(ctx, o, (const double2 *)in->data, (const double2 *)U->data, dagger,
                           cb, HOP_ONLY, 0.0, nullptr, nullptr, nullptr, nullptr);
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

extern "C" int wd_diag(wd_ctx *ctx, wd_field *out, const wd_field *in, double mass)
{
    if (!ctx)
        return WD_EARG;
    int rc = check_f64_op(ctx, out, in, nullptr);
    if (rc)
        return rc;
    rc = wd_launch_scale(ctx, (double2 *)out->data, (const double2 *)in->data, 4.0 + mass,
                         out->rhs_elems);
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

extern "C" int wd_apply_M(wd_ctx *ctx, wd_field *out, const wd_field *in, const wd_gauge *U,
                          double mass, int dagger)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, U != nullptr, "apply_M: null gauge");
    int rc = check_f64_op(ctx, out, in, U);
    if (rc)
        return rc;
    rc = wd_launch_hop_f64(ctx, (double2 *)out->data, (const double2 *)in->data,
                           (const double2 *)U->data, dagger, -1, HOP_M, 4.0 + mass,
                           (const double2 *)in->data, nullptr, nullptr, nullptr);
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

extern "C" int wd_apply_MdagM(wd_ctx *ctx, wd_field *out, const wd_field *in, wd_field *tmp,
                              const wd_gauge *U, double mass)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, U != nullptr, "apply_MdagM: null gauge");
    int rc = check_f64_op(ctx, out, in, U);
    if (!rc)
        rc = check_f64_op(ctx, tmp, in, U);
    if (rc)
        return rc;
    WD_ARG_CHECK(ctx, tmp != out, "apply_MdagM: tmp and out must be distinct");
    rc = wd_launch_hop_f64(ctx, (double2 *)tmp->data, (const double2 *)in->data,
                           (const double2 *)U->data, 0, -1, HOP_M, 4.0 + mass,
                           (const double2 *)in->data, nullptr, nullptr, nullptr);
    if (!rc)
        rc = wd_launch_hop_f64(ctx, (double2 *)out->data, (const double2 *)tmp->data,
                               (const double2 *)U->data, 1, -1, HOP_M, 4.0 + mass,
                               (const double2 *)tmp->data, nullptr, nullptr, nullptr);
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

// ---- reductions -------------------------------------------------------------------------
extern "C" int wd_norm2(wd_ctx *ctx, const wd_field *f, double *out)
{
    if (!ctx || !f || !out)
        return WD_EARG;
    WD_ARG_CHECK(ctx, f->prec == WD_F64 && f->nrhs == 1, "norm2: single-rhs F64 field required");
    double2 *res = (double2 *)ctx->dscalars;
    int rc = wd_launch_dot(ctx, (const double2 *)f->data, nullptr, f->rhs_elems, res);
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->hpinned, res, sizeof(double2), hipMemcpyDeviceToHost,
                                     ctx->stream));
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *out = ((double2 *)ctx->hpinned)->x;
    return WD_OK;
}

extern "C" int wd_dot(wd_ctx *ctx, const wd_field *a, const wd_field *b, double out[2])
{
    if (!ctx || !a || !b || !out)
        return WD_EARG;
    WD_ARG_CHECK(ctx, a->prec == WD_F64 && b->prec == WD_F64 && a->nrhs == 1 && b->nrhs == 1,
                 "dot: single-rhs F64 fields required");
    double2 *res = (double2 *)ctx->dscalars;
    int rc = wd_launch_dot(ctx, (const double2 *)a->data, (const double2 *)b->data, a->rhs_elems,
                           res);
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->hpinned, res, sizeof(double2), hipMemcpyDeviceToHost,
                                     ctx->stream));
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    out[0] = ((double2 *)ctx->hpinned)->x;
    out[1] = ((double2 *)ctx->hpinned)->y;
    return WD_OK;
}

// ---- precision conversion -----------------------------------------------------------------
namespace {

// F64 SoA -> F32 SoA, elementwise over one rhs (24 * stride complex)
__global__ __launch_bounds__(kBlock) void k_f64_to_f32(float2 *__restrict__ dst,
                                                       const double2 *__restrict__ src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBlock) {
        const double2 v = src[i];
        dst[i] = make_float2((float)v.x, (float)v.y);
    }
}

__global__ __launch_bounds__(kBlock) void k_f32_to_f64(double2 *__restrict__ dst,
                                                       const float2 *__restrict__ src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBlock) {
        const float2 v = src[i];
        dst[i] = make_double2(v.x, v.y);
    }
}

// F64 SoA -> F16 SoA + per-site scale; one thread per (parity, site)
__global__ __launch_bounds__(kBlock) void k_f64_to_f16(__half2 *__restrict__ dst,
                                                       float *__restrict__ scale,
                                                       const double2 *__restrict__ src, int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x; // [0, 2*stride)
    if (e >= 2 * stride)
        return;
    const int p = e / stride, i = e - p * stride;
    const double2 *s = src + (size_t)p * 12 * stride + i;
    double m = 0.0;
    for (int k = 0; k < 12; ++k) {
        const double2 v = s[(size_t)k * stride];
        m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
    }
    const float sc = (float)m;
    const float inv = sc > 0.f ? 1.0f / sc : 0.f;
    scale[e] = sc;
    __half2 *d = dst + (size_t)p * 12 * stride + i;
    for (int k = 0; k < 12; ++k) {
        const double2 v = s[(size_t)k * stride];
        d[(size_t)k * stride] = __floats2half2_rn((float)v.x * inv, (float)v.y * inv);
    }
}

__global__ __launch_bounds__(kBlock) void k_f16_to_f64(double2 *__restrict__ dst,
                                                       const __half2 *__restrict__ src,
                                                       const float *__restrict__ scale,
                                                       int stride)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= 2 * stride)
        return;
    const int p = e / stride, i = e - p * stride;
    const float sc = scale[e];
    const __half2 *s = src + (size_t)p * 12 * stride + i;
    double2 *d = dst + (size_t)p * 12 * stride + i;
    for (int k = 0; k < 12; ++k) {
        const float2 v = __half22float2(s[(size_t)k * stride]);
        d[(size_t)k * stride] = make_double2((double)v.x * sc, (double)v.y * sc);
    }
}

} // namespace

// dst[rhs] = src[rhs] with precision conversion; F64 <-> F32 and F64 <-> F16,
// or a plain copy for equal precisions.
int wd_convert_rhs(wd_ctx *ctx, wd_field *dst, int rd, const wd_field *src, int rs)
{
    const size_t n = src->rhs_elems;
    const int grid_e = wd_elem_grid(n);
    const int grid_s = (int)((2 * ctx->stride + kBlock - 1) / kBlock);
    void *d = rhs_ptr(dst, rd);
    const void *s = rhs_ptr(src, rs);
    if (dst->prec == src->prec) {
        WD_HIP_CHECK(ctx, hipMemcpyAsync(d, s, n * prec_bytes_complex(src->prec),
                                         hipMemcpyDeviceToDevice, ctx->stream));
        if (dst->scale)
            WD_HIP_CHECK(ctx, hipMemcpyAsync(rhs_scale(dst, rd), rhs_scale(src, rs),
                                             sizeof(float) * 2 * ctx->stride,
                                             hipMemcpyDeviceToDevice, ctx->stream));
        return WD_OK;
    }
    if (src->prec == WD_F64 && dst->prec == WD_F32)
        hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_e), dim3(kBlock), 0, ctx->stream, (float2 *)d,
                           (const double2 *)s, n);
    else if (src->prec == WD_F32 && dst->prec == WD_F64)
        hipLaunchKernelGGL(k_f32_to_f64, dim3(grid_e), dim3(kBlock), 0, ctx->stream, (double2 *)d,
                           (const float2 *)s, n);
    else if (src->prec == WD_F64 && dst->prec == WD_F16)
        hipLaunchKernelGGL(k_f64_to_f16, dim3(grid_s), dim3(kBlock), 0, ctx->stream,
                           (__half2 *)d, rhs_scale(dst, rd), (const double2 *)s,
                           (int)ctx->stride);
    else if (src->prec == WD_F16 && dst->prec == WD_F64)
        hipLaunchKernelGGL(k_f16_to_f64, dim3(grid_s), dim3(kBlock), 0, ctx->stream,
                           (double2 *)d, (const __half2 *)s, rhs_scale(src, rs),
                           (int)ctx->stride);
    else
        return wd_set_error(ctx, WD_EARG, "copy: F32 <-> F16 conversion is not provided");
    return wd_check_launch(ctx, "convert");
}

extern "C" int wd_field_copy(wd_ctx *ctx, wd_field *dst, const wd_field *src)
{
    if (!ctx || !dst || !src)
        return WD_EARG;
    WD_ARG_CHECK(ctx, dst != src, "copy: dst and src must be distinct");
    WD_ARG_CHECK(ctx, dst->nrhs == src->nrhs, "copy: nrhs mismatch");
    for (int r = 0; r < src->nrhs; ++r) {
        int rc = wd_convert_rhs(ctx, dst, r, src, r);
        if (rc)
            return rc;
    }
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

// ---- measurement ----------------------------------------------------------------------------
extern "C" int wd_time_op(wd_ctx *ctx, int op, wd_field *out, const wd_field *in,
                          const wd_gauge *U, double mass, int dagger, int cb, int iters,
                          double *ms_per_launch)
{
    if (!ctx || !ms_per_launch)
        return WD_EARG;
    WD_ARG_CHECK(ctx, iters >= 1, "time_op: iters < 1");
    WD_ARG_CHECK(ctx, U != nullptr, "time_op: null gauge");
    int rc = check_f64_op(ctx, out, in, U);
    if (rc)
        return rc;
    WD_ARG_CHECK(ctx, op == WD_OP_HOP || op == WD_OP_M, "time_op: unknown op");
    WD_ARG_CHECK(ctx, cb == 0 || cb == 1, "time_op: cb must be 0 or 1");
    hipEvent_t e0, e1;
    WD_HIP_CHECK(ctx, hipEventCreate(&e0));
    WD_HIP_CHECK(ctx, hipEventCreate(&e1));
    WD_HIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
    for (int k = 0; k < iters && !rc; ++k) {
        if (op == WD_OP_HOP)
            rc = wd_launch_hop_f64(ctx, (double2 *)out->data, (const double2 *)in->data,
                                   (const double2 *)U->data, dagger, cb, HOP_ONLY, 0.0, nullptr,
                                   nullptr, nullptr, nullptr);
        else
            rc = wd_launch_hop_f64(ctx, (double2 *)out->data, (const double2 *)in->data,
                                   (const double2 *)U->data, dagger, -1, HOP_M, 4.0 + mass,
                                   (const double2 *)in->data, nullptr, nullptr, nullptr);
    }
    hipError_t er = hipEventRecord(e1, ctx->stream);
    hipError_t es = hipEventSynchronize(e1);
    float ms = 0.f;
    hipError_t ee = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, er);
    WD_HIP_CHECK(ctx, es);
    WD_HIP_CHECK(ctx, ee);
    *ms_per_launch = (double)ms / iters;
    return WD_OK;
}

extern "C" int wd_set_option(wd_ctx *ctx, int option, int value)
{
    if (!ctx)
        return WD_EARG;
    switch (option) {
    case WD_OPT_XCD_REMAP:
        ctx->xcd_remap = value ? 1 : 0;
        return WD_OK;
    default:
        return wd_set_error(ctx, WD_EARG, "set_option: unknown option");
    }
}
