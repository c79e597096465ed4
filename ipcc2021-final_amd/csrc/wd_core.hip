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
    if (wd_malloc(ctx, &ctx->staging, bytes) != hipSuccess) {
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
    {
        const long slice = (long)c->geom.Lxh * L[1] * L[2]; // sites per time slice and parity
        c->geom.tD = (slice % kBlock == 0) ? (int)(slice / kBlock) : 0;
        // default column: 4 z-planes (4 * Lxh * Ly sites); with the parities
        // side by side, measured on MI355X within 3 % of the best width at
        // 16^3x32 ... 64^4
        const int w4 = (int)((4L * c->geom.Lxh * L[1] + kBlock / 2) / kBlock);
        c->geom.tW = (c->geom.tD > 0 && w4 >= 1 && w4 < c->geom.tD) ? w4 : 0;
    }
    c->xcd_remap = 1;
    c->parity_il = 1;
    c->ts_tb = wd_auto_tstream(c);
    c->f64_pairs = 1;
    c->loop_prefetch = 0;
    c->loop_systems = 0;
    // partials: enough for a both-parity hop launch and the streaming
    // reduction grids (kStreamPartials, 256 KiB)
    const int hop_blocks = (int)(2 * ((Vh + kBlock - 1) / kBlock));
    c->max_partials = hop_blocks > kStreamPartials ? hop_blocks : kStreamPartials;
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
    // the allocation-free correlator's scratch address on this device; a
    // failure here only disables that entry point (it then returns an error)
    (void)wd_contract_setup(device);
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
    wd_work_release(c);
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

// ---- device memory ----------------------------------------------------------------
hipError_t wd_malloc(wd_ctx *ctx, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && !ctx->work.empty()) {
        (void)hipGetLastError();
        wd_work_release(ctx);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
    }
    return e;
}

void *wd_work_get(wd_ctx *ctx, size_t bytes)
{
    for (size_t i = 0; i < ctx->work.size(); ++i)
        if (ctx->work[i].first == bytes) {
            void *p = ctx->work[i].second;
            ctx->work[i] = ctx->work.back();
            ctx->work.pop_back();
            return p;
        }
    void *p = nullptr;
    return wd_malloc(ctx, &p, bytes) == hipSuccess ? p : nullptr;
}

void wd_work_put(wd_ctx *ctx, void *p, size_t bytes)
{
    if (p)
        ctx->work.emplace_back(bytes, p);
}

void wd_work_release(wd_ctx *ctx)
{
    if (ctx->work.empty())
        return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &w : ctx->work)
        (void)hipFree(w.second);
    ctx->work.clear();
}

extern "C" int wd_release_workspace(wd_ctx *ctx)
{
    if (!ctx)
        return WD_EARG;
    wd_work_release(ctx);
    return WD_OK;
}

extern "C" int wd_workspace_bytes(const wd_ctx *ctx, size_t *bytes)
{
    if (!ctx || !bytes)
        return WD_EARG;
    size_t t = 0;
    for (const auto &w : ctx->work)
        t += w.first;
    *bytes = t;
    return WD_OK;
}

// ---- gauge --------------------------------------------------------------------------
extern "C" int wd_gauge_create(wd_ctx *ctx, wd_gauge **U, int prec, int recon)
{
    if (!ctx || !U)
        return WD_EARG;
    *U = nullptr;
    WD_ARG_CHECK(ctx, prec == WD_F64 || prec == WD_F32 || prec == WD_F16, "gauge: bad precision");
    WD_ARG_CHECK(ctx, recon == WD_RECON_18 || recon == WD_RECON_12 || recon == WD_RECON_12T,
                 "gauge: bad recon");
    WD_ARG_CHECK(ctx, !(prec == WD_F16 && recon == WD_RECON_18),
                 "gauge: F16 links are stored 12-real only");
    WD_ARG_CHECK(ctx, !(prec == WD_F64 && recon == WD_RECON_12T),
                 "gauge: temporal-gauge links are for the sloppy formats (F32, F16)");
    wd_gauge *g = new (std::nothrow) wd_gauge();
    if (!g)
        return WD_ENOMEM;
    g->ctx = ctx;
    g->prec = prec;
    g->recon = recon;
    const int ncomp = recon == WD_RECON_18 ? 9 : 6;
    g->bytes = (size_t)4 * 2 * ncomp * ctx->stride * prec_bytes_complex(prec);
    if (wd_malloc(ctx, &g->data, g->bytes) != hipSuccess) {
        delete g;
        return wd_set_error(ctx, WD_ENOMEM, "gauge allocation failed");
    }
    const size_t obytes = sizeof(double2) * 2 * 9 * ctx->stride;
    if (recon == WD_RECON_12T && wd_malloc(ctx, (void **)&g->omega, obytes) != hipSuccess) {
        wd_gauge_destroy(g);
        return wd_set_error(ctx, WD_ENOMEM, "gauge allocation failed");
    }
    int rc = wd_launch_zero(ctx, g->data, g->bytes);
    if (!rc && g->omega)
        rc = wd_launch_zero(ctx, g->omega, obytes);
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
    if (U->omega)
        (void)hipFree(U->omega);
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
    for (int mu = 0; mu < 4; ++mu)
        WD_ARG_CHECK(ctx, A[mu] != nullptr, "gauge upload: null direction array");
    const bool tg = U->recon == WD_RECON_12T;
    // temporal gauge: the t links first (Omega), then every direction rotated
    for (int k = 0; k < 4; ++k) {
        const int mu = tg ? (k + 3) % 4 : k;
        WD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->staging, A[mu], bytes, hipMemcpyHostToDevice,
                                         ctx->stream));
        const double2 *src = (const double2 *)ctx->staging;
        if (tg) {
            rc = mu == 3 ? wd_launch_tg_omega(ctx, U, src) : 0;
            if (!rc)
                rc = wd_launch_tg_links(ctx, U, mu, src);
        } else
            rc = wd_launch_gauge_from_ref(ctx, U, mu, src);
        if (rc)
            return rc;
    }
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

// ---- fields -------------------------------------------------------------------------
static int field_create(wd_ctx *ctx, wd_field **f, int prec, int nrhs, bool work)
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
    x->work = work;
    x->rhs_elems = (size_t)24 * ctx->stride;
    x->bytes = x->rhs_elems * nrhs * prec_bytes_complex(prec);
    const size_t sb = prec == WD_F16 ? sizeof(float) * 2 * ctx->stride * nrhs : 0;
    bool ok;
    if (work) {
        x->data = wd_work_get(ctx, x->bytes);
        x->scale = sb ? (float *)wd_work_get(ctx, sb) : nullptr;
        ok = x->data && (!sb || x->scale);
    } else {
        ok = wd_malloc(ctx, &x->data, x->bytes) == hipSuccess &&
             (!sb || wd_malloc(ctx, (void **)&x->scale, sb) == hipSuccess);
    }
    if (!ok) {
        wd_field_destroy(x);
        return wd_set_error(ctx, WD_ENOMEM, "field allocation failed");
    }
    // zeroed in stream order; a user field also waits for it (its caller may
    // touch it from the host next), a solver's work field does not
    int rc = wd_launch_zero(ctx, x->data, x->bytes);
    if (!rc && x->scale)
        rc = wd_launch_zero(ctx, x->scale, sb);
    if (!rc && !work && hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = wd_set_error(ctx, WD_EHIP, "field zeroing failed");
    if (rc) {
        wd_field_destroy(x);
        return rc;
    }
    *f = x;
    return WD_OK;
}

extern "C" int wd_field_create(wd_ctx *ctx, wd_field **f, int prec, int nrhs)
{
    return field_create(ctx, f, prec, nrhs, false);
}

int wd_field_create_work(wd_ctx *ctx, wd_field **f, int prec, int nrhs)
{
    return field_create(ctx, f, prec, nrhs, true);
}

extern "C" void wd_field_destroy(wd_field *f)
{
    if (!f)
        return;
    wd_ctx *ctx = f->ctx;
    if (f->work) {
        // stream order keeps the next user of the blocks behind this field's work
        wd_work_put(ctx, f->data, f->bytes);
        wd_work_put(ctx, f->scale, sizeof(float) * 2 * ctx->stride * f->nrhs);
    } else {
        (void)hipStreamSynchronize(ctx->stream);
        if (f->data)
            (void)hipFree(f->data);
        if (f->scale)
            (void)hipFree(f->scale);
    }
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
    rc = wd_launch_spinor_from_ref(ctx, f, rhs, (const double2 *)ctx->staging);
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
    rc = wd_launch_spinor_to_ref(ctx, f, rhs, (double2 *)ctx->staging);
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
        WD_ARG_CHECK(ctx, U->prec == WD_F64, "operator: F64 gauge required");
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
    rc = wd_launch_hop_f64(ctx, o, (const double2 *)in->data, f64_links(U), dagger,
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
                           f64_links(U), dagger, -1, HOP_M, 4.0 + mass,
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
                           f64_links(U), 0, -1, HOP_M, 4.0 + mass,
                           (const double2 *)in->data, nullptr, nullptr, nullptr);
    if (!rc)
        rc = wd_launch_hop_f64(ctx, (double2 *)out->data, (const double2 *)tmp->data,
                               f64_links(U), 1, -1, HOP_M, 4.0 + mass,
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
                                   f64_links(U), dagger, cb, HOP_ONLY, 0.0, nullptr,
                                   nullptr, nullptr, nullptr);
        else
            rc = wd_launch_hop_f64(ctx, (double2 *)out->data, (const double2 *)in->data,
                                   f64_links(U), dagger, -1, HOP_M, 4.0 + mass,
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

extern "C" int wd_get_option(wd_ctx *ctx, int option, int *value)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, value != nullptr, "get_option: null value");
    switch (option) {
    case WD_OPT_XCD_REMAP:
        *value = ctx->xcd_remap;
        return WD_OK;
    case WD_OPT_PARITY_ORDER:
        *value = ctx->parity_il;
        return WD_OK;
    case WD_OPT_BLOCK_TILE:
        *value = ctx->geom.tW;
        return WD_OK;
    case WD_OPT_TSTREAM:
        *value = ctx->ts_tb;
        return WD_OK;
    case WD_OPT_F64_PAIRS:
        *value = ctx->f64_pairs;
        return WD_OK;
    case WD_OPT_LOOP_PREFETCH:
        *value = ctx->loop_prefetch;
        return WD_OK;
    case WD_OPT_LOOP_SYSTEMS:
        *value = ctx->loop_systems;
        return WD_OK;
    default:
        return wd_set_error(ctx, WD_EARG, "get_option: unknown option");
    }
}

extern "C" int wd_set_option(wd_ctx *ctx, int option, int value)
{
    if (!ctx)
        return WD_EARG;
    switch (option) {
    case WD_OPT_XCD_REMAP:
        ctx->xcd_remap = value ? 1 : 0;
        return WD_OK;
    case WD_OPT_PARITY_ORDER:
        ctx->parity_il = value ? 1 : 0;
        return WD_OK;
    case WD_OPT_TSTREAM:
        WD_ARG_CHECK(ctx, value >= 0, "set_option: t-run length must be >= 0");
        if (value >= 2)
            WD_ARG_CHECK(ctx, ctx->L[3] % value == 0, "set_option: Lt must be a multiple of the t-run length");
        ctx->ts_tb = value == 0 ? 0 : value == 1 ? wd_auto_tstream(ctx) : value;
        return WD_OK;
    case WD_OPT_F64_PAIRS:
        ctx->f64_pairs = value ? 1 : 0;
        return WD_OK;
    case WD_OPT_LOOP_PREFETCH:
        WD_ARG_CHECK(ctx, value == 0 || value == 1 || value == 3,
                     "set_option: loop prefetch must be 0, 1 or 3");
        ctx->loop_prefetch = value;
        return WD_OK;
    case WD_OPT_LOOP_SYSTEMS:
        WD_ARG_CHECK(ctx, value >= 0, "set_option: systems per workgroup must be >= 0");
        ctx->loop_systems = value;
        return WD_OK;
    case WD_OPT_BLOCK_TILE:
        WD_ARG_CHECK(ctx, value >= 0, "set_option: block tile width must be >= 0");
        // tiling needs block-aligned time slices; a width >= D is the natural order
        ctx->geom.tW = (ctx->geom.tD > 0 && value < ctx->geom.tD) ? value : 0;
        return WD_OK;
    default:
        return wd_set_error(ctx, WD_EARG, "set_option: unknown option");
    }
}
