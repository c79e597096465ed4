// wd_cg.hip -- FP64 CGNR on M^dagger M, the reference's CGinvert
// (src/invert.cpp:25-121) as a device-resident loop.
//
// One iteration is six launches, all reading their scalars from device memory
// so the host never waits inside the loop:
//   1 k_cg_xpay      p = r + beta p            (beta = 0 on the first iteration:
//                                               p starts at 0, so p = r exactly)
//   2 hop  M         qq = M p                  (fused (4+m) diagonal + hop)
//   3 hop  M^dag+dot q = M^dag qq, partial (p, q) per workgroup
//   4 k_cg_alpha     alpha = rho / (p, q)      (complex, as the reference)
//   5 k_cg_update    x += alpha p, r -= alpha q, partial |r|^2
//   6 k_cg_beta      |r|, history, stop test |r| < acc, beta = |r|^2 / rho, rho = |r|^2
// rho = (r, r) is carried from step 6 to the next iteration's step 4; the
// reference recomputes the same sum with vector_p (src/invert.cpp:68, :82).
// Once the stop flag is set every later launch returns immediately, so the
// host launches `check_every` iterations (optionally one captured hipGraph)
// between reads of the flag.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <vector>

#include "wd_blas.h"
#include "wd_internal.h"

namespace {

struct CGState {
    double2 alpha;
    double2 beta;
    double rho;    // |r|^2 of the current residual
    double acc;    // stop threshold on |r|
    double rsd;    // |r| after the last iteration
    int iter;      // iterations completed
    int stop;      // 0 running, 1 converged, 2 breakdown, 3 iteration limit
    int max_iter;  // the reference's `max`: iterations f = 1 .. max-1
    int pad;
};

typedef double d2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ d2v ldv(const double2 *p, size_t i)
{
    return *reinterpret_cast<const d2v *>(p + i);
}
__device__ __forceinline__ void stv(double2 *p, size_t i, d2v v)
{
    *reinterpret_cast<d2v *>(p + i) = v;
}

// p = r + beta p (complex beta)
__global__ __launch_bounds__(kBlock) void k_cg_xpay(double2 *__restrict__ p,
                                                    const double2 *__restrict__ r,
                                                    const CGState *__restrict__ st, size_t n)
{
    if (st->stop)
        return;
    const double br = st->beta.x, bi = st->beta.y;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBlock) {
        const d2v pv = ldv(p, i), rv = ldv(r, i);
        d2v o;
        o.x = rv.x + (br * pv.x - bi * pv.y);
        o.y = rv.y + (br * pv.y + bi * pv.x);
        stv(p, i, o);
    }
}

// alpha = rho / (p, q); breakdown when (p, q) is not finite or its real part <= 0
__global__ __launch_bounds__(1024) void k_cg_alpha(const double2 *__restrict__ partials, int count,
                                                   CGState *__restrict__ st)
{
    if (st->stop)
        return;
    const double2 pq = sum_partials_1024(partials, count);
    if (threadIdx.x == 0) {
        if (!isfinite(pq.x) || !isfinite(pq.y) || !(pq.x > 0.0)) {
            st->stop = 2;
            return;
        }
        const double den = pq.x * pq.x + pq.y * pq.y;
        // (rho + 0i) / pq
        st->alpha = make_double2(st->rho * pq.x / den, -st->rho * pq.y / den);
    }
}

// x += alpha p ; r -= alpha q ; partial |r|^2
__global__ __launch_bounds__(kBlock) void k_cg_update(double2 *__restrict__ x,
                                                      double2 *__restrict__ r,
                                                      const double2 *__restrict__ p,
                                                      const double2 *__restrict__ q,
                                                      const CGState *__restrict__ st, size_t n,
                                                      double2 *__restrict__ partials)
{
    if (st->stop)
        return;
    const double ar = st->alpha.x, ai = st->alpha.y;
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBlock) {
        const d2v pv = ldv(p, i), qv = ldv(q, i);
        d2v xv = ldv(x, i), rv = ldv(r, i);
        xv.x += ar * pv.x - ai * pv.y;
        xv.y += ar * pv.y + ai * pv.x;
        rv.x -= ar * qv.x - ai * qv.y;
        rv.y -= ar * qv.y + ai * qv.x;
        stv(x, i, xv);
        stv(r, i, rv);
        s = fma(rv.x, rv.x, s);
        s = fma(rv.y, rv.y, s);
    }
    block_reduce_store(s, 0.0, partials + blockIdx.x);
}

__global__ __launch_bounds__(1024) void k_cg_beta(const double2 *__restrict__ partials, int count,
                                                  CGState *__restrict__ st,
                                                  double *__restrict__ hist)
{
    if (st->stop)
        return;
    const double2 s = sum_partials_1024(partials, count);
    if (threadIdx.x == 0) {
        const double rr = s.x;
        const double rsd = sqrt(rr);
        const int f = st->iter + 1;
        st->iter = f;
        st->rsd = rsd;
        hist[f - 1] = rsd;
        if (rsd < st->acc) {
            st->stop = 1;
        } else if (!isfinite(rr)) {
            st->stop = 2;
        } else if (f >= st->max_iter - 1) {
            st->stop = 3;
        } else {
            st->beta = make_double2(rr / st->rho, 0.0);
            st->rho = rr;
        }
    }
}

// rho = |r|^2 from the partials, beta = 0, stop flag cleared
__global__ __launch_bounds__(1024) void k_cg_init(const double2 *__restrict__ partials, int count,
                                                  CGState *__restrict__ st, double acc,
                                                  int max_iter)
{
    const double2 s = sum_partials_1024(partials, count);
    if (threadIdx.x == 0) {
        st->rho = s.x;
        st->beta = make_double2(0.0, 0.0);
        st->alpha = make_double2(0.0, 0.0);
        st->acc = acc;
        st->rsd = sqrt(s.x);
        st->iter = 0;
        st->max_iter = max_iter;
        st->stop = max_iter <= 1 ? 3 : 0;
    }
}

struct Buf {
    wd_ctx *ctx = nullptr;
    double2 *p = nullptr;
    size_t n = 0;
    ~Buf()
    {
        if (p)
            (void)hipFree(p);
    }
    int alloc(wd_ctx *c, size_t elems)
    {
        ctx = c;
        n = elems;
        if (hipMalloc((void **)&p, sizeof(double2) * elems) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return wd_set_error(c, WD_ENOMEM, "cg: work vector allocation failed");
        }
        return wd_launch_zero(c, p, sizeof(double2) * elems);
    }
};

} // namespace

extern "C" void wd_cg_params_default(wd_cg_params *p)
{
    if (!p)
        return;
    p->max_iter = 500;    // src/main.cpp:47
    p->tol = 1.0e-9;      // src/main.cpp:48
    p->tol_relative = 0;  // |r| < acc, src/invert.cpp:108
    p->use_x0 = 1;
    p->check_every = 8;
    p->use_graph = 1;
    p->rsd_history = nullptr;
}

extern "C" int wd_cg(wd_ctx *ctx, wd_field *x, const wd_field *b, const wd_gauge *U, double mass,
                     const wd_cg_params *params, wd_cg_report *report)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, x && b && U && params, "cg: null argument");
    WD_ARG_CHECK(ctx, x != b, "cg: x and b must be distinct fields");
    WD_ARG_CHECK(ctx, x->prec == WD_F64 && b->prec == WD_F64 && x->nrhs == 1 && b->nrhs == 1,
                 "cg: single-rhs F64 fields required");
    WD_ARG_CHECK(ctx, U->prec == WD_F64 && U->recon == WD_RECON_18,
                 "cg: F64 18-real gauge required");
    WD_ARG_CHECK(ctx, params->max_iter >= 1, "cg: max_iter < 1");
    WD_ARG_CHECK(ctx, params->check_every >= 1, "cg: check_every < 1");
    WD_ARG_CHECK(ctx, params->tol >= 0.0, "cg: negative tolerance");

    const auto t0 = std::chrono::steady_clock::now();
    const size_t n = x->rhs_elems;
    const double a = 4.0 + mass;
    int rc;
    Buf r, p, q, qq, Mdb, hist;
    if ((rc = r.alloc(ctx, n)) || (rc = p.alloc(ctx, n)) || (rc = q.alloc(ctx, n)) ||
        (rc = qq.alloc(ctx, n)) || (rc = Mdb.alloc(ctx, n)))
        return rc;
    const int max_iter = params->max_iter;
    if ((rc = hist.alloc(ctx, (size_t)(max_iter + 1) / 2 + 1)))
        return rc;
    double *d_hist = (double *)hist.p;
    CGState *st = (CGState *)ctx->dscalars;
    double2 *d_scalar = (double2 *)((char *)ctx->dscalars + 256);
    double2 *partials = (double2 *)ctx->partials;
    const double2 *Ud = (const double2 *)U->data;
    double2 *xd = (double2 *)x->data;
    const double2 *bd = (const double2 *)b->data;
    long applies = 0;

    // Mdb = M^dagger b (src/invert.cpp:41)
    if ((rc = wd_launch_hop_f64(ctx, Mdb.p, bd, Ud, 1, -1, HOP_M, a, bd, nullptr, nullptr,
                                nullptr)))
        return rc;
    ++applies;
    // |M^dagger b| for the relative stop test and the report
    if ((rc = wd_launch_dot(ctx, Mdb.p, nullptr, n, d_scalar)))
        return rc;
    WD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->hpinned, d_scalar, sizeof(double2),
                                     hipMemcpyDeviceToHost, ctx->stream));
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const double rhs_norm = std::sqrt(((double2 *)ctx->hpinned)->x);
    const double acc = params->tol_relative ? params->tol * rhs_norm : params->tol;

    // r = Mdb - M^dagger M x0 (src/invert.cpp:53-58); x0 = 0 gives r = Mdb
    if (!params->use_x0) {
        if ((rc = wd_launch_zero(ctx, xd, sizeof(double2) * n)))
            return rc;
        WD_HIP_CHECK(ctx, hipMemcpyAsync(r.p, Mdb.p, sizeof(double2) * n, hipMemcpyDeviceToDevice,
                                         ctx->stream));
    } else {
        if ((rc = wd_launch_hop_f64(ctx, qq.p, xd, Ud, 0, -1, HOP_M, a, xd, nullptr, nullptr,
                                    nullptr)) ||
            (rc = wd_launch_hop_f64(ctx, q.p, qq.p, Ud, 1, -1, HOP_M, a, qq.p, nullptr, nullptr,
                                    nullptr)) ||
            (rc = wd_launch_sub(ctx, r.p, Mdb.p, q.p, n)))
            return rc;
        applies += 2;
    }
    // rho = |r|^2, beta = 0, p = 0
    const int rg = wd_red_grid(n), eg = wd_elem_grid(n);
    {
        double2 *tmp_res = d_scalar + 1;
        if ((rc = wd_launch_dot(ctx, r.p, nullptr, n, tmp_res)))
            return rc;
        // the dot leaves its partials in ctx->partials; k_cg_init re-sums them
        hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1024), 0, ctx->stream, partials, rg, st, acc,
                           max_iter);
        if ((rc = wd_check_launch(ctx, "k_cg_init")))
            return rc;
    }
    const int hop_blocks = wd_hop_grid(ctx, 2);
    const int *stop = &st->stop;

    auto launch_iteration = [&]() -> int {
        int e;
        hipLaunchKernelGGL(k_cg_xpay, dim3(eg), dim3(kBlock), 0, ctx->stream, p.p, r.p, st, n);
        if ((e = wd_check_launch(ctx, "k_cg_xpay")))
            return e;
        if ((e = wd_launch_hop_f64(ctx, qq.p, p.p, Ud, 0, -1, HOP_M, a, p.p, nullptr, nullptr,
                                   stop)))
            return e;
        if ((e = wd_launch_hop_f64(ctx, q.p, qq.p, Ud, 1, -1, HOP_M_DOT, a, qq.p, p.p, partials,
                                   stop)))
            return e;
        hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1024), 0, ctx->stream, partials, hop_blocks,
                           st);
        if ((e = wd_check_launch(ctx, "k_cg_alpha")))
            return e;
        hipLaunchKernelGGL(k_cg_update, dim3(rg), dim3(kBlock), 0, ctx->stream, xd, r.p, p.p, q.p,
                           st, n, partials);
        if ((e = wd_check_launch(ctx, "k_cg_update")))
            return e;
        hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1024), 0, ctx->stream, partials, rg, st,
                           d_hist);
        return wd_check_launch(ctx, "k_cg_beta");
    };

    const int chunk = params->check_every;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    if (params->use_graph) {
        WD_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int e = 0;
        for (int k = 0; k < chunk && !e; ++k)
            e = launch_iteration();
        hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
        if (e)
            return e;
        WD_HIP_CHECK(ctx, ce);
        ce = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return wd_set_error(ctx, WD_EHIP,
                                std::string("cg: graph instantiate: ") + hipGetErrorString(ce));
        }
    }

    CGState hs = {};
    int launched = 0;
    int status = WD_OK;
    while (true) {
        // stop flag set before the loop (max_iter <= 1)
        WD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->hpinned, st, sizeof(CGState), hipMemcpyDeviceToHost,
                                         ctx->stream));
        WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        hs = *(CGState *)ctx->hpinned;
        if (hs.stop || launched >= max_iter - 1)
            break;
        if (gexec) {
            hipError_t ge = hipGraphLaunch(gexec, ctx->stream);
            if (ge != hipSuccess) {
                status = wd_set_error(ctx, WD_EHIP,
                                      std::string("cg: graph launch: ") + hipGetErrorString(ge));
                break;
            }
        } else {
            for (int k = 0; k < chunk && !status; ++k)
                status = launch_iteration();
            if (status)
                break;
        }
        launched += chunk;
        applies += 2L * chunk;
    }
    if (gexec)
        (void)hipGraphExecDestroy(gexec);
    if (graph)
        (void)hipGraphDestroy(graph);
    if (status)
        return status;
    // operator applies actually executed: 2 per completed iteration
    applies = applies - 2L * launched + 2L * hs.iter;

    // true residual |M^dagger b - M^dagger M x| in FP64
    if ((rc = wd_launch_hop_f64(ctx, qq.p, xd, Ud, 0, -1, HOP_M, a, xd, nullptr, nullptr,
                                nullptr)) ||
        (rc = wd_launch_hop_f64(ctx, q.p, qq.p, Ud, 1, -1, HOP_M, a, qq.p, nullptr, nullptr,
                                nullptr)) ||
        (rc = wd_launch_sub(ctx, r.p, Mdb.p, q.p, n)) ||
        (rc = wd_launch_dot(ctx, r.p, nullptr, n, d_scalar)))
        return rc;
    WD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->hpinned, d_scalar, sizeof(double2),
                                     hipMemcpyDeviceToHost, ctx->stream));
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const double true_rsd = std::sqrt(((double2 *)ctx->hpinned)->x);
    if (params->rsd_history && hs.iter > 0)
        WD_HIP_CHECK(ctx, hipMemcpy(params->rsd_history, d_hist, sizeof(double) * hs.iter,
                                    hipMemcpyDeviceToHost));
    const auto t1 = std::chrono::steady_clock::now();
    if (report) {
        report->iterations = hs.iter;
        report->converged = hs.stop == 1;
        report->rsd = hs.rsd;
        report->true_rsd = true_rsd;
        report->rhs_norm = rhs_norm;
        report->seconds = std::chrono::duration<double>(t1 - t0).count();
        report->operator_applies = applies + 2;
    }
    if (hs.stop == 2)
        return wd_set_error(ctx, WD_EBREAKDOWN, "cg: breakdown ((p, A p) not positive/finite)");
    return hs.stop == 1 ? WD_OK : WD_ENOTCONV;
}
