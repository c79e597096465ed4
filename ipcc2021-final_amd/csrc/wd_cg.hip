// wd_cg.hip -- FP64 CGNR on M^dagger M, the reference's CGinvert
// (src/invert.cpp:25-121) as a device-resident loop.
//
// One iteration is five launches, all reading their scalars from device memory
// so the host never waits inside the loop:
//   1 k_cg_xpay      x += alpha p (the previous iteration's update, deferred to
//                    here where p is read anyway), then p = r + beta p
//                    (alpha = beta = 0 and p = 0 on the first iteration)
//   2 hop  M+norm    qq = M p, partial |qq|^2 per workgroup (HOP_M_NORM)
//   3 k_cg_alpha     alpha = rho / |M p|^2     ((p, M^dag M p) of the reference,
//                    src/invert.cpp:72-77, in the form whose value is real)
//   4 hop  M^dag+axpy r -= alpha M^dag qq, partial |r|^2 (HOP_M_AXPY; the
//                    reference's q = M^dag M p is never stored)
//   5 k_cg_beta      |r|, history, stop test |r| < acc, beta = |r|^2 / rho, rho = |r|^2
// rho = (r, r) is carried from step 6 to the next iteration's step 4; the
// reference recomputes the same sum with vector_p (src/invert.cpp:68, :82).
// Deferring x += alpha p changes no value (same alpha, same p) and saves one
// pass over p per iteration; k_cg_xfinal applies the last pending update after
// the loop.  Once the stop flag is set every later launch returns immediately,
// so the host launches `check_every` iterations (optionally one captured
// hipGraph) between reads of the flag.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// Streaming accesses: every vector is touched once per kernel and is far
// larger than the caches, so loads and stores are non-temporal.
__device__ __forceinline__ d2v ldv(const double2 *p, size_t i)
{
    return __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p + i));
}
__device__ __forceinline__ void stv(double2 *p, size_t i, d2v v)
{
    __builtin_nontemporal_store(v, reinterpret_cast<d2v *>(p + i));
}

// v + c w for complex c = (cr, ci)
__device__ __forceinline__ d2v caxpy(d2v v, double cr, double ci, d2v w)
{
    d2v o;
    o.x = v.x + (cr * w.x - ci * w.y);
    o.y = v.y + (cr * w.y + ci * w.x);
    return o;
}

// x += alpha p ; p = r + beta p.  Two elements per trip, all loads issued
// before any use, to keep enough bytes in flight per wavefront.
__global__ __launch_bounds__(kBlock) void k_cg_xpay(double2 *__restrict__ x,
                                                    double2 *__restrict__ p,
                                                    const double2 *__restrict__ r,
                                                    const CGState *__restrict__ st, size_t n)
{
    if (st->stop)
        return;
    const double ar = st->alpha.x, ai = st->alpha.y;
    const double br = st->beta.x, bi = st->beta.y;
    const size_t step = (size_t)gridDim.x * kBlock;
    size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
    for (; i + step < n; i += 2 * step) {
        const d2v p0 = ldv(p, i), p1 = ldv(p, i + step);
        const d2v x0 = ldv(x, i), x1 = ldv(x, i + step);
        const d2v r0 = ldv(r, i), r1 = ldv(r, i + step);
        stv(x, i, caxpy(x0, ar, ai, p0));
        stv(x, i + step, caxpy(x1, ar, ai, p1));
        stv(p, i, caxpy(r0, br, bi, p0));
        stv(p, i + step, caxpy(r1, br, bi, p1));
    }
    if (i < n) {
        const d2v p0 = ldv(p, i), x0 = ldv(x, i), r0 = ldv(r, i);
        stv(x, i, caxpy(x0, ar, ai, p0));
        stv(p, i, caxpy(r0, br, bi, p0));
    }
}

// x += alpha p once after the loop: the update of the last iteration (alpha
// is 0 when that iteration broke down before computing it)
__global__ __launch_bounds__(kBlock) void k_cg_xfinal(double2 *__restrict__ x,
                                                      const double2 *__restrict__ p,
                                                      const CGState *__restrict__ st, size_t n)
{
    const double ar = st->alpha.x, ai = st->alpha.y;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBlock)
        stv(x, i, caxpy(ldv(x, i), ar, ai, ldv(p, i)));
}

// alpha = rho / (p, M^dag M p) = rho / |M p|^2 (partials of |M p|^2);
// breakdown when |M p|^2 is not finite or not positive
__global__ __launch_bounds__(1024) void k_cg_alpha(const double2 *__restrict__ partials, int count,
                                                   CGState *__restrict__ st)
{
    if (st->stop)
        return;
    const double2 pq = sum_partials_1024(partials, count);
    if (threadIdx.x == 0) {
        if (!isfinite(pq.x) || !(pq.x > 0.0)) {
            st->alpha = make_double2(0.0, 0.0); // no x update for this iteration
            st->stop = 2;
            return;
        }
        st->alpha = make_double2(st->rho / pq.x, 0.0);
    }
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

// a work vector of one solve, taken from and handed back to the context's
// work cache
struct Buf {
    wd_ctx *ctx = nullptr;
    double2 *p = nullptr;
    size_t n = 0;
    ~Buf()
    {
        if (p)
            wd_work_put(ctx, p, sizeof(double2) * n);
    }
    int alloc(wd_ctx *c, size_t elems)
    {
        ctx = c;
        n = elems;
        p = (double2 *)wd_work_get(c, sizeof(double2) * elems);
        if (!p)
            return wd_set_error(c, WD_ENOMEM, "cg: work vector allocation failed");
        return wd_launch_zero(c, p, sizeof(double2) * elems);
    }
};

// Runs `iteration` (the launches of one CG iteration, each a no-op once
// st->stop is set) until the device sets the stop flag: `chunk` iterations
// between host reads of the flag, replayed from one captured hipGraph when
// use_graph.  Returns the final state in *hs and the iterations launched.
template <class F>
int run_cg_loop(wd_ctx *ctx, CGState *st, F &&iteration, int chunk, int use_graph, int max_iter,
                CGState *hs, int *launched_out)
{
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    if (use_graph) {
        WD_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int e = 0;
        for (int k = 0; k < chunk && !e; ++k)
            e = iteration();
        hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
        if (e) {
            if (graph)
                (void)hipGraphDestroy(graph);
            return e;
        }
        WD_HIP_CHECK(ctx, ce);
        ce = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return wd_set_error(ctx, WD_EHIP,
                                std::string("cg: graph instantiate: ") + hipGetErrorString(ce));
        }
    }
    int launched = 0;
    int status = WD_OK;
    while (true) {
        // stop flag set before the loop (max_iter <= 1)
        hipError_t ce = hipMemcpyAsync(ctx->hpinned, st, sizeof(CGState), hipMemcpyDeviceToHost,
                                       ctx->stream);
        if (ce == hipSuccess)
            ce = hipStreamSynchronize(ctx->stream);
        if (ce != hipSuccess) {
            status = wd_set_error(ctx, WD_EHIP, std::string("cg: poll: ") + hipGetErrorString(ce));
            break;
        }
        *hs = *(CGState *)ctx->hpinned;
        if (hs->stop || launched >= max_iter - 1)
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
                status = iteration();
            if (status)
                break;
        }
        launched += chunk;
    }
    if (gexec)
        (void)hipGraphExecDestroy(gexec);
    if (graph)
        (void)hipGraphDestroy(graph);
    *launched_out = launched;
    return status;
}

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
    WD_ARG_CHECK(ctx, U->prec == WD_F64, "cg: F64 gauge required");
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
    const F64Links Ud = f64_links(U);
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
    // the streaming kernel (2 elements per thread and trip): one trip per
    // thread for xpay.  Measured on MI355X: ~81 % of HBM peak from about 16k
    // workgroups up, ~68 % at 1k.
    const size_t per_blk = (size_t)kBlock * 2;
    const int sg = (int)std::max<size_t>(1, std::min<size_t>((n + per_blk - 1) / per_blk, 65536));
    const int *stop = &st->stop;

    auto launch_iteration = [&]() -> int {
        int e;
        hipLaunchKernelGGL(k_cg_xpay, dim3(sg), dim3(kBlock), 0, ctx->stream, xd, p.p, r.p, st,
                           n);
        if ((e = wd_check_launch(ctx, "k_cg_xpay")))
            return e;
        if ((e = wd_launch_hop_f64(ctx, qq.p, p.p, Ud, 0, -1, HOP_M_NORM, a, p.p, nullptr,
                                   partials, stop)))
            return e;
        hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1024), 0, ctx->stream, partials, hop_blocks,
                           st);
        if ((e = wd_check_launch(ctx, "k_cg_alpha")) ||
            (e = wd_launch_hop_f64_axpy(ctx, r.p, qq.p, Ud, 1, -1, a, qq.p, &st->alpha.x,
                                        partials, stop)))
            return e;
        hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1024), 0, ctx->stream, partials, hop_blocks,
                           st, d_hist);
        return wd_check_launch(ctx, "k_cg_beta");
    };

    CGState hs = {};
    int launched = 0;
    int status = run_cg_loop(ctx, st, launch_iteration, params->check_every, params->use_graph,
                             max_iter, &hs, &launched);
    if (status)
        return status;
    hipLaunchKernelGGL(k_cg_xfinal, dim3(eg), dim3(kBlock), 0, ctx->stream, xd, p.p, st, n);
    if ((rc = wd_check_launch(ctx, "k_cg_xfinal")))
        return rc;
    // operator applies executed in the loop: 2 per completed iteration
    applies += 2L * hs.iter;

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
    if (params->rsd_history && hs.iter > 0) {
        WD_HIP_CHECK(ctx, hipMemcpyAsync(params->rsd_history, d_hist, sizeof(double) * hs.iter,
                                         hipMemcpyDeviceToHost, ctx->stream));
        WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
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

// ---- Schur even-odd preconditioned CGNR ------------------------------------------
// On (even, odd) sites M = [[a, -D_eo/2], [-D_oe/2, a]] with a = 4 + m.  With
// x_o = (b_o + D_oe x_e / 2) / a, the even half solves the Schur system
//     A x_e = bh,   A = a^2 - D_eo D_oe / 4,   bh = a b_e + D_eo b_o / 2
// (A is a times the Schur complement a - D_eo D_oe / (4a); problem statement
// p.11-12, which the reference describes but does not implement).  CGNR runs
// on A^dag A, A^dag = a^2 - D^dag_eo D^dag_oe / 4, with the kernels of wd_cg on
// the even halves of the vectors; one application of A is two one-parity hops,
// the cost of one full-lattice M.  The stop test of the reference, the FP64
// |M^dag b - M^dag M x| of the full system, is checked after reconstructing
// x_o; when the Schur residual threshold (scaled from the ratio of the two
// right-hand-side norms) was not tight enough, CG continues from x_e with a
// proportionally tighter threshold.
extern "C" int wd_cg_eo(wd_ctx *ctx, wd_field *x, const wd_field *b, const wd_gauge *U,
                        double mass, const wd_cg_params *params, wd_cg_report *report)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, x && b && U && params, "cg_eo: null argument");
    WD_ARG_CHECK(ctx, x != b, "cg_eo: x and b must be distinct fields");
    WD_ARG_CHECK(ctx, x->prec == WD_F64 && b->prec == WD_F64 && x->nrhs == 1 && b->nrhs == 1,
                 "cg_eo: single-rhs F64 fields required");
    WD_ARG_CHECK(ctx, U->prec == WD_F64, "cg_eo: F64 gauge required");
    WD_ARG_CHECK(ctx, params->max_iter >= 1, "cg_eo: max_iter < 1");
    WD_ARG_CHECK(ctx, params->check_every >= 1, "cg_eo: check_every < 1");
    WD_ARG_CHECK(ctx, params->tol >= 0.0, "cg_eo: negative tolerance");
    WD_ARG_CHECK(ctx, 4.0 + mass != 0.0, "cg_eo: 4 + mass must not vanish");

    const auto t0 = std::chrono::steady_clock::now();
    const size_t n = x->rhs_elems;  // full vector (both parities)
    const size_t ne = n / 2;        // even half
    const double a = 4.0 + mass, a2 = a * a;
    int rc;
    Buf r, p, q, qq, t, Mdb, bh, rs, hist;
    if ((rc = r.alloc(ctx, n)) || (rc = p.alloc(ctx, ne)) || (rc = q.alloc(ctx, n)) ||
        (rc = qq.alloc(ctx, n)) || (rc = t.alloc(ctx, n)) || (rc = Mdb.alloc(ctx, n)) ||
        (rc = bh.alloc(ctx, ne)) || (rc = rs.alloc(ctx, ne)))
        return rc;
    const int max_iter = params->max_iter;
    if ((rc = hist.alloc(ctx, (size_t)(max_iter + 1) / 2 + 1)))
        return rc;
    double *d_hist = (double *)hist.p;
    CGState *st = (CGState *)ctx->dscalars;
    double2 *d_scalar = (double2 *)((char *)ctx->dscalars + 256);
    double2 *partials = (double2 *)ctx->partials;
    const F64Links Ud = f64_links(U);
    double2 *xd = (double2 *)x->data;
    const double2 *bd = (const double2 *)b->data;
    long applies = 0;

    auto read_norm = [&](const double2 *v, size_t len, double *out) -> int {
        int e = wd_launch_dot(ctx, v, nullptr, len, d_scalar);
        if (e)
            return e;
        WD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->hpinned, d_scalar, sizeof(double2),
                                         hipMemcpyDeviceToHost, ctx->stream));
        WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        *out = std::sqrt(((double2 *)ctx->hpinned)->x);
        return WD_OK;
    };
    // out_e = A in_e (dag 0) or A^dag in_e; HOP_M_DOT adds the (dotv, out) partials
    auto apply_A = [&](int dag, const double2 *in, double2 *out, int mode, const double2 *dotv,
                       const int *stop) -> int {
        int e = wd_launch_hop_f64c(ctx, t.p, in, Ud, dag, 1, HOP_ONLY, 0.0, 0.5, nullptr,
                                   nullptr, nullptr, stop);
        if (e)
            return e;
        return wd_launch_hop_f64c(ctx, out, t.p, Ud, dag, 0, mode, a2, -0.5, in, dotv,
                                  mode == HOP_ONLY || mode == HOP_M ? nullptr : partials, stop);
    };
    // |M^dag b - M^dag M x| of the full system
    auto full_residual = [&](double *out) -> int {
        int e;
        if ((e = wd_launch_hop_f64(ctx, qq.p, xd, Ud, 0, -1, HOP_M, a, xd, nullptr, nullptr,
                                   nullptr)) ||
            (e = wd_launch_hop_f64(ctx, q.p, qq.p, Ud, 1, -1, HOP_M, a, qq.p, nullptr, nullptr,
                                   nullptr)) ||
            (e = wd_launch_sub(ctx, r.p, Mdb.p, q.p, n)))
            return e;
        applies += 2;
        return read_norm(r.p, n, out);
    };

    // full-system right-hand side M^dag b (the stop test) and its norm
    if ((rc = wd_launch_hop_f64(ctx, Mdb.p, bd, Ud, 1, -1, HOP_M, a, bd, nullptr, nullptr,
                                nullptr)))
        return rc;
    ++applies;
    double rhs_norm = 0.0;
    if ((rc = read_norm(Mdb.p, n, &rhs_norm)))
        return rc;
    const double target = params->tol_relative ? params->tol * rhs_norm : params->tol;
    if (!(rhs_norm > 0.0)) {
        // M^dag b = 0: x = 0 is the solution (b = 0 for an invertible M)
        if ((rc = wd_launch_zero(ctx, xd, sizeof(double2) * n)))
            return rc;
        WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        if (report) {
            *report = wd_cg_report{};
            report->converged = 1;
            report->operator_applies = applies;
            report->seconds =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return WD_OK;
    }
    // bh = a b_e + D_eo b_o / 2 ; rs = A^dag bh
    if ((rc = wd_launch_hop_f64c(ctx, bh.p, bd, Ud, 0, 0, HOP_M, a, 0.5, bd, nullptr, nullptr,
                                 nullptr)) ||
        (rc = apply_A(1, bh.p, rs.p, HOP_M, nullptr, nullptr)))
        return rc;
    ++applies;
    double schur_norm = 0.0;
    if ((rc = read_norm(rs.p, ne, &schur_norm)))
        return rc;
    if (!params->use_x0 && (rc = wd_launch_zero(ctx, xd, sizeof(double2) * n)))
        return rc;

    const int hop_blocks = wd_hop_grid(ctx, 1);
    const size_t per_blk = (size_t)kBlock * 2;
    const int sg = (int)std::max<size_t>(1, std::min<size_t>((ne + per_blk - 1) / per_blk, 65536));
    const int eg = wd_elem_grid(ne), rg = wd_red_grid(ne);
    const int *stop = &st->stop;
    double *hist_at = d_hist;

    auto launch_iteration = [&]() -> int {
        int e;
        hipLaunchKernelGGL(k_cg_xpay, dim3(sg), dim3(kBlock), 0, ctx->stream, xd, p.p, r.p, st,
                           ne);
        if ((e = wd_check_launch(ctx, "k_cg_xpay")))
            return e;
        // qq = A p with |A p|^2; alpha; r -= alpha A^dag qq with |r|^2 (the
        // second A^dag launch updates r instead of storing A^dag A p)
        if ((e = apply_A(0, p.p, qq.p, HOP_M_NORM, nullptr, stop)))
            return e;
        hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1024), 0, ctx->stream, partials, hop_blocks,
                           st);
        if ((e = wd_check_launch(ctx, "k_cg_alpha")) ||
            (e = wd_launch_hop_f64c(ctx, t.p, qq.p, Ud, 1, 1, HOP_ONLY, 0.0, 0.5, nullptr, nullptr,
                                    nullptr, stop)) ||
            (e = wd_launch_hop_f64_axpy(ctx, r.p, t.p, Ud, 1, 0, a2, qq.p, &st->alpha.x, partials,
                                        stop)))
            return e;
        hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1024), 0, ctx->stream, partials, hop_blocks,
                           st, hist_at);
        return wd_check_launch(ctx, "k_cg_beta");
    };

    // Schur threshold from the ratio of the right-hand sides, halved for margin
    double acc = rhs_norm > 0.0 ? 0.5 * target * schur_norm / rhs_norm : 0.0;
    int total = 0;
    int last_stop = 0;
    double schur_rsd = schur_norm, true_rsd = 0.0;
    bool first = true;
    for (int round = 0; round < 4; ++round) {
        // r = rs - A^dag A x_e (r = rs when x_e = 0)
        if (first && !params->use_x0) {
            WD_HIP_CHECK(ctx, hipMemcpyAsync(r.p, rs.p, sizeof(double2) * ne,
                                             hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            if ((rc = apply_A(0, xd, qq.p, HOP_M, nullptr, nullptr)) ||
                (rc = apply_A(1, qq.p, q.p, HOP_M, nullptr, nullptr)) ||
                (rc = wd_launch_sub(ctx, r.p, rs.p, q.p, ne)))
                return rc;
            applies += 2;
        }
        first = false;
        if ((rc = wd_launch_dot(ctx, r.p, nullptr, ne, d_scalar + 1)))
            return rc;
        hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1024), 0, ctx->stream, partials, rg, st, acc,
                           max_iter - total);
        if ((rc = wd_check_launch(ctx, "k_cg_init")))
            return rc;
        hist_at = d_hist + total;
        CGState hs = {};
        int launched = 0;
        if ((rc = run_cg_loop(ctx, st, launch_iteration, params->check_every, params->use_graph,
                              max_iter - total, &hs, &launched)))
            return rc;
        hipLaunchKernelGGL(k_cg_xfinal, dim3(eg), dim3(kBlock), 0, ctx->stream, xd, p.p, st, ne);
        if ((rc = wd_check_launch(ctx, "k_cg_xfinal")))
            return rc;
        total += hs.iter;
        applies += 2L * hs.iter;
        schur_rsd = hs.rsd;
        last_stop = hs.stop;
        // x_o = b_o / a + D_oe x_e / (2a), then the full-system residual
        if ((rc = wd_launch_hop_f64c(ctx, xd, xd, Ud, 0, 1, HOP_M, 1.0 / a, 0.5 / a, bd, nullptr,
                                     nullptr, nullptr)) ||
            (rc = full_residual(&true_rsd)))
            return rc;
        if (true_rsd <= target || last_stop != 1 || total >= max_iter - 1 || !(true_rsd > 0.0))
            break;
        acc *= 0.5 * target / true_rsd;
    }
    if (params->rsd_history && total > 0) {
        WD_HIP_CHECK(ctx, hipMemcpyAsync(params->rsd_history, d_hist, sizeof(double) * total,
                                         hipMemcpyDeviceToHost, ctx->stream));
        WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    const auto t1 = std::chrono::steady_clock::now();
    const bool converged = true_rsd <= target && last_stop != 2;
    if (report) {
        report->iterations = total;
        report->converged = converged;
        report->rsd = schur_rsd;
        report->true_rsd = true_rsd;
        report->rhs_norm = rhs_norm;
        report->seconds = std::chrono::duration<double>(t1 - t0).count();
        report->operator_applies = applies;
    }
    if (last_stop == 2)
        return wd_set_error(ctx, WD_EBREAKDOWN, "cg_eo: breakdown ((p, A p) not positive/finite)");
    return converged ? WD_OK : WD_ENOTCONV;
}
