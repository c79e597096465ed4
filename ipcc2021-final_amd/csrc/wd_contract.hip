// wd_contract.hip -- meson contraction of point-source propagators.
//
// The pion two-point function from the 12 (spin, colour) point-source
// solutions S_r = M^-1 delta_r of one source site x0:
//     C(dt) = sum_r sum_{x : t(x) = t0 + dt mod Lt} sum_{s,c} |S_r(x)_{s c}|^2,
// the contraction of the reference's refs/python/contraction.py:25-45 for the
// pion, whose gamma5 insertions turn into |S|^2 by gamma5-hermiticity of M.
//
// One workgroup per (time slice, parity, system) sums |psi|^2 over the
// slice's Lxh*Ly*Lz sites of that parity block (a contiguous index range, so
// the loads are coalesced); the host adds the per-(system, parity) sums in a
// fixed order, so the result is deterministic.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "wd_fields.h"
#include "wd_internal.h"

using namespace wdf;

// ---- scratch of the allocation-free correlator ----------------------------------
// kCorrPartials per-(system, parity, slice) sums, then the kCorrMaxLt results
// of one call.  Every call on a GPU shares that GPU's copy of the array (calls
// that overlap on one GPU race) and writes every element it reads, so nothing
// needs clearing between calls.
constexpr int kCorrPartials = 16384;
constexpr int kCorrMaxLt = 512; // the results are staged through the context's 4 KiB pinned buffer
static __device__ __align__(256) double scratch[kCorrPartials + kCorrMaxLt];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch`, per device

} // namespace

// Looks up the device address of `scratch` on the current device, once per
// device.  In host code the array's name is a host-side stand-in; kernels and
// copies are given this address instead.  Called by wd_create with the
// context's device current.
int wd_contract_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return WD_EARG;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device])
        return WD_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return WD_EHIP;
    }
    g_scratch_ptr[device] = static_cast<double *>(p);
    return WD_OK;
}

namespace {

__global__ __launch_bounds__(kBlock) void k_slice_norm2(const double2 *__restrict__ S,
                                                        size_t rhs_elems, Geom g, int rhs0,
                                                        double *__restrict__ out)
{
    const int t = blockIdx.x, p = blockIdx.y, r = blockIdx.z;
    const int slice = g.Lxh * g.Ly * g.Lz;
    const double2 *base = S + (size_t)(rhs0 + r) * rhs_elems + (size_t)p * 12 * g.stride;
    double acc = 0.0;
    for (int k = threadIdx.x; k < slice; k += kBlock) {
        cpx<double> v[4][3];
        Spinor<WD_F64>::load(base, nullptr, g.stride, (unsigned)(t * slice + k), v);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                acc = fma(v[s][c].r, v[s][c].r, acc);
                acc = fma(v[s][c].i, v[s][c].i, acc);
            }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        acc += __shfl_xor(acc, off, 64);
    __shared__ double lds[kBlock / 64];
    if ((threadIdx.x & 63) == 0)
        lds[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < kBlock / 64; ++w)
            tot += lds[w];
        out[((size_t)r * 2 + p) * gridDim.x + t] = tot;
    }
}

// Ct[dt] = sum over systems r and parities p (in that order, from 0.0) of the
// slice sums of t = (t0 + dt) mod Lt: the host-side sum of wd_pion_correlator,
// same order, so the two entry points agree bit for bit.
__global__ void k_corr_sum(const double *__restrict__ part, int rhs_count, int Lt, int t0,
                           double *__restrict__ Ct)
{
    const int dt = blockIdx.x * blockDim.x + threadIdx.x;
    if (dt >= Lt)
        return;
    const int t = (t0 + dt) % Lt;
    double c = 0.0;
    for (int r = 0; r < rhs_count; ++r)
        for (int p = 0; p < 2; ++p)
            c += part[((size_t)r * 2 + p) * Lt + t];
    Ct[dt] = c;
}

} // namespace

extern "C" int wd_pion_correlator_noalloc(wd_ctx *ctx, const wd_field *S, int rhs_first,
                                          int rhs_count, int t0, double *Ct)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, S && Ct, "pion correlator: null argument");
    WD_ARG_CHECK(ctx, S->ctx == ctx, "pion correlator: field from another context");
    WD_ARG_CHECK(ctx, S->prec == WD_F64, "pion correlator: propagators must be F64");
    WD_ARG_CHECK(ctx, rhs_first >= 0 && rhs_count >= 1 && rhs_first + rhs_count <= S->nrhs,
                 "pion correlator: system range outside the field");
    const int Lt = ctx->L[3];
    WD_ARG_CHECK(ctx, t0 >= 0 && t0 < Lt, "pion correlator: source time outside the lattice");
    WD_ARG_CHECK(ctx, Lt <= kCorrMaxLt,
                 "pion correlator (no allocation): Lt above the limit of 512 time slices");
    WD_ARG_CHECK(ctx, (long)rhs_count * 2 * Lt <= kCorrPartials,
                 "pion correlator (no allocation): rhs_count * 2 * Lt above the scratch limit "
                 "of 16384 partial sums");
    double *part = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device]
                                                                   : nullptr;
    WD_ARG_CHECK(ctx, part != nullptr, "pion correlator (no allocation): scratch not set up");
    double *res = part + kCorrPartials;
    hipLaunchKernelGGL(k_slice_norm2, dim3(Lt, 2, rhs_count), dim3(kBlock), 0, ctx->stream,
                       (const double2 *)S->data, S->rhs_elems, ctx->geom, rhs_first, part);
    int rc = wd_check_launch(ctx, "k_slice_norm2");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_corr_sum, dim3((Lt + 63) / 64), dim3(64), 0, ctx->stream,
                       (const double *)part, rhs_count, Lt, t0, res);
    if ((rc = wd_check_launch(ctx, "k_corr_sum")))
        return rc;
    WD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->hpinned, res, sizeof(double) * Lt,
                                     hipMemcpyDeviceToHost, ctx->stream));
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const double *h = static_cast<const double *>(ctx->hpinned);
    for (int dt = 0; dt < Lt; ++dt)
        Ct[dt] = h[dt];
    return WD_OK;
}

extern "C" int wd_pion_correlator(wd_ctx *ctx, const wd_field *S, int rhs_first, int rhs_count,
                                  int t0, double *Ct)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, S && Ct, "pion correlator: null argument");
    WD_ARG_CHECK(ctx, S->prec == WD_F64, "pion correlator: propagators must be F64");
    WD_ARG_CHECK(ctx, rhs_first >= 0 && rhs_count >= 1 && rhs_first + rhs_count <= S->nrhs,
                 "pion correlator: system range outside the field");
    const int Lt = ctx->L[3];
    WD_ARG_CHECK(ctx, t0 >= 0 && t0 < Lt, "pion correlator: source time outside the lattice");
    const size_t n = (size_t)rhs_count * 2 * Lt;
    double *d = (double *)wd_work_get(ctx, sizeof(double) * n);
    if (!d)
        return wd_set_error(ctx, WD_ENOMEM, "pion correlator: allocation failed");
    hipLaunchKernelGGL(k_slice_norm2, dim3(Lt, 2, rhs_count), dim3(kBlock), 0, ctx->stream,
                       (const double2 *)S->data, S->rhs_elems, ctx->geom, rhs_first, d);
    int rc = wd_check_launch(ctx, "k_slice_norm2");
    std::vector<double> h(n);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(h.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost,
                                      ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess)
            rc = wd_set_error(ctx, WD_EHIP, std::string("pion correlator: ") + hipGetErrorString(e));
    }
    wd_work_put(ctx, d, sizeof(double) * n);
    if (rc)
        return rc;
    for (int dt = 0; dt < Lt; ++dt) {
        const int t = (t0 + dt) % Lt;
        double c = 0.0;
        for (int r = 0; r < rhs_count; ++r)
            for (int p = 0; p < 2; ++p)
                c += h[((size_t)r * 2 + p) * Lt + t];
        Ct[dt] = c;
    }
    return WD_OK;
}
