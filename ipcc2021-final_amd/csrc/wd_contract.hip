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

#include <vector>

#include "wd_fields.h"
#include "wd_internal.h"

using namespace wdf;

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

} // namespace

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
