// wd_linkprod.hip -- the MFMA-vs-VALU experiment for the link product of the
// multi-system operator: C = U h for one SU(3) link U (3x3 complex, F16) and
// the half spinors h of N systems at the same site (3 colours x 2 spins per
// system, F16), accumulated in FP32 -- the "3x3 by 3x2N" product.
//
// Two kernels compute the same products from the same F16 inputs:
//   VALU: one site per lane, the arithmetic of the loop sloppy kernel (each
//         complex product two v_dot2_f32_f16 on the packed (re, im) link entry
//         and the operand forms (hr, -hi), (hi, hr) of h);
//   MFMA: v_mfma_f32_16x16x16f16 with two sites per tile.  Each site's link in
//         its real form [Ur -Ui; Ui Ur] (6x6) sits on the diagonal of the
//         16x16 A operand (12 of 16 rows and columns used); the B operand's 16
//         columns are 8 systems x 2 spins, its rows the (re, im) x colour parts
//         of the two sites' half spinors; one instruction forms 2 sites x 8
//         systems.  A site's link is the same for all its systems, so A is
//         built once per site pair; N > 8 takes ceil(N/8) column tiles.
// `inner` repeats the product on registers (each repetition accumulates, so
// none can be hoisted), which makes a launch compute-bound for inner >> 1;
// the operands are gathered from memory once per launch.
//
// This is a measurement, not a path of the solver: the operator keeps one
// site per lane, where the half spinors are formed, and an MFMA operand wants
// a site's colours spread over lanes.
#include "wd_internal.h"

namespace {

typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef _Float16 h4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ h2v as_h2(unsigned w) { return __builtin_bit_cast(h2v, w); }

// U: [site][a][b] packed (re, im); h: [site][system][spin][colour];
// out: [site][system][spin][colour] (re, im) as floats.
__global__ __launch_bounds__(kBlock) void k_linkprod_valu(const unsigned *__restrict__ U,
                                                          const unsigned *__restrict__ h,
                                                          float *__restrict__ out, int nsites,
                                                          int nsys, int inner)
{
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= nsites)
        return;
    unsigned W[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            W[a][b] = U[(size_t)s * 9 + a * 3 + b];
    for (int r = 0; r < nsys; ++r) {
        const unsigned *hs = h + ((size_t)s * nsys + r) * 6;
        // operand forms: re = dot(U, (hr, -hi)), im = dot(U, (hi, hr))
        h2v x[2][3], y[2][3];
#pragma unroll
        for (int sp = 0; sp < 2; ++sp)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const h2v v = as_h2(hs[sp * 3 + c]);
                x[sp][c] = h2v{v.x, -v.y};
                y[sp][c] = h2v{v.y, v.x};
            }
        float acc[3][4];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                acc[a][q] = 0.f;
        for (int k = 0; k < inner; ++k)
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const h2v u = as_h2(W[a][b]);
                    acc[a][0] = __builtin_amdgcn_fdot2(u, x[0][b], acc[a][0], false);
                    acc[a][1] = __builtin_amdgcn_fdot2(u, y[0][b], acc[a][1], false);
                    acc[a][2] = __builtin_amdgcn_fdot2(u, x[1][b], acc[a][2], false);
                    acc[a][3] = __builtin_amdgcn_fdot2(u, y[1][b], acc[a][3], false);
                }
        float *o = out + ((size_t)s * nsys + r) * 12;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o[2 * a] = acc[a][0];
            o[2 * a + 1] = acc[a][1];
            o[6 + 2 * a] = acc[a][2];
            o[6 + 2 * a + 1] = acc[a][3];
        }
    }
}

// Real-form row / column index i in [0, 12) of the two-site tile: site i / 6,
// part (i % 6) / 3 (0 = re, 1 = im), colour i % 3.
__global__ __launch_bounds__(kBlock) void k_linkprod_mfma(const unsigned *__restrict__ U,
                                                          const unsigned *__restrict__ h,
                                                          float *__restrict__ out, int nsites,
                                                          int nsys, int inner)
{
    const int lane = threadIdx.x & 63;
    const int pair = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (2 * pair >= nsites)
        return; // uniform over the wavefront
    const int s0 = 2 * pair;
    const int lo = lane & 15, kb = 4 * (lane >> 4);
    // A operand: lane holds row lo, columns kb .. kb + 3
    h4v A;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = kb + j;
        _Float16 v = 0;
        if (lo < 12 && k < 12 && lo / 6 == k / 6) {
            const int site = s0 + lo / 6, p = (lo % 6) / 3, a = lo % 3, q = (k % 6) / 3,
                      b = k % 3;
            const h2v u = as_h2(U[(size_t)site * 9 + a * 3 + b]);
            v = (p == q) ? u.x : (p == 0 ? (_Float16)(-u.y) : u.y);
        }
        A[j] = v;
    }
    for (int r0 = 0; r0 < nsys; r0 += 8) {
        // B operand: lane holds column lo = (system r0 + lo / 2, spin lo % 2),
        // rows kb .. kb + 3
        const int r = r0 + lo / 2, sp = lo & 1;
        h4v B;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = kb + j;
            _Float16 v = 0;
            if (k < 12 && r < nsys) {
                const int site = s0 + k / 6, q = (k % 6) / 3, b = k % 3;
                const h2v hv = as_h2(h[(((size_t)site * nsys + r) * 2 + sp) * 3 + b]);
                v = q ? hv.y : hv.x;
            }
            B[j] = v;
        }
        // four independent accumulation chains hide the instruction latency
        f4v D0 = {0.f, 0.f, 0.f, 0.f}, D1 = D0, D2 = D0, D3 = D0;
        int k = 0;
        for (; k + 4 <= inner; k += 4) {
            D0 = __builtin_amdgcn_mfma_f32_16x16x16f16(A, B, D0, 0, 0, 0);
            D1 = __builtin_amdgcn_mfma_f32_16x16x16f16(A, B, D1, 0, 0, 0);
            D2 = __builtin_amdgcn_mfma_f32_16x16x16f16(A, B, D2, 0, 0, 0);
            D3 = __builtin_amdgcn_mfma_f32_16x16x16f16(A, B, D3, 0, 0, 0);
        }
        for (; k < inner; ++k)
            D0 = __builtin_amdgcn_mfma_f32_16x16x16f16(A, B, D0, 0, 0, 0);
        const f4v D = (D0 + D1) + (D2 + D3);
        // D: lane holds column lo, rows kb .. kb + 3
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = kb + j;
            if (i < 12 && r < nsys) {
                const int site = s0 + i / 6, p = (i % 6) / 3, a = i % 3;
                out[(((size_t)site * nsys + r) * 2 + sp) * 6 + 2 * a + p] = D[j];
            }
        }
    }
}

} // namespace

extern "C" int wd_link_product(wd_ctx *ctx, int impl, int nsites, int nsys, const uint32_t *U,
                               const uint32_t *h, int inner, int iters, float *out,
                               double *ms_per_launch)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, impl == WD_LINKPROD_VALU || impl == WD_LINKPROD_MFMA,
                 "link_product: unknown implementation");
    WD_ARG_CHECK(ctx, nsites >= 2 && nsites % 2 == 0, "link_product: nsites must be even and >= 2");
    WD_ARG_CHECK(ctx, nsys >= 1 && nsys <= 64, "link_product: nsys must be in 1 .. 64");
    WD_ARG_CHECK(ctx, inner >= 1 && iters >= 1, "link_product: inner and iters must be >= 1");
    WD_ARG_CHECK(ctx, U && h && out && ms_per_launch, "link_product: null pointer");
    WD_ARG_CHECK(ctx, (size_t)nsites * nsys <= ((size_t)1 << 31) / 12,
                 "link_product: nsites * nsys too large");
    const size_t ub = (size_t)nsites * 9 * 4, hb = (size_t)nsites * nsys * 6 * 4,
                 ob = (size_t)nsites * nsys * 12 * 4;
    void *dU = wd_work_get(ctx, ub), *dh = wd_work_get(ctx, hb), *dout = wd_work_get(ctx, ob);
    int rc = WD_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!dU || !dh || !dout) {
        rc = wd_set_error(ctx, WD_ENOMEM, "link_product: out of device memory");
    } else if (hipMemcpyAsync(dU, U, ub, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
               hipMemcpyAsync(dh, h, hb, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
               hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = wd_set_error(ctx, WD_EHIP, "link_product: copy or event creation failed");
    } else {
        const int grid = impl == WD_LINKPROD_VALU ? (nsites + kBlock - 1) / kBlock
                                                  : (nsites / 2 + kBlock / 64 - 1) / (kBlock / 64);
        float ms = 0.f;
        for (int k = 0; k <= iters && !rc; ++k) {
            if (k == 1 && hipEventRecord(e0, ctx->stream) != hipSuccess)
                rc = wd_set_error(ctx, WD_EHIP, "link_product: event record failed");
            if (impl == WD_LINKPROD_VALU)
                hipLaunchKernelGGL(k_linkprod_valu, dim3(grid), dim3(kBlock), 0, ctx->stream,
                                   (const unsigned *)dU, (const unsigned *)dh, (float *)dout,
                                   nsites, nsys, inner);
            else
                hipLaunchKernelGGL(k_linkprod_mfma, dim3(grid), dim3(kBlock), 0, ctx->stream,
                                   (const unsigned *)dU, (const unsigned *)dh, (float *)dout,
                                   nsites, nsys, inner);
            if (!rc)
                rc = wd_check_launch(ctx, "k_linkprod");
        }
        if (!rc && (hipEventRecord(e1, ctx->stream) != hipSuccess ||
                    hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = wd_set_error(ctx, WD_EHIP, "link_product: timing failed");
        if (!rc && (hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                    hipStreamSynchronize(ctx->stream) != hipSuccess))
            rc = wd_set_error(ctx, WD_EHIP, "link_product: copy back failed");
        if (!rc)
            *ms_per_launch = (double)ms / iters;
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(ctx->stream);
    wd_work_put(ctx, dU, ub);
    wd_work_put(ctx, dh, hb);
    wd_work_put(ctx, dout, ob);
    return rc;
}
