// wd_hop_sloppy.hip -- FP32-arithmetic Wilson operator on many right-hand
// sides: spinors stored F16 (with a per-site scale) or F32, links F32 / F16,
// 12- or 18-real.
//
// Work decomposition: a workgroup owns kBlock consecutive sites of one parity
// and one group of R right-hand sides (R = rhs per thread, a template
// parameter chosen by the autotuner).  Each thread loads every link of its
// site once into registers and applies it to its R right-hand sides, so link
// traffic is shared R ways in registers; the G = ceil(nrhs/R) groups of one
// site range are consecutive logical workgroups and land on the same XCD, so
// the remaining link re-reads hit that XCD's L2.  Right-hand sides whose
// `active` flag is 0 (converged systems) are skipped per workgroup (uniform
// branch).  Per-rhs diagonal coefficients a_r = 4 + m_r allow several masses
// in one launch.
//
// Modes as for the FP64 kernel:
//   HOP_ONLY  out = -1/2 hop(in)
//   HOP_M     out = a_r in - 1/2 hop(in)
//   HOP_M_DOT HOP_M plus per-(rhs, workgroup) FP64 partials of sum conj(dotv) out
#include "wd_fields.h"
#include "wd_internal.h"
#include "wd_sloppy.h"

using namespace wdf;

namespace {

template <int R>
__device__ __forceinline__ void pin_accs(cpx<float> (&acc)[R][4][3])
{
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                asm volatile("" : "+v"(acc[r][s][c].r), "+v"(acc[r][s][c].i));
}

template <int SF, int GF, int R, int MU, bool FWD, bool DAG>
__device__ __forceinline__ void dir_rhs(cpx<float> (&acc)[R][4][3], const SloppyArgs &a,
                                        const bool (&act)[R], const int (&rhs)[R],
                                        const char *Ubase, int p, int i, int j, float sgn,
                                        int stride)
{
    cpx<float> W[3][3];
    // forward: U_mu(x) at (parity p, site i); backward: U_mu(x - mu) at (1 - p, j)
    const size_t blk = (size_t)gauge_chunks(GF) * gauge_chunk_bytes(GF) * stride;
    const int lpar = FWD ? p : 1 - p;
    // one group: each link is read once, stream it past L2; several groups:
    // the next groups of this site range re-read it from L2 (uniform branch)
    if (a.groups > 1)
        load_link<GF, false>(Ubase + (size_t)(MU * 2 + lpar) * blk, stride, (unsigned)(FWD ? i : j),
                             W);
    else
        load_link<GF, true>(Ubase + (size_t)(MU * 2 + lpar) * blk, stride, (unsigned)(FWD ? i : j),
                            W);
    using D = Dir<MU, FWD, DAG>;
    const size_t pbytes = (size_t)spinor_chunks(SF) * 16 * stride;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!act[r])
            continue;
        const char *in_r = a.in + (size_t)rhs[r] * a.rhs_bytes + (size_t)(1 - p) * pbytes;
        const float *sc = a.in_scale ? a.in_scale + (size_t)rhs[r] * 2 * stride +
                                           (size_t)(1 - p) * stride
                                     : nullptr;
        cpx<float> psi[4][3], h[2][3], res[2][3];
        Spinor<SF>::load(in_r, sc, stride, (unsigned)j, psi);
        project<D>(psi, h);
        su3_mul_half<!FWD>(W, h, res);
        reconstruct<D>(acc[r], res, sgn);
    }
    pin_accs<R>(acc);
}

// Block-wide sum of (sr, si) stored by thread 0 to dst; safe to call repeatedly.
__device__ __forceinline__ void block_sum_store(double sr, double si, double2 *dst)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    __shared__ double lr[kBlock / 64], li[kBlock / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lr[w] = sr;
        li[w] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, ti = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) {
            tr += lr[k];
            ti += li[k];
        }
        *dst = make_double2(tr, ti);
    }
    __syncthreads();
}

template <int SF, int GF, int R, bool DAG, int MODE>
__global__ __launch_bounds__(kBlock) void k_hop_sloppy(SloppyArgs a, Geom g, int cb_first,
                                                       int nparity, int remap)
{
    if (a.stop && *a.stop)
        return;
    const int lb = logical_block(remap);
    const int grp = lb % a.groups;
    const int sb = lb / a.groups;
    // both parities alternate over the same site range (shared L2 reuse of in[p])
    const int p = (nparity == 2) ? (sb & 1) : cb_first;
    const int i = ((nparity == 2) ? (sb >> 1) : sb) * kBlock + threadIdx.x;
    const int stride = g.stride;
    bool act[R];
    int rhs[R];
    bool any = false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        rhs[r] = grp * R + r;
        act[r] = rhs[r] < a.nrhs && (a.active == nullptr || a.active[rhs[r]] != 0);
        any = any || act[r];
    }
    if (!any)
        return; // uniform over the workgroup
    const bool site_ok = i < g.Vh;

    cpx<float> acc[R][4][3];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                acc[r][s][c] = {0.f, 0.f};

    if (site_ok) {
        const int Lx = 2 * g.Lxh;
        int rr = i / g.Lxh;
        const int xh = i - rr * g.Lxh;
        const int y = rr % g.Ly;
        rr /= g.Ly;
        const int z = rr % g.Lz;
        const int t = rr / g.Lz;
        const int x = 2 * xh + ((p + y + z + t) & 1);
        const int row = i - xh;
        const float bx = (float)g.bc[0], by = (float)g.bc[1], bz = (float)g.bc[2],
                    bt = (float)g.bc[3];
        const char *U = a.U;
        {
            const int xf = (x + 1 == Lx) ? 0 : x + 1, xb = (x == 0) ? Lx - 1 : x - 1;
            dir_rhs<SF, GF, R, 0, true, DAG>(acc, a, act, rhs, U, p, i, row + (xf >> 1),
                                             x == Lx - 1 ? bx : 1.f, stride);
            dir_rhs<SF, GF, R, 0, false, DAG>(acc, a, act, rhs, U, p, i, row + (xb >> 1),
                                              x == 0 ? bx : 1.f, stride);
        }
        {
            const int s = g.Lxh;
            const int jf = (y + 1 == g.Ly) ? i - (g.Ly - 1) * s : i + s;
            const int jb = (y == 0) ? i + (g.Ly - 1) * s : i - s;
            dir_rhs<SF, GF, R, 1, true, DAG>(acc, a, act, rhs, U, p, i, jf,
                                             y == g.Ly - 1 ? by : 1.f, stride);
            dir_rhs<SF, GF, R, 1, false, DAG>(acc, a, act, rhs, U, p, i, jb, y == 0 ? by : 1.f,
                                              stride);
        }
        {
            const int s = g.Lxh * g.Ly;
            const int jf = (z + 1 == g.Lz) ? i - (g.Lz - 1) * s : i + s;
            const int jb = (z == 0) ? i + (g.Lz - 1) * s : i - s;
            dir_rhs<SF, GF, R, 2, true, DAG>(acc, a, act, rhs, U, p, i, jf,
                                             z == g.Lz - 1 ? bz : 1.f, stride);
            dir_rhs<SF, GF, R, 2, false, DAG>(acc, a, act, rhs, U, p, i, jb, z == 0 ? bz : 1.f,
                                              stride);
        }
        {
            const int s = g.Lxh * g.Ly * g.Lz;
            const int jf = (t + 1 == g.Lt) ? i - (g.Lt - 1) * s : i + s;
            const int jb = (t == 0) ? i + (g.Lt - 1) * s : i - s;
            dir_rhs<SF, GF, R, 3, true, DAG>(acc, a, act, rhs, U, p, i, jf,
                                             t == g.Lt - 1 ? bt : 1.f, stride);
            dir_rhs<SF, GF, R, 3, false, DAG>(acc, a, act, rhs, U, p, i, jb, t == 0 ? bt : 1.f,
                                              stride);
        }
    }

    const size_t pbytes = (size_t)spinor_chunks(SF) * 16 * stride;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!act[r])
            continue;
        double pr = 0.0, pi = 0.0;
        if (site_ok) {
            const size_t off = (size_t)rhs[r] * a.rhs_bytes + (size_t)p * pbytes;
            const size_t soff = (size_t)rhs[r] * 2 * stride + (size_t)p * stride;
            cpx<float> res[4][3];
            if constexpr (MODE == HOP_ONLY) {
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        res[s][c] = {-0.5f * acc[r][s][c].r, -0.5f * acc[r][s][c].i};
            } else {
                cpx<float> v[4][3];
                Spinor<SF>::load(a.same + off, a.same_scale ? a.same_scale + soff : nullptr,
                                 stride, (unsigned)i, v);
                const float ar = a.diag[rhs[r]];
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        res[s][c] = {fmaf(ar, v[s][c].r, -0.5f * acc[r][s][c].r),
                                     fmaf(ar, v[s][c].i, -0.5f * acc[r][s][c].i)};
            }
            Spinor<SF>::store(a.out + off, a.out_scale ? a.out_scale + soff : nullptr, stride,
                              (unsigned)i, res);
            if constexpr (MODE == HOP_M_DOT) {
                // the partial uses the stored (rounded) values, as a later
                // reduction over the field would
                Spinor<SF>::load(a.out + off, a.out_scale ? a.out_scale + soff : nullptr,
                                 stride, (unsigned)i, res);
                cpx<float> d[4][3];
                Spinor<SF>::load(a.dotv + off, a.dotv_scale ? a.dotv_scale + soff : nullptr,
                                 stride, (unsigned)i, d);
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double dr = d[s][c].r, di = d[s][c].i;
                        const double qr = res[s][c].r, qi = res[s][c].i;
                        pr = fma(dr, qr, pr);
                        pr = fma(di, qi, pr);
                        pi = fma(dr, qi, pi);
                        pi = fma(-di, qr, pi);
                    }
            }
        }
        if constexpr (MODE == HOP_M_DOT)
            block_sum_store(pr, pi, a.partials + (size_t)rhs[r] * a.nsb + sb);
    }
}

template <int SF, int GF, int R, bool DAG>
void launch_r(wd_ctx *ctx, const SloppyArgs &a, int cb_first, int nparity, int mode, int grid)
{
    const Geom &g = ctx->geom;
    switch (mode) {
    case HOP_ONLY:
        hipLaunchKernelGGL((k_hop_sloppy<SF, GF, R, DAG, HOP_ONLY>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, a, g, cb_first, nparity, ctx->xcd_remap);
        break;
    case HOP_M:
        hipLaunchKernelGGL((k_hop_sloppy<SF, GF, R, DAG, HOP_M>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, a, g, cb_first, nparity, ctx->xcd_remap);
        break;
    default:
        hipLaunchKernelGGL((k_hop_sloppy<SF, GF, R, DAG, HOP_M_DOT>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, a, g, cb_first, nparity, ctx->xcd_remap);
        break;
    }
}

template <int SF, int GF, bool DAG>
void launch_gf(wd_ctx *ctx, const SloppyArgs &a, int R, int cb_first, int nparity, int mode,
               int grid)
{
    switch (R) {
    case 1:
        launch_r<SF, GF, 1, DAG>(ctx, a, cb_first, nparity, mode, grid);
        break;
    case 2:
        launch_r<SF, GF, 2, DAG>(ctx, a, cb_first, nparity, mode, grid);
        break;
    default:
        launch_r<SF, GF, 4, DAG>(ctx, a, cb_first, nparity, mode, grid);
        break;
    }
}

template <int SF, bool DAG>
void launch_sf(wd_ctx *ctx, const SloppyArgs &a, int gf, int R, int cb_first, int nparity,
               int mode, int grid)
{
    switch (gf) {
    case G_F32_18:
        launch_gf<SF, G_F32_18, DAG>(ctx, a, R, cb_first, nparity, mode, grid);
        break;
    case G_F32_12:
        launch_gf<SF, G_F32_12, DAG>(ctx, a, R, cb_first, nparity, mode, grid);
        break;
    default:
        launch_gf<SF, G_F16_12, DAG>(ctx, a, R, cb_first, nparity, mode, grid);
        break;
    }
}

} // namespace

bool wd_sloppy_supported(int spinor_prec, int gauge_prec, int recon)
{
    const bool s = spinor_prec == WD_F32 || spinor_prec == WD_F16;
    const bool g = (gauge_prec == WD_F32) || (gauge_prec == WD_F16 && recon == WD_RECON_12);
    return s && g;
}

int wd_sloppy_groups(int nrhs, int R) { return (nrhs + R - 1) / R; }

int wd_launch_hop_sloppy(wd_ctx *ctx, SloppyArgs a, int spinor_prec, int gauge_prec, int recon,
                         int R, int dagger, int cb, int mode)
{
    const Geom &g = ctx->geom;
    WD_ARG_CHECK(ctx, R == 1 || R == 2 || R == 4, "sloppy hop: rhs per thread must be 1, 2 or 4");
    WD_ARG_CHECK(ctx, wd_sloppy_supported(spinor_prec, gauge_prec, recon),
                 "sloppy hop: unsupported precision combination");
    const int nparity = (cb < 0) ? 2 : 1;
    a.bpp = (g.Vh + kBlock - 1) / kBlock;
    a.nsb = a.bpp * nparity;
    a.groups = wd_sloppy_groups(a.nrhs, R);
    const long grid = (long)a.nsb * a.groups;
    WD_ARG_CHECK(ctx, grid < (1L << 31), "sloppy hop: grid too large");
    const int cb_first = (cb < 0) ? 0 : cb;
    const int gf = gauge_fmt(gauge_prec, recon);
    if (spinor_prec == WD_F16) {
        if (dagger)
            launch_sf<WD_F16, true>(ctx, a, gf, R, cb_first, nparity, mode, (int)grid);
        else
            launch_sf<WD_F16, false>(ctx, a, gf, R, cb_first, nparity, mode, (int)grid);
    } else {
        if (dagger)
            launch_sf<WD_F32, true>(ctx, a, gf, R, cb_first, nparity, mode, (int)grid);
        else
            launch_sf<WD_F32, false>(ctx, a, gf, R, cb_first, nparity, mode, (int)grid);
    }
    return wd_check_launch(ctx, "k_hop_sloppy");
}
