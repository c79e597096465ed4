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
// `active` flag is not 1 (converged systems, systems waiting for a reliable
// update) are skipped per workgroup (uniform branch).  Per-rhs diagonal coefficients a_r = 4 + m_r allow several masses
// in one launch.
//
// Instruction budget: with F16 spinors the kernel issues about as many VALU
// cycles as it moves bytes (measured ~70 % VALU issue on MI355X at one rhs),
// so the inner loop is written for instruction count: the half spinor is
// projected straight from the packed halves (one v_fma_mix_f32 per real, no
// conversions), the site scale, the boundary phase and the -1/2 ride in one
// factor of the 24 reconstruction FMAs, and loads use a uniform base plus a
// 32-bit lane offset (no 64-bit address arithmetic).
//
// Modes as for the FP64 kernel:
//   HOP_ONLY  out = -1/2 hop(in)
//   HOP_M     out = a_r in - 1/2 hop(in)
//   HOP_M_DOT HOP_M plus per-(rhs, workgroup) FP64 partials of sum conj(dotv) out
//             (per site in FP32 from the stored F16 values, whose products are
//             exact in FP32; FP64 across sites)
//   HOP_M_DOT_NORM  HOP_M_DOT plus the partials of |out|^2 (BiCGStab's |t|^2
//             from the launch that writes t; no dagger form)
//   HOP_M_NORM      HOP_M plus the partials (|out|^2, 0) (the CG's |A p|^2 from
//             the launch that writes A p; no dagger form)
//   HOP_M_AXPY      out = dotv - axpy_r (M in) with the partials (|out|^2, 0)
//             of the stored out (the CG's r -= alpha A^dag A p in the launch
//             that computes A^dag A p, which is never stored; out may be the
//             field dotv, each site is read and written by its own thread;
//             dagger form only)
#include "wd_fields.h"
#include "wd_internal.h"
#include "wd_sloppy.h"
#include "wd_sloppy_dev.h"

using namespace wdf;
using namespace wdsl;

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

// One hop direction for the R right-hand sides of this thread.  in_b / isc_b:
// per-rhs base of the neighbour parity (spinor data, F16 scales); hs = -1/2
// times the boundary phase of this hop; unit: the link is the identity (t
// links of a temporal-gauge field off the last slice), neither loaded nor
// applied.
template <int SF, int GF, int R, int MU, bool FWD, bool DAG>
__device__ __forceinline__ void dir_rhs(cpx<float> (&acc)[R][4][3], const SloppyArgs &a,
                                        const bool (&act)[R], const char *const (&in_b)[R],
                                        const float *const (&isc_b)[R], int p, int i, int j,
                                        float hs, int stride, bool unit = false)
{
    cpx<float> W[3][3];
    // forward: U_mu(x) at (parity p, site i); backward: U_mu(x - mu) at (1 - p, j)
    const size_t blk = (size_t)gauge_chunks(GF) * gauge_chunk_bytes(GF) * stride;
    const int lpar = FWD ? p : 1 - p;
    const char *Ub = a.U + (size_t)(MU * 2 + lpar) * blk;
    // links read once per launch stream past L2; temporal when the next
    // groups of this site range, or the other parity, re-read them (uniform)
    if (unit) {
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y < 3; ++y)
                W[x][y] = {0.f, 0.f};
    } else if (a.tlinks) {
        load_link32<GF, false>(Ub, stride, (unsigned)(FWD ? i : j), W);
    } else {
        load_link32<GF, true>(Ub, stride, (unsigned)(FWD ? i : j), W);
    }
    using D = Dir<MU, FWD, DAG>;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!act[r])
            continue;
        cpx<float> h[2][3], res[2][3];
        float f = hs;
        if constexpr (SF == WD_F16) {
            unsigned w[12];
            load16(in_b[r], stride, (unsigned)j, w);
            f *= load_scale(isc_b[r], (unsigned)j);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                h[0][c] = proj16<D::P0>(w[c], w[3 * D::S0 + c], a.one);
                h[1][c] = proj16<D::P1>(w[3 + c], w[3 * D::S1 + c], a.one);
            }
        } else {
            cpx<float> psi[4][3];
            load32(in_b[r], stride, (unsigned)j, psi);
            project<D>(psi, h);
        }
        if (unit) {
            reconstruct_f<D>(acc[r], h, f);
        } else {
            su3_mul_half<!FWD>(W, h, res);
            reconstruct_f<D>(acc[r], res, f);
        }
    }
    pin_accs<R>(acc);
}

template <int SF, int GF, int R, bool DAG, int MODE>
__global__ __launch_bounds__(kBlock) void k_hop_sloppy(SloppyArgs a, Geom g, int cb_first,
                                                       int nparity, int remap)
{
    if (a.stop && *a.stop)
        return;
    const int lb = logical_block(remap);
    // Work order.  chunk == 0: the groups of one site block are consecutive
    // (links shared in L2 at once).  chunk == C: C site blocks form a chunk and
    // the groups take turns over the whole chunk, so between two reads of a
    // system's t-neighbour only that group's spinors pass through L2 (the
    // links of the chunk are re-read per group, from L2/MALL).
    int grp, sb;
    if (a.chunk > 0 && a.groups > 1) {
        const int per = a.chunk * a.groups;
        const int c = lb / per;
        const int cl = min(a.chunk, a.nsb - c * a.chunk); // last chunk may be short
        const int rem = lb - c * per;
        grp = rem / cl;
        sb = c * a.chunk + (rem - grp * cl);
    } else {
        grp = lb % a.groups;
        sb = lb / a.groups;
    }
    // both parities alternate over the same site range (shared L2 reuse of in[p])
    const int p = (nparity == 2) ? (sb & 1) : cb_first;
    const int blk = block_order((nparity == 2) ? (sb >> 1) : sb, g);
    const int nat = (nparity == 2) ? 2 * blk + p : blk; // natural index (partials)
    const int i = blk * kBlock + threadIdx.x;
    const int stride = g.stride;
    bool act[R];
    int rhs[R];
    bool any = false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        rhs[r] = grp * R + r;
        act[r] = rhs[r] < a.nrhs && (a.active == nullptr || a.active[rhs[r]] == 1);
        any = any || act[r];
    }
    if (!any)
        return; // uniform over the workgroup
    const bool site_ok = i < g.Vh;

    // per-rhs bases (uniform): spinor data and F16 scales of the neighbour parity
    const size_t pbytes = (size_t)spinor_chunks(SF) * 16 * stride;
    const char *in_b[R];
    const float *isc_b[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        in_b[r] = a.in + (size_t)rhs[r] * a.rhs_bytes + (size_t)(1 - p) * pbytes;
        isc_b[r] = (SF == WD_F16) ? a.in_scale + (size_t)rhs[r] * 2 * stride +
                                        (size_t)(1 - p) * stride
                                  : nullptr;
    }

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
        // hop coefficient (-1/2 for M) times the boundary phase of a crossing hop
        const float h0 = a.hc;
        const float hx = h0 * (float)g.bc[0], hy = h0 * (float)g.bc[1],
                    hz = h0 * (float)g.bc[2], ht = h0 * (float)g.bc[3];
        {
            const int xf = (x + 1 == Lx) ? 0 : x + 1, xb = (x == 0) ? Lx - 1 : x - 1;
            dir_rhs<SF, GF, R, 0, true, DAG>(acc, a, act, in_b, isc_b, p, i, row + (xf >> 1),
                                             x == Lx - 1 ? hx : h0, stride);
            dir_rhs<SF, GF, R, 0, false, DAG>(acc, a, act, in_b, isc_b, p, i, row + (xb >> 1),
                                              x == 0 ? hx : h0, stride);
        }
        {
            const int s = g.Lxh;
            const int jf = (y + 1 == g.Ly) ? i - (g.Ly - 1) * s : i + s;
            const int jb = (y == 0) ? i + (g.Ly - 1) * s : i - s;
            dir_rhs<SF, GF, R, 1, true, DAG>(acc, a, act, in_b, isc_b, p, i, jf,
                                             y == g.Ly - 1 ? hy : h0, stride);
            dir_rhs<SF, GF, R, 1, false, DAG>(acc, a, act, in_b, isc_b, p, i, jb,
                                              y == 0 ? hy : h0, stride);
        }
        {
            const int s = g.Lxh * g.Ly;
            const int jf = (z + 1 == g.Lz) ? i - (g.Lz - 1) * s : i + s;
            const int jb = (z == 0) ? i + (g.Lz - 1) * s : i - s;
            dir_rhs<SF, GF, R, 2, true, DAG>(acc, a, act, in_b, isc_b, p, i, jf,
                                             z == g.Lz - 1 ? hz : h0, stride);
            dir_rhs<SF, GF, R, 2, false, DAG>(acc, a, act, in_b, isc_b, p, i, jb,
                                              z == 0 ? hz : h0, stride);
        }
        {
            const int s = g.Lxh * g.Ly * g.Lz;
            const int jf = (t + 1 == g.Lt) ? i - (g.Lt - 1) * s : i + s;
            const int jb = (t == 0) ? i + (g.Lt - 1) * s : i - s;
            // temporal gauge: the forward t link is the identity off the last
            // slice, the backward one (U_t of slice t - 1) off the first
            dir_rhs<SF, GF, R, 3, true, DAG>(acc, a, act, in_b, isc_b, p, i, jf,
                                             t == g.Lt - 1 ? ht : h0, stride,
                                             a.tg && t != g.Lt - 1);
            dir_rhs<SF, GF, R, 3, false, DAG>(acc, a, act, in_b, isc_b, p, i, jb,
                                              t == 0 ? ht : h0, stride, a.tg && t != 0);
        }
    }

#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!act[r])
            continue;
        double pr = 0.0, pi = 0.0, nn = 0.0;
        if (site_ok) {
            const size_t off = (size_t)rhs[r] * a.rhs_bytes + (size_t)p * pbytes;
            const size_t soff = (size_t)rhs[r] * 2 * stride + (size_t)p * stride;
            const unsigned ui = (unsigned)i;
            // acc already holds -1/2 hop(in); HOP_M adds a_r * in at the site
            if constexpr (MODE != HOP_ONLY) {
                const float ar = a.diag[rhs[r]];
                if constexpr (SF == WD_F16) {
                    unsigned v[12];
                    load16(a.same + off, stride, ui, v);
                    const float as = ar * load_scale(a.same_scale + soff, ui);
#pragma unroll
                    for (int k = 0; k < 12; ++k) {
                        acc[r][k / 3][k % 3].r = fmaf(re16(v[k]), as, acc[r][k / 3][k % 3].r);
                        acc[r][k / 3][k % 3].i = fmaf(im16(v[k]), as, acc[r][k / 3][k % 3].i);
                    }
                } else {
                    cpx<float> v[4][3];
                    load32(a.same + off, stride, ui, v);
#pragma unroll
                    for (int s = 0; s < 4; ++s)
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            acc[r][s][c] = {fmaf(ar, v[s][c].r, acc[r][s][c].r),
                                            fmaf(ar, v[s][c].i, acc[r][s][c].i)};
                }
            }
            // HOP_M_AXPY: out = dotv - axpy_r (M in), the elementwise update
            // fused into the launch that computes M in; F16: with M in as it
            // would be stored (the arithmetic of an update that reads it back)
            if constexpr (MODE == HOP_M_AXPY) {
                const float cr = a.axpy[rhs[r]];
                if constexpr (SF == WD_F16) {
                    unsigned q[12], v[12];
                    const float sq = pack16(acc[r], q);
                    load16(a.dotv + off, stride, ui, v);
                    const float sv = load_scale(a.dotv_scale + soff, ui);
#pragma unroll
                    for (int k = 0; k < 12; ++k)
                        acc[r][k / 3][k % 3] = {fmaf(-cr, re16(q[k]) * sq, re16(v[k]) * sv),
                                                fmaf(-cr, im16(q[k]) * sq, im16(v[k]) * sv)};
                } else {
                    cpx<float> v[4][3];
                    load32(a.dotv + off, stride, ui, v);
#pragma unroll
                    for (int s = 0; s < 4; ++s)
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            acc[r][s][c] = {fmaf(-cr, acc[r][s][c].r, v[s][c].r),
                                            fmaf(-cr, acc[r][s][c].i, v[s][c].i)};
                }
            }
            if constexpr (SF == WD_F16) {
                unsigned q[12];
                const float sq = store16(a.out + off, a.out_scale + soff, stride, ui, acc[r], q);
                if constexpr (MODE == HOP_M_DOT_NORM || MODE == HOP_M_NORM || MODE == HOP_M_AXPY) {
                    // |q|^2 of the stored values: f16 squares are exact in FP32
                    float nq = 0.f;
#pragma unroll
                    for (int k = 0; k < 12; ++k) {
                        nq = fmaf(re16(q[k]), re16(q[k]), nq);
                        nq = fmaf(im16(q[k]), im16(q[k]), nq);
                    }
                    nn = (double)nq * ((double)sq * (double)sq);
                }
                if constexpr (MODE == HOP_M_DOT || MODE == HOP_M_DOT_NORM) {
                    // sum conj(d) q over the stored (rounded) values: f16 x f16
                    // products are exact in FP32, scales applied once
                    unsigned d[12];
                    load16(a.dotv + off, stride, ui, d);
                    const float sd = load_scale(a.dotv_scale + soff, ui);
                    float tr = 0.f, ti = 0.f;
#pragma unroll
                    for (int k = 0; k < 12; ++k) {
                        tr = fmaf(re16(d[k]), re16(q[k]), tr);
                        tr = fmaf(im16(d[k]), im16(q[k]), tr);
                        ti = fmaf(re16(d[k]), im16(q[k]), ti);
                        ti = fmaf(-im16(d[k]), re16(q[k]), ti);
                    }
                    const double ss = (double)sd * (double)sq;
                    pr = (double)tr * ss;
                    pi = (double)ti * ss;
                }
            } else {
                store32(a.out + off, stride, ui, acc[r]);
                if constexpr (MODE == HOP_M_DOT_NORM || MODE == HOP_M_NORM || MODE == HOP_M_AXPY) {
#pragma unroll
                    for (int s = 0; s < 4; ++s)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const double qr = acc[r][s][c].r, qi = acc[r][s][c].i;
                            nn = fma(qr, qr, nn);
                            nn = fma(qi, qi, nn);
                        }
                }
                if constexpr (MODE == HOP_M_DOT || MODE == HOP_M_DOT_NORM) {
                    cpx<float> d[4][3];
                    load32(a.dotv + off, stride, ui, d);
#pragma unroll
                    for (int s = 0; s < 4; ++s)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const double dr = d[s][c].r, di = d[s][c].i;
                            const double qr = acc[r][s][c].r, qi = acc[r][s][c].i;
                            pr = fma(dr, qr, pr);
                            pr = fma(di, qi, pr);
                            pi = fma(dr, qi, pi);
                            pi = fma(-di, qr, pi);
                        }
                }
            }
        }
        if constexpr (MODE == HOP_M_DOT || MODE == HOP_M_DOT_NORM)
            block_sum_store(pr, pi, a.partials + (size_t)rhs[r] * a.nsb + nat);
        if constexpr (MODE == HOP_M_DOT_NORM)
            block_sum_store(nn, 0.0, a.partials2 + (size_t)rhs[r] * a.nsb + nat);
        if constexpr (MODE == HOP_M_NORM || MODE == HOP_M_AXPY)
            block_sum_store(nn, 0.0, a.partials + (size_t)rhs[r] * a.nsb + nat);
    }
}

template <int SF, int GF, int R, bool DAG>
void launch_r(wd_ctx *ctx, const SloppyArgs &a, const Geom &g, int cb_first, int nparity, int mode,
              int grid)
{
    switch (mode) {
    case HOP_ONLY:
        hipLaunchKernelGGL((k_hop_sloppy<SF, GF, R, DAG, HOP_ONLY>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, a, g, cb_first, nparity, ctx->xcd_remap);
        break;
    case HOP_M:
        hipLaunchKernelGGL((k_hop_sloppy<SF, GF, R, DAG, HOP_M>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, a, g, cb_first, nparity, ctx->xcd_remap);
        break;
    case HOP_M_DOT_NORM:
        // (s, t) and |t|^2 of BiCGStab's t = A s: no dagger form is needed
        if constexpr (!DAG)
            hipLaunchKernelGGL((k_hop_sloppy<SF, GF, R, false, HOP_M_DOT_NORM>), dim3(grid),
                               dim3(kBlock), 0, ctx->stream, a, g, cb_first, nparity,
                               ctx->xcd_remap);
        break;
    case HOP_M_NORM:
        // |A p|^2 of the CG's A p: no dagger form is needed
        if constexpr (!DAG)
            hipLaunchKernelGGL((k_hop_sloppy<SF, GF, R, false, HOP_M_NORM>), dim3(grid),
                               dim3(kBlock), 0, ctx->stream, a, g, cb_first, nparity,
                               ctx->xcd_remap);
        break;
    case HOP_M_AXPY:
        // the CG's r -= alpha A^dag (A p): the dagger form only
        if constexpr (DAG)
            hipLaunchKernelGGL((k_hop_sloppy<SF, GF, R, true, HOP_M_AXPY>), dim3(grid),
                               dim3(kBlock), 0, ctx->stream, a, g, cb_first, nparity,
                               ctx->xcd_remap);
        break;
    default:
        hipLaunchKernelGGL((k_hop_sloppy<SF, GF, R, DAG, HOP_M_DOT>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, a, g, cb_first, nparity, ctx->xcd_remap);
        break;
    }
}

template <int SF, int GF, bool DAG>
void launch_gf(wd_ctx *ctx, const SloppyArgs &a, const Geom &g, int R, int cb_first, int nparity,
               int mode, int grid)
{
    switch (R) {
    case 1:
        launch_r<SF, GF, 1, DAG>(ctx, a, g, cb_first, nparity, mode, grid);
        break;
    case 2:
        launch_r<SF, GF, 2, DAG>(ctx, a, g, cb_first, nparity, mode, grid);
        break;
    default:
        launch_r<SF, GF, 4, DAG>(ctx, a, g, cb_first, nparity, mode, grid);
        break;
    }
}

template <int SF, bool DAG>
void launch_sf(wd_ctx *ctx, const SloppyArgs &a, const Geom &g, int gf, int R, int cb_first,
               int nparity, int mode, int grid)
{
    switch (gf) {
    case G_F32_18:
        launch_gf<SF, G_F32_18, DAG>(ctx, a, g, R, cb_first, nparity, mode, grid);
        break;
    case G_F32_12:
        launch_gf<SF, G_F32_12, DAG>(ctx, a, g, R, cb_first, nparity, mode, grid);
        break;
    default:
        launch_gf<SF, G_F16_12, DAG>(ctx, a, g, R, cb_first, nparity, mode, grid);
        break;
    }
}

} // namespace

bool wd_sloppy_supported(int spinor_prec, int gauge_prec, int recon)
{
    const bool s = spinor_prec == WD_F32 || spinor_prec == WD_F16;
    const bool r12 = recon == WD_RECON_12 || recon == WD_RECON_12T;
    const bool g = (gauge_prec == WD_F32) || (gauge_prec == WD_F16 && r12);
    return s && g;
}

bool wd_sloppy_loop_supported(int spinor_prec, int gauge_prec, int recon)
{
    return spinor_prec == WD_F16 && (gauge_prec == WD_F16 || gauge_prec == WD_F32) &&
           (recon == WD_RECON_12 || recon == WD_RECON_12T);
}

int wd_sloppy_groups(int nrhs, int R) { return R == WD_RHS_LOOP ? 1 : (nrhs + R - 1) / R; }

int wd_launch_hop_sloppy(wd_ctx *ctx, SloppyArgs a, int spinor_prec, int gauge_prec, int recon,
                         int R, int dagger, int cb, int mode)
{
    // columns of the z-t tiled order twice as wide as the FP64 kernels' (the
    // FP32/F16 working set per site is smaller; measured faster on MI355X),
    // or the natural site order when the autotuner found it faster (a.tile < 0)
    Geom g = ctx->geom;
    if (a.tile < 0)
        g.tW = 0;
    else if (g.tW > 0)
        g.tW = 2 * g.tW < g.tD ? 2 * g.tW : 0;
    WD_ARG_CHECK(ctx, R == 1 || R == 2 || R == 4 || R == WD_RHS_LOOP,
                 "sloppy hop: rhs per thread must be 1, 2, 4 or WD_RHS_LOOP");
    WD_ARG_CHECK(ctx, R != WD_RHS_LOOP || wd_sloppy_loop_supported(spinor_prec, gauge_prec, recon),
                 "sloppy hop: WD_RHS_LOOP needs F16 spinors and F16 12-real links");
    WD_ARG_CHECK(ctx, a.chunk >= 0, "sloppy hop: site chunk must be >= 0");
    // every per-parity array (at most 6 x 16 bytes per site) must fit 32-bit offsets
    WD_ARG_CHECK(ctx, 96.0 * g.stride < 4294967296.0, "sloppy hop: lattice too large");
    WD_ARG_CHECK(ctx, a.hc != 0.f, "sloppy hop: hop coefficient not set");
    WD_ARG_CHECK(ctx, mode != HOP_M_DOT_NORM || (!dagger && a.partials2),
                 "sloppy hop: the norm mode has no dagger form and needs partials2");
    WD_ARG_CHECK(ctx, mode != HOP_M_NORM || (!dagger && a.partials),
                 "sloppy hop: the norm-only mode has no dagger form and needs partials");
    WD_ARG_CHECK(ctx, mode != HOP_M_AXPY || (dagger && a.partials && a.dotv && a.axpy),
                 "sloppy hop: the axpy mode is the dagger form only and needs partials, dotv "
                 "and axpy");
    a.one = 1.0f;
    a.tg = recon == WD_RECON_12T;
    WD_ARG_CHECK(ctx, wd_sloppy_supported(spinor_prec, gauge_prec, recon),
                 "sloppy hop: unsupported precision combination");
    const int nparity = (cb < 0) ? 2 : 1;
    a.bpp = (g.Vh + kBlock - 1) / kBlock;
    a.nsb = a.bpp * nparity;
    a.groups = wd_sloppy_groups(a.nrhs, R);
    a.tlinks = a.groups > 1 || (nparity == 2 && ctx->parity_il);
    const long grid = (long)a.nsb * a.groups;
    WD_ARG_CHECK(ctx, grid < (1L << 31), "sloppy hop: grid too large");
    if (R == WD_RHS_LOOP)
        return wd_launch_hop_sloppy_loop(ctx, a, g, dagger, cb, mode, gauge_prec);
    const int cb_first = (cb < 0) ? 0 : cb;
    const int gf = gauge_fmt(gauge_prec, recon);
    if (spinor_prec == WD_F16) {
        if (dagger)
            launch_sf<WD_F16, true>(ctx, a, g, gf, R, cb_first, nparity, mode, (int)grid);
        else
            launch_sf<WD_F16, false>(ctx, a, g, gf, R, cb_first, nparity, mode, (int)grid);
    } else {
        if (dagger)
            launch_sf<WD_F32, true>(ctx, a, g, gf, R, cb_first, nparity, mode, (int)grid);
        else
            launch_sf<WD_F32, false>(ctx, a, g, gf, R, cb_first, nparity, mode, (int)grid);
    }
    return wd_check_launch(ctx, "k_hop_sloppy");
}
