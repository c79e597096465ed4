// wd_hop_f64.hip -- FP64 Wilson hopping term and fused M / M^dagger kernels.
//
// One thread per output site.  For each of the 8 neighbours the thread loads
// the full neighbour spinor (12 complex, coalesced across the wavefront), the
// link (9 complex, streamed with non-temporal loads because every link is
// read exactly once per half-lattice pass), spin-projects to a half spinor,
// multiplies by U or U^dagger and reconstructs into a register accumulator.
// The -1/2 of the hopping term (src/dslash.cpp:552 `* half`, with the minus of
// `tmp = -(...)`) is applied once in the epilogue; scaling by 0.5 is exact.
//
// Modes (template):
//   HOP_ONLY  out[p] = -1/2 hop(in)                       DslashoffdNew
//   HOP_M     out[p] = a * same[p] - 1/2 hop(in)          DslashNew, a = 4 + m
//   HOP_M_DOT HOP_M plus a per-workgroup partial of sum conj(dotv) * out, the
//             (p, q) product of the CG step fused into q = M^dagger (M p).
//   HOP_M_NORM HOP_M plus a per-workgroup partial (|out|^2, 0): the CG's
//             |M p|^2 = (p, M^dagger M p) from the launch that writes M p.
//   HOP_M_AXPY out = dotv - alpha (a * same - 1/2 hop(in)) (alpha real, read
//             from device memory; out may be dotv, each site read and written
//             by its own thread) with the partial (|out|^2, 0): the CG's
//             r -= alpha M^dagger M p without storing M^dagger M p.
#include "wd_blas.h"
#include "wd_dirac.h"
#include "wd_fields.h"
#include "wd_internal.h"

using wdf::f64_idx;

namespace {

typedef double d2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ cpx<double> ld(const double2 *__restrict__ p, size_t idx)
{
    const d2v v = *reinterpret_cast<const d2v *>(p + idx);
    return {v.x, v.y};
}

__device__ __forceinline__ cpx<double> ld_nt(const double2 *__restrict__ p, size_t idx)
{
    const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p + idx));
    return {v.x, v.y};
}

__device__ __forceinline__ void st(double2 *__restrict__ p, size_t idx, cpx<double> z)
{
    d2v v;
    v.x = z.r;
    v.y = z.i;
    // streamed out: non-temporal, so the output does not evict the input
    // spinors the neighbouring workgroups still read from L2
    __builtin_nontemporal_store(v, reinterpret_cast<d2v *>(p + idx));
}

__device__ __forceinline__ void load_spinor(const double2 *__restrict__ base, int stride, int j,
                                            cpx<double> (&psi)[4][3])
{
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            psi[s][c] = ld(base, f64_idx(s * 3 + c, (unsigned)j, 12));
}

// NT: non-temporal (each link read once per launch); temporal when the two
// parities of a site range run side by side and share links through L2.
// LN: complex numbers stored per link, 9 (18-real, the reference's format) or
// 6 (12-real: rows 0 and 1, row 2 rebuilt as conj(row0 x row1), SU(3) links).
template <int LN, bool NT>
__device__ __forceinline__ void load_link(const double2 *__restrict__ base, int stride, int j,
                                          cpx<double> (&U)[3][3])
{
#pragma unroll
    for (int k = 0; k < LN; ++k)
        U[k / 3][k % 3] = NT ? ld_nt(base, f64_idx(k, (unsigned)j, LN))
                             : ld(base, f64_idx(k, (unsigned)j, LN));
    if constexpr (LN == 6)
        reconstruct_row2(U);
}

// Forces the accumulator to be materialised here so the compiler cannot
// defer the arithmetic of a direction (and keep its loads live) past this point.
__device__ __forceinline__ void pin_acc(cpx<double> (&acc)[4][3])
{
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            asm volatile("" : "+v"(acc[s][c].r), "+v"(acc[s][c].i));
}

// One direction pair (forward and backward) of the hop.
template <int MU, bool DAG, bool NTL, int LN>
__device__ __forceinline__ void hop_mu(cpx<double> (&acc)[4][3], const double2 *__restrict__ in_o,
                                       const double2 *__restrict__ U, const Geom &g, int p, int i,
                                       int jf, int jb, double sf, double sb)
{
    const int stride = g.stride;
    {
        cpx<double> psi[4][3], W[3][3], h[2][3], r[2][3];
        load_spinor(in_o, stride, jf, psi);
        load_link<LN, NTL>(U + (long)((MU * 2 + p) * LN) * stride, stride, i, W);
        using D = Dir<MU, true, DAG>;
        project<D>(psi, h);
        su3_mul_half<false>(W, h, r);
        reconstruct<D>(acc, r, sf);
    }
    pin_acc(acc);
    {
        cpx<double> psi[4][3], W[3][3], h[2][3], r[2][3];
        load_spinor(in_o, stride, jb, psi);
        load_link<LN, NTL>(U + (long)((MU * 2 + (1 - p)) * LN) * stride, stride, jb, W);
        using D = Dir<MU, false, DAG>;
        project<D>(psi, h);
        su3_mul_half<true>(W, h, r);
        reconstruct<D>(acc, r, sb);
    }
    pin_acc(acc);
}

// IL (both parities only): the two parities of a site block are consecutive
// logical workgroups, so a link is read for its forward use by one and for
// its backward use by the other shortly after (temporal link loads); else the
// first half of the workgroups computes parity 0 and the second parity 1.
template <bool DAG, int MODE, bool IL, int LN>
// waves_per_eu(1, 1) lets the scheduler keep more of a direction's loads in
// flight (~150 VGPRs, still 3 waves per SIMD at run time); measured faster
// than the default budget on MI355X.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_hop_f64(double2 *__restrict__ out,
                                                    const double2 *__restrict__ in,
                                                    const double2 *__restrict__ U, Geom g,
                                                    int cb_first, int nparity, double a,
                                                    double hc,
                                                    const double2 *__restrict__ same,
                                                    const double2 *__restrict__ dotv,
                                                    double2 *__restrict__ partials, int remap,
                                                    const int *__restrict__ stop, F64Batch b)
{
    // solver loops launch ahead of their convergence check; a set stop flag
    // turns the remaining launches into no-ops
    if (stop && *stop)
        return;
    // right-hand side blockIdx.y of a batch: masked systems are skipped,
    // each may carry its own diagonal coefficient a_r = 4 + m_r
    {
        const int rhs = blockIdx.y;
        if (b.mask && !b.mask[rhs])
            return;
        if (b.a_rhs)
            a = b.a_rhs[rhs];
        if (b.c_rhs)
            hc = b.c_rhs[rhs];
        const size_t ro = (size_t)rhs * b.rhs_elems;
        out += ro;
        in += ro;
        if (same)
            same += ro;
        if (dotv)
            dotv += ro;
        if (partials)
            partials += (size_t)rhs * gridDim.x;
    }
    const int lb = logical_block(remap);
    const int bpp = gridDim.x / nparity;
    int p, blk;
    if (IL) {
        p = lb & 1;
        blk = block_order(lb >> 1, g);
    } else {
        p = (nparity == 2) ? lb / bpp : cb_first;
        blk = block_order(lb - (nparity == 2 ? p * bpp : 0), g);
    }
    const int i = blk * kBlock + threadIdx.x;
    const bool active = i < g.Vh;

    cpx<double> res[4][3];
    if (active) {
        // coordinates of site i of parity p
        const int Lx = 2 * g.Lxh;
        int rr = i / g.Lxh;
        const int xh = i - rr * g.Lxh;
        const int y = rr % g.Ly;
        rr /= g.Ly;
        const int z = rr % g.Lz;
        const int t = rr / g.Lz;
        const int x = 2 * xh + ((p + y + z + t) & 1);
        const int row = i - xh; // index of (xh = 0, y, z, t)

        const double2 *in_o = in + (long)(1 - p) * 12 * g.stride;
        cpx<double> acc[4][3];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                acc[s][c] = {0.0, 0.0};

        {
            const int xf = (x + 1 == Lx) ? 0 : x + 1, xb = (x == 0) ? Lx - 1 : x - 1;
            hop_mu<0, DAG, !IL, LN>(acc, in_o, U, g, p, i, row + (xf >> 1), row + (xb >> 1),
                           x == Lx - 1 ? g.bc[0] : 1.0, x == 0 ? g.bc[0] : 1.0);
        }
        {
            const int s = g.Lxh;
            const int jf = (y + 1 == g.Ly) ? i - (g.Ly - 1) * s : i + s;
            const int jb = (y == 0) ? i + (g.Ly - 1) * s : i - s;
            hop_mu<1, DAG, !IL, LN>(acc, in_o, U, g, p, i, jf, jb, y == g.Ly - 1 ? g.bc[1] : 1.0,
                           y == 0 ? g.bc[1] : 1.0);
        }
        {
            const int s = g.Lxh * g.Ly;
            const int jf = (z + 1 == g.Lz) ? i - (g.Lz - 1) * s : i + s;
            const int jb = (z == 0) ? i + (g.Lz - 1) * s : i - s;
            hop_mu<2, DAG, !IL, LN>(acc, in_o, U, g, p, i, jf, jb, z == g.Lz - 1 ? g.bc[2] : 1.0,
                           z == 0 ? g.bc[2] : 1.0);
        }
        {
            const int s = g.Lxh * g.Ly * g.Lz;
            const int jf = (t + 1 == g.Lt) ? i - (g.Lt - 1) * s : i + s;
            const int jb = (t == 0) ? i + (g.Lt - 1) * s : i - s;
            hop_mu<3, DAG, !IL, LN>(acc, in_o, U, g, p, i, jf, jb, t == g.Lt - 1 ? g.bc[3] : 1.0,
                           t == 0 ? g.bc[3] : 1.0);
        }

        double2 *out_p = out + (long)p * 12 * g.stride;
        if constexpr (MODE == HOP_ONLY) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    res[s][c] = {hc * acc[s][c].r, hc * acc[s][c].i};
                    st(out_p, f64_idx(s * 3 + c, (unsigned)i, 12), res[s][c]);
                }
        } else {
            const double2 *same_p = same + (long)p * 12 * g.stride;
            // HOP_M_AXPY: out = dotv - alpha res (res = M in is not stored)
            const double2 *dv_p = dotv + (long)p * 12 * g.stride;
            const double al = MODE == HOP_M_AXPY ? *b.axpy : 0.0;
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const cpx<double> v = ld(same_p, f64_idx(s * 3 + c, (unsigned)i, 12));
                    res[s][c] = {fma(a, v.r, hc * acc[s][c].r), fma(a, v.i, hc * acc[s][c].i)};
                    if constexpr (MODE == HOP_M_AXPY) {
                        const cpx<double> d = ld(dv_p, f64_idx(s * 3 + c, (unsigned)i, 12));
                        res[s][c] = {fma(-al, res[s][c].r, d.r), fma(-al, res[s][c].i, d.i)};
                    }
                    st(out_p, f64_idx(s * 3 + c, (unsigned)i, 12), res[s][c]);
                }
        }
    }

    if constexpr (MODE == HOP_M_NORM || MODE == HOP_M_AXPY) {
        // partial (|out|^2, 0) of the stored values for this workgroup
        double nn = 0.0;
        if (active) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    nn = fma(res[s][c].r, res[s][c].r, nn);
                    nn = fma(res[s][c].i, res[s][c].i, nn);
                }
        }
        // (partials already point at this system's row)
        block_reduce_store(nn, 0.0, partials + (nparity == 2 ? p * bpp : 0) + blk);
    }

    if constexpr (MODE == HOP_M_DOT) {
        // partial of sum conj(dotv) * out for this workgroup, fixed order
        double pr = 0.0, pi = 0.0;
        if (active) {
            const double2 *dv = dotv + (long)p * 12 * g.stride;
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const cpx<double> v = ld(dv, f64_idx(s * 3 + c, (unsigned)i, 12));
                    pr = fma(v.r, res[s][c].r, pr);
                    pr = fma(v.i, res[s][c].i, pr);
                    pi = fma(v.r, res[s][c].i, pi);
                    pi = fma(-v.i, res[s][c].r, pi);
                }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            pr += __shfl_xor(pr, off, 64);
            pi += __shfl_xor(pi, off, 64);
        }
        __shared__ double sr[kBlock / 64], si[kBlock / 64];
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            sr[w] = pr;
            si[w] = pi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double tr = 0.0, ti = 0.0;
            for (int k = 0; k < kBlock / 64; ++k) {
                tr += sr[k];
                ti += si[k];
            }
            // natural block order, so the sum does not depend on the work order
            partials[(nparity == 2 ? p * bpp : 0) + blk] = make_double2(tr, ti);
        }
    }
}


// One direction pair for two systems of a batch: each link is loaded once and
// applied to both; per system the operations are those of hop_mu, in the
// same order, so each system's result is bit-identical to k_hop_f64's.
// use0 / use1 are uniform over the workgroup.
template <int MU, bool DAG, bool NTL, int LN>
__device__ __forceinline__ void hop_mu2(cpx<double> (&acc0)[4][3], cpx<double> (&acc1)[4][3],
                                        bool use0, bool use1,
                                        const double2 *__restrict__ in0,
                                        const double2 *__restrict__ in1,
                                        const double2 *__restrict__ U, const Geom &g, int p,
                                        int i, int jf, int jb, double sf, double sb)
{
    const int stride = g.stride;
    {
        cpx<double> W[3][3];
        load_link<LN, NTL>(U + (long)((MU * 2 + p) * LN) * stride, stride, i, W);
        using D = Dir<MU, true, DAG>;
        if (use0) {
            cpx<double> psi[4][3], h[2][3], r[2][3];
            load_spinor(in0, stride, jf, psi);
            project<D>(psi, h);
            su3_mul_half<false>(W, h, r);
            reconstruct<D>(acc0, r, sf);
        }
        if (use1) {
            cpx<double> psi[4][3], h[2][3], r[2][3];
            load_spinor(in1, stride, jf, psi);
            project<D>(psi, h);
            su3_mul_half<false>(W, h, r);
            reconstruct<D>(acc1, r, sf);
        }
    }
    pin_acc(acc0);
    pin_acc(acc1);
    {
        cpx<double> W[3][3];
        load_link<LN, NTL>(U + (long)((MU * 2 + (1 - p)) * LN) * stride, stride, jb, W);
        using D = Dir<MU, false, DAG>;
        if (use0) {
            cpx<double> psi[4][3], h[2][3], r[2][3];
            load_spinor(in0, stride, jb, psi);
            project<D>(psi, h);
            su3_mul_half<true>(W, h, r);
            reconstruct<D>(acc0, r, sb);
        }
        if (use1) {
            cpx<double> psi[4][3], h[2][3], r[2][3];
            load_spinor(in1, stride, jb, psi);
            project<D>(psi, h);
            su3_mul_half<true>(W, h, r);
            reconstruct<D>(acc1, r, sb);
        }
    }
    pin_acc(acc0);
    pin_acc(acc1);
}

__device__ __forceinline__ void store_out(double2 *__restrict__ out_p,
                                          const double2 *__restrict__ same_p, int i, double a,
                                          double hc, const cpx<double> (&acc)[4][3], bool with_same)
{
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cpx<double> res;
            if (with_same) {
                const cpx<double> v = ld(same_p, f64_idx(s * 3 + c, (unsigned)i, 12));
                res = {fma(a, v.r, hc * acc[s][c].r), fma(a, v.i, hc * acc[s][c].i)};
            } else {
                res = {hc * acc[s][c].r, hc * acc[s][c].i};
            }
            st(out_p, f64_idx(s * 3 + c, (unsigned)i, 12), res);
        }
}

// Batched hop of two systems per thread (systems 2 blockIdx.y and
// 2 blockIdx.y + 1): the link traffic, three quarters of an FP64 hop's bytes,
// is paid once per pair.  HOP_ONLY and HOP_M; same work order, masks and
// per-system coefficients as k_hop_f64 with a batch.
template <bool DAG, int MODE, bool IL, int LN>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_hop_f64_pair(
    double2 *__restrict__ out, const double2 *__restrict__ in, const double2 *__restrict__ U,
    Geom g, int cb_first, int nparity, double a, double hc, const double2 *__restrict__ same,
    int remap, const int *__restrict__ stop, F64Batch b, int nrhs)
{
    static_assert(MODE == HOP_ONLY || MODE == HOP_M, "pairs: HOP_ONLY and HOP_M only");
    if (stop && *stop)
        return;
    const int r0 = 2 * blockIdx.y, r1 = r0 + 1;
    const bool use0 = !b.mask || b.mask[r0];
    const bool use1 = r1 < nrhs && (!b.mask || b.mask[r1]);
    if (!use0 && !use1)
        return;
    const double a0 = b.a_rhs ? b.a_rhs[r0] : a;
    const double a1 = (b.a_rhs && use1) ? b.a_rhs[r1] : a;
    const double hc0 = b.c_rhs ? b.c_rhs[r0] : hc;
    const double hc1 = (b.c_rhs && use1) ? b.c_rhs[r1] : hc;
    const size_t ro0 = (size_t)r0 * b.rhs_elems;
    const size_t ro1 = use1 ? (size_t)r1 * b.rhs_elems : ro0; // never read when !use1

    const int lb = logical_block(remap);
    const int bpp = gridDim.x / nparity;
    int p, blk;
    if (IL) {
        p = lb & 1;
        blk = block_order(lb >> 1, g);
    } else {
        p = (nparity == 2) ? lb / bpp : cb_first;
        blk = block_order(lb - (nparity == 2 ? p * bpp : 0), g);
    }
    const int i = blk * kBlock + threadIdx.x;
    if (i >= g.Vh)
        return;
    const int Lx = 2 * g.Lxh;
    int rr = i / g.Lxh;
    const int xh = i - rr * g.Lxh;
    const int y = rr % g.Ly;
    rr /= g.Ly;
    const int z = rr % g.Lz;
    const int t = rr / g.Lz;
    const int x = 2 * xh + ((p + y + z + t) & 1);
    const int row = i - xh;

    const long po = (long)(1 - p) * 12 * g.stride;
    const double2 *in0 = in + ro0 + po, *in1 = in + ro1 + po;
    cpx<double> acc0[4][3], acc1[4][3];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            acc0[s][c] = acc1[s][c] = {0.0, 0.0};
    {
        const int xf = (x + 1 == Lx) ? 0 : x + 1, xb = (x == 0) ? Lx - 1 : x - 1;
        hop_mu2<0, DAG, !IL, LN>(acc0, acc1, use0, use1, in0, in1, U, g, p, i, row + (xf >> 1),
                             row + (xb >> 1), x == Lx - 1 ? g.bc[0] : 1.0, x == 0 ? g.bc[0] : 1.0);
    }
    {
        const int s = g.Lxh;
        const int jf = (y + 1 == g.Ly) ? i - (g.Ly - 1) * s : i + s;
        const int jb = (y == 0) ? i + (g.Ly - 1) * s : i - s;
        hop_mu2<1, DAG, !IL, LN>(acc0, acc1, use0, use1, in0, in1, U, g, p, i, jf, jb,
                             y == g.Ly - 1 ? g.bc[1] : 1.0, y == 0 ? g.bc[1] : 1.0);
    }
    {
        const int s = g.Lxh * g.Ly;
        const int jf = (z + 1 == g.Lz) ? i - (g.Lz - 1) * s : i + s;
        const int jb = (z == 0) ? i + (g.Lz - 1) * s : i - s;
        hop_mu2<2, DAG, !IL, LN>(acc0, acc1, use0, use1, in0, in1, U, g, p, i, jf, jb,
                             z == g.Lz - 1 ? g.bc[2] : 1.0, z == 0 ? g.bc[2] : 1.0);
    }
    {
        const int s = g.Lxh * g.Ly * g.Lz;
        const int jf = (t + 1 == g.Lt) ? i - (g.Lt - 1) * s : i + s;
        const int jb = (t == 0) ? i + (g.Lt - 1) * s : i - s;
        hop_mu2<3, DAG, !IL, LN>(acc0, acc1, use0, use1, in0, in1, U, g, p, i, jf, jb,
                             t == g.Lt - 1 ? g.bc[3] : 1.0, t == 0 ? g.bc[3] : 1.0);
    }
    const long oo = (long)p * 12 * g.stride;
    constexpr bool WS = MODE == HOP_M;
    if (use0)
        store_out(out + ro0 + oo, WS ? same + ro0 + oo : nullptr, i, a0, hc0, acc0, WS);
    if (use1)
        store_out(out + ro1 + oo, WS ? same + ro1 + oo : nullptr, i, a1, hc1, acc1, WS);
}

// t-streaming kernel for the both-parity M / M^dagger launches of one system
// (HOP_M, HOP_M_DOT, HOP_M_NORM, HOP_M_AXPY).  A thread
// owns the column (x, y, z) for Tb consecutive t.  Site (x, y, z, t) has
// parity (x + y + z + t) & 1, so the column's next site is the forward t
// neighbour: the t link of a site serves its forward hop and, as U^dagger,
// the backward hop of the next site, and the site's input spinor is the
// backward t neighbour of the next site; both stay in registers (the
// backward t half spinor of site t + 1 is computed at step t).  The columns
// of type q, (x + y + z) & 1 == q, give every lane of a wavefront the parity
// (q + t) & 1 at step t, so a wavefront reads consecutive indices; the two
// types of a column range are consecutive logical workgroups and read each
// other's sites as x, y, z neighbours.  Directions are accumulated in the
// order of k_hop_f64 (x, y, z, t forward, t backward), so out is bit-identical
// to k_hop_f64's.  HOP_M_DOT partials are per workgroup, indexed by the work
// unit (t run, column block, type).
template <bool DAG, int MODE, int LN>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_hop_f64_ts(
    double2 *__restrict__ out, const double2 *__restrict__ in, const double2 *__restrict__ U,
    Geom g, double a, double hc, const double2 *__restrict__ same,
    const double2 *__restrict__ dotv, double2 *__restrict__ partials, int Tb, int ncb, int nwork,
    int remap, const int *__restrict__ stop, const double *__restrict__ axpy)
{
    if (stop && *stop)
        return;
    // the grid is padded to a multiple of the XCD count so the XCD mapping
    // applies (the two column types of a range then share one L2)
    const int lb = logical_block(remap);
    if (lb >= nwork)
        return;
    const int q = lb & 1;
    const int cbk = (lb >> 1) % ncb, k = (lb >> 1) / ncb;
    const int ncol = g.Lxh * g.Ly * g.Lz; // columns of one type
    const int c0 = cbk * kBlock + threadIdx.x;
    // lanes past the last column shadow a valid one and store nothing, so
    // every thread reaches the workgroup reduction of HOP_M_DOT
    const bool live = c0 < ncol;
    const int c = live ? c0 : ncol - 1;
    const int Lx = 2 * g.Lxh;
    const int xh = c % g.Lxh;
    const int y = (c / g.Lxh) % g.Ly;
    const int z = c / (g.Lxh * g.Ly);
    const int x = 2 * xh + ((y + z + q) & 1);
    const int slice = ncol;
    const int t0 = k * Tb;
    const int stride = g.stride;
    const int xf = (x + 1 == Lx) ? 0 : x + 1, xb = (x == 0) ? Lx - 1 : x - 1;
    const double sxf = x == Lx - 1 ? g.bc[0] : 1.0, sxb = x == 0 ? g.bc[0] : 1.0;
    const double syf = y == g.Ly - 1 ? g.bc[1] : 1.0, syb = y == 0 ? g.bc[1] : 1.0;
    const double szf = z == g.Lz - 1 ? g.bc[2] : 1.0, szb = z == 0 ? g.bc[2] : 1.0;
    const int dyf = (y + 1 == g.Ly) ? -(g.Ly - 1) * g.Lxh : g.Lxh;
    const int dyb = (y == 0) ? (g.Ly - 1) * g.Lxh : -g.Lxh;
    const int sz = g.Lxh * g.Ly;
    const int dzf = (z + 1 == g.Lz) ? -(g.Lz - 1) * sz : sz;
    const int dzb = (z == 0) ? (g.Lz - 1) * sz : -sz;

    // backward t half spinor of the first site, from (x, y, z, t0 - 1)
    cpx<double> rc[2][3];
    {
        const int tb = (t0 == 0) ? g.Lt - 1 : t0 - 1;
        const int pb = (q + tb) & 1;
        const int ib = ((tb * g.Lz + z) * g.Ly + y) * g.Lxh + xh;
        cpx<double> psi[4][3], W[3][3], h[2][3];
        load_spinor(in + (long)pb * 12 * stride, stride, ib, psi);
        load_link<LN, true>(U + (long)((3 * 2 + pb) * LN) * stride, stride, ib, W);
        using D = Dir<3, false, DAG>;
        project<D>(psi, h);
        su3_mul_half<true>(W, h, rc);
        const double sb = (t0 == 0) ? g.bc[3] : 1.0;
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                rc[s][cc] = {sb * rc[s][cc].r, sb * rc[s][cc].i};
    }

    // one spinor per thread (12 complex, 48 KiB per workgroup): the input
    // spinor of the column's next site, carried from one step to the next
    // (18-real links only)
    __shared__ double2 carry[LN == 9 ? 12 : 1][kBlock];
    double pr = 0.0, pi = 0.0;
    for (int t = t0; t < t0 + Tb; ++t) {
        const int p = (q + t) & 1;
        const int i = ((t * g.Lz + z) * g.Ly + y) * g.Lxh + xh;
        const int row = i - xh;
        const double2 *in_o = in + (long)(1 - p) * 12 * stride;
        cpx<double> acc[4][3];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                acc[s][cc] = {0.0, 0.0};
        hop_mu<0, DAG, false, LN>(acc, in_o, U, g, p, i, row + (xf >> 1), row + (xb >> 1), sxf, sxb);
        hop_mu<1, DAG, false, LN>(acc, in_o, U, g, p, i, i + dyf, i + dyb, syf, syb);
        hop_mu<2, DAG, false, LN>(acc, in_o, U, g, p, i, i + dzf, i + dzb, szf, szb);
        // forward t with the site's t link, kept for the next site's backward
        // hop; no other thread reads it, so it loads non-temporal and leaves
        // L2 to the spinors and x, y, z links the neighbouring columns share
        // (32^3x64 M 0.447 -> 0.437 ms on MI355X)
        cpx<double> W[3][3];
        load_link<LN, true>(U + (long)((3 * 2 + p) * LN) * stride, stride, i, W);
        // the forward t neighbour is the column's next site, whose input
        // spinor the next step needs again (diagonal term, backward hop)
        cpx<double> psi[4][3];
        {
            const int jf = (t + 1 == g.Lt) ? i - (g.Lt - 1) * slice : i + slice;
            cpx<double> h[2][3], r[2][3];
            load_spinor(in_o, stride, jf, psi);
            using D = Dir<3, true, DAG>;
            project<D>(psi, h);
            su3_mul_half<false>(W, h, r);
            reconstruct<D>(acc, r, t == g.Lt - 1 ? g.bc[3] : 1.0);
        }
        pin_acc(acc);
        {
            using D = Dir<3, false, DAG>;
            reconstruct<D>(acc, rc, 1.0);
        }
        pin_acc(acc);
        // the site's input spinor: `same` of HOP_M (when same == in) and the
        // backward t neighbour of the next site; after the run's first step
        // it is the previous step's forward t neighbour, kept in this
        // thread's LDS slots (read before the slots take this step's).
        // Measured on MI355X at 32^3x64: M on 18-real links 0.442 -> 0.435
        // ms; on 12-real links 0.354 -> 0.369 ms, so those reload it.
        constexpr bool CARRY = LN == 9;
        cpx<double> v[4][3];
        if (CARRY && t > t0) {
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                const double2 e = carry[k][threadIdx.x];
                v[k / 3][k % 3] = {e.x, e.y};
            }
        } else {
            load_spinor(in + (long)p * 12 * stride, stride, i, v);
        }
        if (CARRY && t + 1 < t0 + Tb) {
#pragma unroll
            for (int k = 0; k < 12; ++k)
                carry[k][threadIdx.x] = make_double2(psi[k / 3][k % 3].r, psi[k / 3][k % 3].i);
        }
        double2 *out_p = out + (long)p * 12 * stride;
        cpx<double> res[4][3];
        {
            const double2 *same_p = same + (long)p * 12 * stride;
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    const cpx<double> sv =
                        (same == in) ? v[s][cc] : ld(same_p, f64_idx(s * 3 + cc, (unsigned)i, 12));
                    res[s][cc] = {fma(a, sv.r, hc * acc[s][cc].r), fma(a, sv.i, hc * acc[s][cc].i)};
                }
        }
        if (live) {
            if constexpr (MODE == HOP_M_AXPY) {
                // out = dotv - alpha res (res = M in is not stored), |out|^2
                const double2 *dv = dotv + (long)p * 12 * stride;
                const double al = *axpy;
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) {
                        const cpx<double> d = ld(dv, f64_idx(s * 3 + cc, (unsigned)i, 12));
                        res[s][cc] = {fma(-al, res[s][cc].r, d.r), fma(-al, res[s][cc].i, d.i)};
                    }
            }
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc)
                    st(out_p, f64_idx(s * 3 + cc, (unsigned)i, 12), res[s][cc]);
            if constexpr (MODE == HOP_M_NORM || MODE == HOP_M_AXPY) {
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) {
                        pr = fma(res[s][cc].r, res[s][cc].r, pr);
                        pr = fma(res[s][cc].i, res[s][cc].i, pr);
                    }
            }
            if constexpr (MODE == HOP_M_DOT) {
                const double2 *dv = dotv + (long)p * 12 * stride;
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) {
                        const cpx<double> d = ld(dv, f64_idx(s * 3 + cc, (unsigned)i, 12));
                        pr = fma(d.r, res[s][cc].r, pr);
                        pr = fma(d.i, res[s][cc].i, pr);
                        pi = fma(d.r, res[s][cc].i, pi);
                        pi = fma(-d.i, res[s][cc].r, pi);
                    }
            }
        }
        if (t + 1 < t0 + Tb) {
            cpx<double> h[2][3];
            using D = Dir<3, false, DAG>;
            project<D>(v, h);
            su3_mul_half<true>(W, h, rc);
        }
    }
    if constexpr (MODE == HOP_M_DOT || MODE == HOP_M_NORM || MODE == HOP_M_AXPY) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            pr += __shfl_xor(pr, off, 64);
            pi += __shfl_xor(pi, off, 64);
        }
        __shared__ double sr[kBlock / 64], si[kBlock / 64];
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            sr[w] = pr;
            si[w] = pi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double tr = 0.0, ti = 0.0;
            for (int kk = 0; kk < kBlock / 64; ++kk) {
                tr += sr[kk];
                ti += si[kk];
            }
            partials[lb] = make_double2(tr, ti);
        }
    }
}

// Workgroups of a t-streaming launch with runs of tb.
static int tstream_grid(const Geom &g, int tb)
{
    const int ncb = (g.Lxh * g.Ly * g.Lz + kBlock - 1) / kBlock;
    return 2 * ncb * (g.Lt / tb);
}

template <bool DAG, int LN>
int launch_mode(wd_ctx *ctx, double2 *out, const double2 *in, const double2 *U, int cb, int mode,
                double a, double hc, const double2 *same, const double2 *dotv, double2 *partials,
                const int *stop, const F64Batch &b, int nrhs)
{
    const Geom &g = ctx->geom;
    const int nparity = (cb < 0) ? 2 : 1;
    const int bpp = (g.Vh + kBlock - 1) / kBlock;
    const int grid = bpp * nparity;
    const int cb_first = (cb < 0) ? 0 : cb;
    const bool red = mode == HOP_M_DOT || mode == HOP_M_NORM || mode == HOP_M_AXPY;
    if (red && partials == ctx->partials && (long)grid * nrhs > ctx->max_partials)
        return wd_set_error(ctx, WD_EARG, "hop: too many workgroups for the partials buffer");
    if (nparity == 2 && nrhs == 1 && ctx->ts_tb > 0 && !b.mask && !b.a_rhs && !b.c_rhs &&
        g.Lt % ctx->ts_tb == 0 && (mode == HOP_M || red)) {
        const int tb = ctx->ts_tb;
        const int ncb = (g.Lxh * g.Ly * g.Lz + kBlock - 1) / kBlock;
        const int gts = tstream_grid(g, tb);
        const int glaunch = (gts + kNumXCD - 1) / kNumXCD * kNumXCD;
        if (red && partials == ctx->partials && gts > ctx->max_partials)
            return wd_set_error(ctx, WD_EARG, "hop: too many workgroups for the partials buffer");
#define WD_HOP_TS_LAUNCH(M)                                                                        \
    hipLaunchKernelGGL((k_hop_f64_ts<DAG, M, LN>), dim3(glaunch), dim3(kBlock), 0, ctx->stream, out, in, \
                       U, g, a, hc, same, dotv, partials, tb, ncb, gts, ctx->xcd_remap, stop, b.axpy)
        if (mode == HOP_M)
            WD_HOP_TS_LAUNCH(HOP_M);
        else if (mode == HOP_M_NORM)
            WD_HOP_TS_LAUNCH(HOP_M_NORM);
        else if (mode == HOP_M_AXPY)
            WD_HOP_TS_LAUNCH(HOP_M_AXPY);
        else
            WD_HOP_TS_LAUNCH(HOP_M_DOT);
#undef WD_HOP_TS_LAUNCH
        return wd_check_launch(ctx, "k_hop_f64_ts");
    }
    const bool il = nparity == 2 && ctx->parity_il;
    if (nrhs >= 2 && ctx->f64_pairs && (mode == HOP_ONLY || mode == HOP_M)) {
        const dim3 gp(grid, (nrhs + 1) / 2);
#define WD_HOP_PAIR_LAUNCH(M, I)                                                                   \
    hipLaunchKernelGGL((k_hop_f64_pair<DAG, M, I, LN>), gp, dim3(kBlock), 0, ctx->stream, out, in, U, \
                       g, cb_first, nparity, a, hc, same, ctx->xcd_remap, stop, b, nrhs)
        if (mode == HOP_ONLY) {
            if (il)
                WD_HOP_PAIR_LAUNCH(HOP_ONLY, true);
            else
                WD_HOP_PAIR_LAUNCH(HOP_ONLY, false);
        } else {
            if (il)
                WD_HOP_PAIR_LAUNCH(HOP_M, true);
            else
                WD_HOP_PAIR_LAUNCH(HOP_M, false);
        }
#undef WD_HOP_PAIR_LAUNCH
        return wd_check_launch(ctx, "k_hop_f64_pair");
    }
    const dim3 gd(grid, nrhs);
#define WD_HOP_F64_LAUNCH(M, I)                                                                    \
    hipLaunchKernelGGL((k_hop_f64<DAG, M, I, LN>), gd, dim3(kBlock), 0, ctx->stream, out, in, U, g,  \
                       cb_first, nparity, a, hc, same, dotv, partials, ctx->xcd_remap, stop, b)
    switch (mode) {
    case HOP_ONLY:
        if (il)
            WD_HOP_F64_LAUNCH(HOP_ONLY, true);
        else
            WD_HOP_F64_LAUNCH(HOP_ONLY, false);
        break;
    case HOP_M:
        if (il)
            WD_HOP_F64_LAUNCH(HOP_M, true);
        else
            WD_HOP_F64_LAUNCH(HOP_M, false);
        break;
    case HOP_M_NORM:
        if (il)
            WD_HOP_F64_LAUNCH(HOP_M_NORM, true);
        else
            WD_HOP_F64_LAUNCH(HOP_M_NORM, false);
        break;
    case HOP_M_AXPY:
        if (il)
            WD_HOP_F64_LAUNCH(HOP_M_AXPY, true);
        else
            WD_HOP_F64_LAUNCH(HOP_M_AXPY, false);
        break;
    default:
        if (il)
            WD_HOP_F64_LAUNCH(HOP_M_DOT, true);
        else
            WD_HOP_F64_LAUNCH(HOP_M_DOT, false);
        break;
    }
#undef WD_HOP_F64_LAUNCH
    return wd_check_launch(ctx, "k_hop_f64");
}

// Dispatch on the link format: 18-real (the reference's) or 12-real links.
template <bool DAG>
int launch_links(wd_ctx *ctx, double2 *out, const double2 *in, F64Links U, int cb, int mode,
                 double a, double hc, const double2 *same, const double2 *dotv,
                 double2 *partials, const int *stop, const F64Batch &b, int nrhs)
{
    if (U.recon == WD_RECON_12)
        return launch_mode<DAG, 6>(ctx, out, in, U.p, cb, mode, a, hc, same, dotv, partials, stop,
                                   b, nrhs);
    return launch_mode<DAG, 9>(ctx, out, in, U.p, cb, mode, a, hc, same, dotv, partials, stop, b,
                               nrhs);
}

} // namespace

// cb in {0, 1}: one output parity; cb = -1: both parities in one launch.
// out = a same + hc hop(in) (HOP_ONLY: hc hop(in)); hc = -1/2 is the Wilson
// operator's coefficient.
int wd_launch_hop_f64c(wd_ctx *ctx, double2 *out, const double2 *in, F64Links U, int dagger,
                       int cb, int mode, double a, double hc, const double2 *same,
                       const double2 *dotv, double2 *partials, const int *stop)
{
    const F64Batch one = {0, nullptr, nullptr, nullptr};
    if (dagger)
        return launch_links<true>(ctx, out, in, U, cb, mode, a, hc, same, dotv, partials, stop,
                                  one, 1);
    return launch_links<false>(ctx, out, in, U, cb, mode, a, hc, same, dotv, partials, stop, one,
                               1);
}

// r = r - alpha (a same - 1/2 hop(in)), HOP_M_AXPY, with the partials of |r|^2;
// alpha (real) in device memory
int wd_launch_hop_f64_axpy(wd_ctx *ctx, double2 *r, const double2 *in, F64Links U, int dagger,
                           int cb, double a, const double2 *same, const double *alpha,
                           double2 *partials, const int *stop)
{
    const F64Batch one = {0, nullptr, nullptr, nullptr, alpha};
    if (dagger)
        return launch_links<true>(ctx, r, in, U, cb, HOP_M_AXPY, a, -0.5, same, r, partials, stop,
                                  one, 1);
    return launch_links<false>(ctx, r, in, U, cb, HOP_M_AXPY, a, -0.5, same, r, partials, stop,
                               one, 1);
}

int wd_launch_hop_f64(wd_ctx *ctx, double2 *out, const double2 *in, F64Links U, int dagger,
                      int cb, int mode, double a, const double2 *same, const double2 *dotv,
                      double2 *partials, const int *stop)
{
    return wd_launch_hop_f64c(ctx, out, in, U, dagger, cb, mode, a, -0.5, same, dotv, partials,
                              stop);
}

// nrhs systems at once (blockIdx.y = rhs); b.mask skips systems, b.a_rhs and
// b.c_rhs give per-system diagonal and hop coefficients (else hc);
// partials are [rhs][workgroup].
int wd_launch_hop_f64_batch_c(wd_ctx *ctx, double2 *out, const double2 *in, F64Links U,
                              int dagger, int cb, int mode, double hc, const double2 *same,
                              const double2 *dotv, double2 *partials, const int *stop,
                              const F64Batch &b, int nrhs)
{
    if (dagger)
        return launch_links<true>(ctx, out, in, U, cb, mode, 0.0, hc, same, dotv, partials, stop,
                                  b, nrhs);
    return launch_links<false>(ctx, out, in, U, cb, mode, 0.0, hc, same, dotv, partials, stop, b,
                               nrhs);
}

int wd_launch_hop_f64_batch(wd_ctx *ctx, double2 *out, const double2 *in, F64Links U,
                            int dagger, int cb, int mode, const double2 *same,
                            const double2 *dotv, double2 *partials, const int *stop,
                            const F64Batch &b, int nrhs)
{
    return wd_launch_hop_f64_batch_c(ctx, out, in, U, dagger, cb, mode, -0.5, same, dotv,
                                     partials, stop, b, nrhs);
}

// Number of workgroups a hop launch of one system over `nparity` parities
// uses (the partials count of HOP_M_DOT).
int wd_hop_grid(const wd_ctx *ctx, int nparity)
{
    if (nparity == 2 && ctx->ts_tb > 0 && ctx->geom.Lt % ctx->ts_tb == 0)
        return tstream_grid(ctx->geom, ctx->ts_tb);
    return ((ctx->geom.Vh + kBlock - 1) / kBlock) * nparity;
}

// Automatic t-run length: 16 where the launch keeps at least 512 workgroups
// (measured on MI355X, fused M against the site-block kernel: 32^3x64 -8 %,
// 40^3x80 -5 %, 48^3x96 -4 %, 64^4 -6 %; with fewer workgroups the
// site-block kernel is faster: 24^3x48 +19 %, 16^3x32 x3.6).
int wd_auto_tstream(const wd_ctx *ctx)
{
    const Geom &g = ctx->geom;
    if (g.Lt % 16 != 0)
        return 0;
    return tstream_grid(g, 16) >= 512 ? 16 : 0;
}
