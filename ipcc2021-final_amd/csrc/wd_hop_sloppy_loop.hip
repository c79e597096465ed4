// wd_hop_sloppy_loop.hip -- the sloppy Wilson operator with every system of a
// site in one thread: F16 spinors (per-site scale), F16 12-real links, SU(3)
// products in packed halves with FP32 accumulation.
//
// Work decomposition: a thread owns one site of one parity and loops over the
// nrhs systems.  Its 8 links (U_mu(x) forward, U_mu(x - mu) backward) are
// loaded once, row 2 rebuilt, and kept in registers as packed halves for all
// systems, so a launch reads every link once however many systems it applies
// to, and the site's neighbour offsets and boundary factors are computed
// once.  All workgroups resident on an XCD walk the systems in the same order
// over a compact site range, so the neighbour spinors of the system being
// processed are shared through L2 while it is small (one system's F16 data of
// the resident sites) -- the kernel with R systems per thread keeps R systems
// of every resident site in flight instead, and their neighbourhoods
// overflow L2 at many systems.  The systems run in groups (at most 12 by
// default), each group one pass of workgroups over all sites: workgroups
// that start later are then fewer systems behind their neighbours, whose
// spinors they read, than in one pass over every system.  Groups are cut
// from the applied systems (`active` == 1) on the device, so masked systems
// cost neither a pass nor a link load.
//
// Arithmetic per direction and system: the half spinor h = psi_a + P psi_b is
// formed with packed F16 adds (one v_pk_add_f16 per complex and operand form,
// the sign and swap patterns ride in the instruction's op_sel / neg
// modifiers), each complex product U h accumulates in FP32 by two
// v_dot2_f32_f16 (the F16 x F16 products are exact in FP32), and the
// reconstruction into the FP32 accumulator is the 24 FMAs of the other
// sloppy kernel, with the site scale, the boundary phase and the hop
// coefficient in one factor.  Against the FP32 products of the R-per-thread
// kernel the only extra roundings are h and the rebuilt row 2 of each link
// (to F16, relative 2^-11, the size of the F16 storage rounding).
//
// Modes as for the other sloppy kernel: HOP_ONLY, HOP_M, HOP_M_DOT (per-system
// FP64 partials of sum conj(dotv) out over the stored values), HOP_M_DOT_NORM
// (plus |out|^2; no dagger form), HOP_M_NORM (|out|^2 only; no dagger form),
// HOP_M_AXPY (out = dotv - axpy_r (M in) and its |out|^2; dagger form only).
// The partials are per wavefront, kLoopWaves per workgroup ([rhs][nsb * kLoopWaves]), summed by shuffles without a
// workgroup barrier per system.
#include "wd_fields.h"
#include "wd_internal.h"
#include "wd_sloppy.h"
#include "wd_sloppy_dev.h"

using namespace wdf;
using namespace wdsl;

namespace {

typedef _Float16 h2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ h2v as_h2(unsigned w) { return __builtin_bit_cast(h2v, w); }

// Operand forms of h = a + P b (P = +1, -1, +i, -i for codes 0..3), each one
// packed add:  F 0: (hr, -hi)   F 1: (hi, hr)   F 2: (hr, hi)   F 3: (hi, -hr)
template <int P, int F>
__device__ __forceinline__ h2v hform(h2v a, h2v b)
{
    // P b as (real, imaginary)
    const _Float16 pr = P == 0 ? b.x : P == 1 ? -b.x : P == 2 ? -b.y : b.y;
    const _Float16 pi = P == 0 ? b.y : P == 1 ? -b.y : P == 2 ? b.x : -b.x;
    if constexpr (F == 0)
        return h2v{a.x, -a.y} + h2v{pr, -pi};
    else if constexpr (F == 1)
        return h2v{a.y, a.x} + h2v{pi, pr};
    else if constexpr (F == 2)
        return h2v{a.x, a.y} + h2v{pr, pi};
    else
        return h2v{a.y, -a.x} + h2v{pi, -pr};
}

// One F16 12-real link as 9 packed complex (rows 0 and 1 as stored, row 2 =
// conj(row0 x row1) computed in FP32 and rounded to F16).
template <bool NT>
__device__ __forceinline__ void load_link_h2(const char *base, int stride, unsigned j,
                                             unsigned (&W)[3][3])
{
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const u2v c = ldg<NT>(reinterpret_cast<const u2v *>(base + ((unsigned)q * (unsigned)stride + j) * 8u));
        W[(2 * q) / 3][(2 * q) % 3] = c.x;
        W[(2 * q + 1) / 3][(2 * q + 1) % 3] = c.y;
    }
    cpx<float> U[3][3];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            U[a][b] = {re16(W[a][b]), im16(W[a][b])};
    reconstruct_row2(U);
#pragma unroll
    for (int b = 0; b < 3; ++b)
        W[2][b] = f2h(U[2][b].r, U[2][b].i);
}

// acc[.][a] += f * reconstruct of colour a of the product (r0 for the half
// spinor's first spin, r1 for its second): the colour-a part of the
// 24-FMA reconstruction, so the product's registers die row by row.
template <class D>
__device__ __forceinline__ void reconstruct_colour(cpx<float> (&acc)[4][3], int a, cpx<float> r0,
                                                   cpx<float> r1, float f)
{
    acc_mulc<0>(acc[0][a], r0, f);
    acc_mulc<0>(acc[1][a], r1, f);
    acc_mulc<D::R0>(acc[D::S0][a], r0, f);
    acc_mulc<D::R1>(acc[D::S1][a], r1, f);
}

// The hop factor of direction d: the (uniform) hop coefficient with its sign
// flipped when the hop crosses an antiperiodic boundary (bit d of sbits).
__device__ __forceinline__ float hop_factor(float hc, unsigned sbits, int d)
{
    return __uint_as_float(__float_as_uint(hc) ^ ((sbits << (31 - d)) & 0x80000000u));
}

// The 12 packed complex and the scale of one neighbour site.
struct Nbr {
    unsigned w[12];
    float s;
};

__device__ __forceinline__ void load_nbr(const char *inb, const float *isc, unsigned j, int stride,
                                         Nbr &n)
{
    load16(inb, stride, j, n.w);
    n.s = load_scale(isc, j);
}

__device__ __forceinline__ void pin_acc(cpx<float> (&acc)[4][3])
{
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            asm volatile("" : "+v"(acc[s][c].r), "+v"(acc[s][c].i));
}

// acc += f * reconstruct(W h) (forward) or f * reconstruct(W^dag h)
// (backward) for direction D, h projected from the neighbour site n, f = hs *
// its scale.  Ends with the accumulator pinned: the loads issued after this
// call in program order (the direction after next) stay behind it, so one
// direction's loads are in flight during each direction's arithmetic --
// issued all at once, the loads of eight directions would need twice the
// registers.
template <int MU, bool FWD, bool DAG>
__device__ __forceinline__ void dir_math(cpx<float> (&acc)[4][3], const unsigned (&W)[3][3],
                                         const Nbr &n, float hs)
{
    using D = Dir<MU, FWD, DAG>;
    const unsigned(&w)[12] = n.w;
    const float f = hs * n.s;
    if constexpr (FWD) {
        // re = sum_b Ur hr - Ui hi = dot(U, (hr, -hi)); im = dot(U, (hi, hr))
        h2v x0[3], y0[3], x1[3], y1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x0[c] = hform<D::P0, 0>(as_h2(w[c]), as_h2(w[3 * D::S0 + c]));
            y0[c] = hform<D::P0, 1>(as_h2(w[c]), as_h2(w[3 * D::S0 + c]));
            x1[c] = hform<D::P1, 0>(as_h2(w[3 + c]), as_h2(w[3 * D::S1 + c]));
            y1[c] = hform<D::P1, 1>(as_h2(w[3 + c]), as_h2(w[3 * D::S1 + c]));
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float r0 = 0.f, i0 = 0.f, r1 = 0.f, i1 = 0.f;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const h2v u = as_h2(W[a][b]);
                r0 = __builtin_amdgcn_fdot2(u, x0[b], r0, false);
                i0 = __builtin_amdgcn_fdot2(u, y0[b], i0, false);
                r1 = __builtin_amdgcn_fdot2(u, x1[b], r1, false);
                i1 = __builtin_amdgcn_fdot2(u, y1[b], i1, false);
            }
            reconstruct_colour<D>(acc, a, {r0, i0}, {r1, i1}, f);
        }
    } else {
        // conj(U_ba) h: re = dot(U_ba, (hr, hi)); im = dot(U_ba, (hi, -hr))
        h2v x0[3], y0[3], x1[3], y1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x0[c] = hform<D::P0, 2>(as_h2(w[c]), as_h2(w[3 * D::S0 + c]));
            y0[c] = hform<D::P0, 3>(as_h2(w[c]), as_h2(w[3 * D::S0 + c]));
            x1[c] = hform<D::P1, 2>(as_h2(w[3 + c]), as_h2(w[3 * D::S1 + c]));
            y1[c] = hform<D::P1, 3>(as_h2(w[3 + c]), as_h2(w[3 * D::S1 + c]));
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float r0 = 0.f, i0 = 0.f, r1 = 0.f, i1 = 0.f;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const h2v u = as_h2(W[b][a]);
                r0 = __builtin_amdgcn_fdot2(u, x0[b], r0, false);
                i0 = __builtin_amdgcn_fdot2(u, y0[b], i0, false);
                r1 = __builtin_amdgcn_fdot2(u, x1[b], r1, false);
                i1 = __builtin_amdgcn_fdot2(u, y1[b], i1, false);
            }
            reconstruct_colour<D>(acc, a, {r0, i0}, {r1, i1}, f);
        }
    }
    pin_acc(acc);
}

// The identity link in the register forms of the two link formats
__device__ __forceinline__ void unit_link(unsigned (&W)[3][3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            W[a][b] = a == b ? 0x3c00u : 0u; // (1.0, 0.0) as packed halves
}
__device__ __forceinline__ void unit_link(cpx<float> (&W)[3][3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            W[a][b] = {a == b ? 1.f : 0.f, 0.f};
}

// Link formats of the loop kernel: F16 12-real links kept as packed halves
// (72 registers, products by packed dot products), or F32 12-real links kept
// as FP32 complex with row 2 rebuilt (144 registers, the FP32 products of the
// R-systems-per-thread kernel, whose results they reproduce bit for bit).
enum LoopLinks { LL_F16 = 0, LL_F32 = 1 };

template <int LF>
struct LinkRegs;
template <>
struct LinkRegs<LL_F16> {
    unsigned W[8][3][3];
};
template <>
struct LinkRegs<LL_F32> {
    cpx<float> W[8][3][3];
};

// acc += f * reconstruct(W h) or f * reconstruct(W^dag h) in FP32, h projected
// from the F16 neighbour n: the per-direction arithmetic of the kernel with
// R systems per thread (same operations in the same order).
template <int MU, bool FWD, bool DAG>
__device__ __forceinline__ void dir_math32(cpx<float> (&acc)[4][3], const cpx<float> (&W)[3][3],
                                           const Nbr &n, float hs, float one)
{
    using D = Dir<MU, FWD, DAG>;
    cpx<float> h[2][3], res[2][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        h[0][c] = proj16<D::P0>(n.w[c], n.w[3 * D::S0 + c], one);
        h[1][c] = proj16<D::P1>(n.w[3 + c], n.w[3 * D::S1 + c], one);
    }
    su3_mul_half<!FWD>(W, h, res);
    reconstruct_f<D>(acc, res, hs * n.s);
    pin_acc(acc);
}

// Sum of (sr, si) over the wavefront, stored by lane 0: the per-system partials
// of this kernel are per wavefront (kBlock / 64 per workgroup), so no
// workgroup barrier is needed per system.
__device__ __forceinline__ void wave_sum_store(double sr, double si, double2 *dst)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    if ((threadIdx.x & 63) == 0)
        *dst = make_double2(sr, si);
}

// The 8 hop directions of one system, d = 2 mu + (backward), with the
// neighbour loads K directions ahead in K + 1 rotating buffers; the loads of
// directions 0 .. PRE - 1 (PRE = K or K + 1) were issued before the walk.
template <bool DAG, int MODE, int K, int LF, int PRE>
struct Walk {
    const SloppyArgs &a;
    const LinkRegs<LF> &L;
    const unsigned (&j)[8];
    const char *inb;
    const float *isc;
    size_t off, soff;
    unsigned ui, sbits;
    int stride;

    template <int D>
    __device__ __forceinline__ void step(cpx<float> (&acc)[4][3], Nbr (&nb)[K + 1]) const
    {
        if constexpr (D + K < 8 && D + K >= PRE)
            load_nbr(inb, isc, j[D + K], stride, nb[(D + K) % (K + 1)]);
        else if constexpr (D + K == 8 && MODE != HOP_ONLY)
            load_nbr(a.same + off, a.same_scale + soff, ui, stride, nb[8 % (K + 1)]);
        if constexpr (LF == LL_F16)
            dir_math<D / 2, D % 2 == 0, DAG>(acc, L.W[D], nb[D % (K + 1)], hop_factor(a.hc, sbits, D));
        else
            dir_math32<D / 2, D % 2 == 0, DAG>(acc, L.W[D], nb[D % (K + 1)],
                                               hop_factor(a.hc, sbits, D), a.one);
    }
};

// Prefetch depth and occupancy: F16 links K = 1 at 3 wavefronts per SIMD,
// K = 3 at 2; F32 links (twice the link registers) K = 1 at 2.
template <int K, int LF>
struct LoopShape;
template <>
struct LoopShape<1, LL_F16> {
    static constexpr int waves = 3;
};
template <>
struct LoopShape<3, LL_F16> {
    static constexpr int waves = 2;
};
template <>
struct LoopShape<1, LL_F32> {
    static constexpr int waves = 2;
};

// Bit l of the result: system base + l is applied (active == 1); one flag per
// lane, so every lane of the wavefront must be active.
__device__ __forceinline__ unsigned long long applied_mask(const int *__restrict__ active, int base,
                                                           int nrhs)
{
    const int r = base + (int)(threadIdx.x & 63);
    return __ballot(r < nrhs && active[r] == 1);
}

// Index of the j-th (from 0) set bit of m; j < popcount(m).
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int j)
{
    const int lane = (int)(threadIdx.x & 63);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    return __ffsll((long long)__ballot(((m >> lane) & 1ull) && rank == j)) - 1;
}

// One site block (kBlock sites of one parity) for systems [r0, r1).
template <bool DAG, int MODE, int K, int LF>
__device__ __forceinline__ void loop_block(const SloppyArgs &a, const Geom &g, int sb, int cb_first,
                                           int nparity, int il, int r0, int r1)
{
    // both parities alternate over the same site range (shared L2 reuse)
    // both parities: side by side (il) or parity 0 first, then parity 1
    const int p = (nparity == 2) ? (il ? (sb & 1) : sb / a.bpp) : cb_first;
    const int blk = block_order((nparity == 2) ? (il ? (sb >> 1) : sb - p * a.bpp) : sb, g);
    const int nat = (nparity == 2) ? 2 * blk + p : blk; // natural index (partials)
    const int i = blk * kBlock + threadIdx.x;
    const int stride = g.stride;
    const bool site_ok = i < g.Vh;
    const size_t pbytes = (size_t)3 * 16 * stride;

    // the site's 8 links, neighbour offsets and hop factors, for all systems
    LinkRegs<LF> L;
    unsigned jn[8];
    unsigned sbits = 0; // bit d: hop d crosses a boundary with phase -1
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
        const int xf = (x + 1 == Lx) ? 0 : x + 1, xb = (x == 0) ? Lx - 1 : x - 1;
        jn[0] = row + (xf >> 1);
        jn[1] = row + (xb >> 1);
        int s = g.Lxh;
        jn[2] = (y + 1 == g.Ly) ? i - (g.Ly - 1) * s : i + s;
        jn[3] = (y == 0) ? i + (g.Ly - 1) * s : i - s;
        s *= g.Ly;
        jn[4] = (z + 1 == g.Lz) ? i - (g.Lz - 1) * s : i + s;
        jn[5] = (z == 0) ? i + (g.Lz - 1) * s : i - s;
        s *= g.Lz;
        jn[6] = (t + 1 == g.Lt) ? i - (g.Lt - 1) * s : i + s;
        jn[7] = (t == 0) ? i + (g.Lt - 1) * s : i - s;
        const bool cross[8] = {x == Lx - 1, x == 0, y == g.Ly - 1, y == 0,
                               z == g.Lz - 1, z == 0, t == g.Lt - 1, t == 0};
#pragma unroll
        for (int d = 0; d < 8; ++d)
            if (cross[d] && g.bc[d >> 1] < 0.0)
                sbits |= 1u << d;
        // temporal gauge: the forward t link is the identity off the last
        // slice, the backward one (U_t of slice t - 1) off the first; such
        // links are not loaded but set to 1 in registers (links load once
        // per pass here, so a separate unit-link path would only cost
        // registers: measured, it made the hop launches spill)
        unsigned tbits = 0; // bit 0 forward, bit 1 backward t link is 1
        if (a.tg)
            tbits = (t != g.Lt - 1 ? 1u : 0u) | (t != 0 ? 2u : 0u);
        // 12-real links, 3 chunks per site: 8 bytes each (F16), 16 (F32)
        const size_t lblk = (size_t)3 * (LF == LL_F16 ? 8 : 16) * stride;
        auto load_dir = [&](int d, const char *base, unsigned idx) {
            if constexpr (LF == LL_F16) {
                if (a.tlinks)
                    load_link_h2<false>(base, stride, idx, L.W[d]);
                else
                    load_link_h2<true>(base, stride, idx, L.W[d]);
            } else {
                if (a.tlinks)
                    load_link32<G_F32_12, false>(base, stride, idx, L.W[d]);
                else
                    load_link32<G_F32_12, true>(base, stride, idx, L.W[d]);
            }
        };
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) {
            // forward: U_mu(x) at (p, i); backward: U_mu(x - mu) at (1 - p, j)
            const char *Uf = a.U + (size_t)(mu * 2 + p) * lblk;
            const char *Ub = a.U + (size_t)(mu * 2 + 1 - p) * lblk;
            if (mu < 3 || !(tbits & 1u))
                load_dir(2 * mu, Uf, (unsigned)i);
            else
                unit_link(L.W[2 * mu]);
            if (mu < 3 || !(tbits & 2u))
                load_dir(2 * mu + 1, Ub, jn[2 * mu + 1]);
            else
                unit_link(L.W[2 * mu + 1]);
        }
    }

    // the applied systems of [r0, r1) as bits of a wavefront-uniform mask, 64
    // systems at a time: a flag read per system would be a vector load whose
    // wait (the loop's first) also drains the previous system's stores
    unsigned long long am = ~0ull;
    int abase = r0;
    if (a.active)
        am = applied_mask(a.active, r0, a.nrhs);
    // the first applied system at or after r (r1 when there is none); uniform
    // over the wavefront, so every lane takes part in the ballot
    auto next_applied = [&](int r) -> int {
        if (!a.active)
            return r;
        for (; r < r1; ++r) {
            if (r - abase >= 64) {
                abase = r;
                am = applied_mask(a.active, r, a.nrhs);
            }
            if ((am >> (r - abase)) & 1ull)
                return r;
        }
        return r1;
    };
    // the neighbour offsets are opaque per system: hoisted out of the loop,
    // their 32 byte offsets (3 chunks + scale per neighbour) would stay live
    // as 64-bit values
    auto opaque_offsets = [&](unsigned (&j)[8]) {
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            j[d] = jn[d];
            asm volatile("" : "+v"(j[d]));
        }
    };
    const char *in_p = a.in + (size_t)(1 - p) * pbytes;
    const float *isc_p = a.in_scale + (size_t)(1 - p) * stride;
    const unsigned ui = (unsigned)i;

    // Neighbour buffers: the loads of direction d + K are issued before the
    // arithmetic of direction d, the site's own spinor (the diagonal term)
    // after the last direction's loads.  The loads of the next system's first
    // PRE directions are issued before this system's stores: loads and
    // stores leave the memory counter in the order they were issued, so a
    // wait for a load issued after the stores would also wait for the
    // stores to complete.
    // (two directions ahead where the registers allow; the |out|^2 and axpy
    // launches would spill with two)
    constexpr int PRE = (K == 1 && MODE != HOP_M_NORM && MODE != HOP_M_AXPY) ? 2 : K;
    Nbr nb[K + 1];
    int r = next_applied(r0);
    if (site_ok && r < r1) {
        unsigned j[8];
        opaque_offsets(j);
        const char *inb = in_p + (size_t)r * a.rhs_bytes;
        const float *isc = isc_p + (size_t)r * 2 * stride;
#pragma unroll
        for (int d = 0; d < PRE; ++d)
            load_nbr(inb, isc, j[d], stride, nb[d]);
    }
#pragma unroll 1
    while (r < r1) {
        const int rn = next_applied(r + 1);
        cpx<float> acc[4][3];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                acc[s][c] = {0.f, 0.f};
        const size_t off = (size_t)r * a.rhs_bytes + (size_t)p * pbytes;
        const size_t soff = (size_t)r * 2 * stride + (size_t)p * stride;
        double pr = 0.0, pi = 0.0, nn = 0.0;
        if (site_ok) {
            const char *inb = in_p + (size_t)r * a.rhs_bytes;
            const float *isc = isc_p + (size_t)r * 2 * stride;
            unsigned j[8];
            opaque_offsets(j);
            const Walk<DAG, MODE, K, LF, PRE> walk{a, L, j, inb, isc, off, soff, ui, sbits, stride};
            walk.template step<0>(acc, nb);
            walk.template step<1>(acc, nb);
            walk.template step<2>(acc, nb);
            walk.template step<3>(acc, nb);
            walk.template step<4>(acc, nb);
            walk.template step<5>(acc, nb);
            walk.template step<6>(acc, nb);
            walk.template step<7>(acc, nb);
            const Nbr &own = nb[8 % (K + 1)];
            // acc holds hc hop(in); HOP_M adds a_r in at the site
            if constexpr (MODE != HOP_ONLY) {
                const float as = a.diag[r] * own.s;
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    acc[k / 3][k % 3].r = fmaf(re16(own.w[k]), as, acc[k / 3][k % 3].r);
                    acc[k / 3][k % 3].i = fmaf(im16(own.w[k]), as, acc[k / 3][k % 3].i);
                }
            }
            unsigned q[12];
            float sq = pack16(acc, q);
            // the dot product's operand (HOP_M_AXPY: the updated field), then
            // the next system's first neighbours, are loaded before the stores
            unsigned dv[12];
            float sd = 0.f;
            if constexpr (MODE == HOP_M_DOT || MODE == HOP_M_DOT_NORM || MODE == HOP_M_AXPY) {
                // (not hoisted above the packing: the axpy launch would spill)
                if constexpr (MODE == HOP_M_AXPY)
                    __builtin_amdgcn_sched_barrier(0);
                load16(a.dotv + off, stride, ui, dv);
                sd = load_scale(a.dotv_scale + soff, ui);
            }
            auto prefetch_next = [&]() {
                if (rn < r1) {
                    const char *inb_n = in_p + (size_t)rn * a.rhs_bytes;
                    const float *isc_n = isc_p + (size_t)rn * 2 * stride;
#pragma unroll
                    for (int d = 0; d < PRE; ++d)
                        load_nbr(inb_n, isc_n, j[d], stride, nb[d]);
                }
            };
            if constexpr (MODE == HOP_M_AXPY) {
                // out = dotv - axpy_r (M in), M in as rounded for storage: the
                // arithmetic of an elementwise update reading M in back; the
                // next system's loads follow (before them, dotv would share
                // the registers with them and spill)
                const float cr = a.axpy[r];
                cpx<float> t[4][3];
#pragma unroll
                for (int k = 0; k < 12; ++k)
                    t[k / 3][k % 3] = {fmaf(-cr, re16(q[k]) * sq, re16(dv[k]) * sd),
                                       fmaf(-cr, im16(q[k]) * sq, im16(dv[k]) * sd)};
                sq = pack16(t, q);
            }
            prefetch_next();
            store16w(a.out + off, a.out_scale + soff, stride, ui, q, sq);
            if constexpr (MODE == HOP_M_DOT_NORM || MODE == HOP_M_NORM || MODE == HOP_M_AXPY) {
                float nq = 0.f;
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    nq = fmaf(re16(q[k]), re16(q[k]), nq);
                    nq = fmaf(im16(q[k]), im16(q[k]), nq);
                }
                nn = (double)nq * ((double)sq * (double)sq);
            }
            if constexpr (MODE == HOP_M_DOT || MODE == HOP_M_DOT_NORM) {
                float tr = 0.f, ti = 0.f;
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    tr = fmaf(re16(dv[k]), re16(q[k]), tr);
                    tr = fmaf(im16(dv[k]), im16(q[k]), tr);
                    ti = fmaf(re16(dv[k]), im16(q[k]), ti);
                    ti = fmaf(-im16(dv[k]), re16(q[k]), ti);
                }
                const double ss = (double)sd * (double)sq;
                pr = (double)tr * ss;
                pi = (double)ti * ss;
            }
        }
        // partials per wavefront: [rhs][nsb * kLoopWaves], wavefront w of
        // block nat at nat * kLoopWaves + w
        const size_t pidx = ((size_t)r * a.nsb + nat) * kLoopWaves + (threadIdx.x >> 6);
        if constexpr (MODE == HOP_M_DOT || MODE == HOP_M_DOT_NORM)
            wave_sum_store(pr, pi, a.partials + pidx);
        if constexpr (MODE == HOP_M_DOT_NORM)
            wave_sum_store(nn, 0.0, a.partials2 + pidx);
        if constexpr (MODE == HOP_M_NORM || MODE == HOP_M_AXPY)
            wave_sum_store(nn, 0.0, a.partials + pidx);
        r = rn;
    }
}

template <bool DAG, int MODE, int K, int LF>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(LoopShape<K, LF>::waves))) void
k_hop_sloppy_loop(SloppyArgs a, Geom g, int cb_first, int nparity, int remap, int per, int il)
{
    if (a.stop && *a.stop)
        return;
    // The systems to apply (active[r] == 1) form balanced groups of at most
    // `per`, one after the other over all site blocks (group-major order).
    // Groups are cut from the applied systems, which masking leaves scattered
    // over the indices, so a group never loads links for one or two systems;
    // the groups past the last exit before any load.  The XCD mapping works
    // within each group, so every XCD owns 1/8 of the sites of every group
    // and the used groups keep the whole chip busy.
    const int grp = blockIdx.x / a.nsb;
    int sb = blockIdx.x - grp * a.nsb;
    if (remap && a.nsb % kNumXCD == 0)
        sb = (sb % kNumXCD) * (a.nsb / kNumXCD) + sb / kNumXCD;
    int nact = a.nrhs;
    if (a.active) {
        nact = 0;
        for (int base = 0; base < a.nrhs; base += 64)
            nact += __popcll(applied_mask(a.active, base, a.nrhs));
    }
    const int ng = (nact + per - 1) / per;
    if (grp >= ng)
        return;
    const int k0 = (int)((long)grp * nact / ng), k1 = (int)((long)(grp + 1) * nact / ng);
    int r0 = k0, r1 = k1;
    if (a.active) {
        // r0: index of the k0-th applied system; r1: one past the (k1 - 1)-th
        r0 = r1 = a.nrhs;
        int k = 0;
        for (int base = 0; base < a.nrhs; base += 64) {
            const unsigned long long m = applied_mask(a.active, base, a.nrhs);
            const int c = __popcll(m);
            if (k0 >= k && k0 < k + c)
                r0 = base + nth_set_bit(m, k0 - k);
            if (k1 - 1 < k + c) {
                r1 = base + nth_set_bit(m, k1 - 1 - k) + 1;
                break;
            }
            k += c;
        }
    }
    loop_block<DAG, MODE, K, LF>(a, g, sb, cb_first, nparity, il, r0, r1);
}

template <bool DAG, int K, int LF>
void launch_mode(wd_ctx *ctx, const SloppyArgs &a, const Geom &g, int cb_first, int nparity,
                 int mode, int grid, int per)
{
    switch (mode) {
    case HOP_ONLY:
        hipLaunchKernelGGL((k_hop_sloppy_loop<DAG, HOP_ONLY, K, LF>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, a, g, cb_first, nparity, ctx->xcd_remap, per, ctx->parity_il);
        break;
    case HOP_M:
        hipLaunchKernelGGL((k_hop_sloppy_loop<DAG, HOP_M, K, LF>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, a, g, cb_first, nparity, ctx->xcd_remap, per, ctx->parity_il);
        break;
    case HOP_M_DOT_NORM:
        if constexpr (!DAG)
            hipLaunchKernelGGL((k_hop_sloppy_loop<false, HOP_M_DOT_NORM, K, LF>), dim3(grid),
                               dim3(kBlock), 0, ctx->stream, a, g, cb_first, nparity,
                               ctx->xcd_remap, per, ctx->parity_il);
        break;
    case HOP_M_NORM:
        if constexpr (!DAG)
            hipLaunchKernelGGL((k_hop_sloppy_loop<false, HOP_M_NORM, K, LF>), dim3(grid),
                               dim3(kBlock), 0, ctx->stream, a, g, cb_first, nparity,
                               ctx->xcd_remap, per, ctx->parity_il);
        break;
    case HOP_M_AXPY:
        if constexpr (DAG)
            hipLaunchKernelGGL((k_hop_sloppy_loop<true, HOP_M_AXPY, K, LF>), dim3(grid),
                               dim3(kBlock), 0, ctx->stream, a, g, cb_first, nparity,
                               ctx->xcd_remap, per, ctx->parity_il);
        break;
    default:
        hipLaunchKernelGGL((k_hop_sloppy_loop<DAG, HOP_M_DOT, K, LF>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, a, g, cb_first, nparity, ctx->xcd_remap, per, ctx->parity_il);
        break;
    }
}

} // namespace

// Automatic system group size of the loop kernel.  Measured on MI355X: 48
// systems at 24^3x48, sloppy M 1.382 ms in one group, 1.290 ms in 4 groups of
// 12 (4-16 per group within 1 %), batched Schur CG 0.605 -> 0.589 s and
// BiCGStab 0.335 -> 0.327 s; 12 systems at 32^3x64 and at 48^3x96 in one
// group (2 groups of 6: CG +3 % / 0 %).
constexpr int kLoopGroup = 12;

int wd_loop_group_cap(const wd_ctx *ctx, int nrhs, int chunk)
{
    const int cap = chunk > 0 ? chunk : ctx->loop_systems > 0 ? ctx->loop_systems : kLoopGroup;
    return nrhs < 1 ? 1 : min(cap, nrhs);
}

// Called by wd_launch_hop_sloppy (which has checked the arguments and filled
// bpp, nsb, tlinks and one) for R = WD_RHS_LOOP.
int wd_launch_hop_sloppy_loop(wd_ctx *ctx, const SloppyArgs &a, const Geom &g, int dagger,
                              int cb, int mode, int gauge_prec)
{
    const int nparity = (cb < 0) ? 2 : 1;
    const int cb_first = (cb < 0) ? 0 : cb;
    // at most `per` systems per group; the kernel balances the groups over the
    // systems each launch applies
    const int per = wd_loop_group_cap(ctx, a.nrhs, a.chunk);
    const int grid = a.nsb * ((a.nrhs + per - 1) / per);
    if (gauge_prec == WD_F32) {
        // F32 links: one direction ahead at 2 wavefronts per SIMD for every mode
        if (dagger)
            launch_mode<true, 1, LL_F32>(ctx, a, g, cb_first, nparity, mode, grid, per);
        else
            launch_mode<false, 1, LL_F32>(ctx, a, g, cb_first, nparity, mode, grid, per);
        return wd_check_launch(ctx, "k_hop_sloppy_loop");
    }
    // automatic prefetch: 1 ahead at 3 wavefronts per SIMD for the hop and M
    // launches; 3 ahead at 2 for the dot launches, which spill 20-23
    // registers at 3 wavefronts per SIMD, and for the axpy launch (which fits
    // at 3 with one direction ahead; measured on MI355X, 48 systems at
    // 24^3x48, Schur CG 0.508 s either way, 3 ahead within noise)
    const bool dot = mode == HOP_M_DOT || mode == HOP_M_DOT_NORM ||
                     mode == HOP_M_AXPY; // HOP_M_NORM fits at 3

    const int K = ctx->loop_prefetch ? ctx->loop_prefetch : dot ? 3 : 1;
    if (K == 3) {
        if (dagger)
            launch_mode<true, 3, LL_F16>(ctx, a, g, cb_first, nparity, mode, grid, per);
        else
            launch_mode<false, 3, LL_F16>(ctx, a, g, cb_first, nparity, mode, grid, per);
    } else {
        if (dagger)
            launch_mode<true, 1, LL_F16>(ctx, a, g, cb_first, nparity, mode, grid, per);
        else
            launch_mode<false, 1, LL_F16>(ctx, a, g, cb_first, nparity, mode, grid, per);
    }
    return wd_check_launch(ctx, "k_hop_sloppy_loop");
}
