// wd_cg_mixed.hip -- mixed-precision CG with reliable updates on many
// right-hand sides (several masses allowed), M^dagger M x = M^dagger b.
//
// Inner iterations run on the FP32-arithmetic operator with F16 (or F32)
// spinor storage and compressed links (wd_hop_sloppy.hip).  FP64 is kept for
// the global reductions, the solution (x += x_sloppy at every reliable update)
// and the true residual r = M^dagger b - M^dagger M x, recomputed with the
// FP64 operator whenever the sloppy residual has dropped by `delta` since the
// last update (reliable updates: Sleijpen & van der Vorst, "Reliable updated
// residuals in hybrid Bi-CG methods", Computing 56 (1996); the form used for
// lattice QCD by Clark et al., Comput. Phys. Commun. 181 (2010) 1517).  A
// system has converged when the FP64 true residual is below tol * |M^dagger b|.
//
// Every system carries its own scalars and flags in device memory; the
// launches form a fixed sequence in which converged systems and systems
// without a pending reliable update are skipped by whole workgroups, so the
// host only polls the flags every `check_every` iterations and the sequence
// can be replayed from a captured hipGraph.  An iteration:
//   xpay    p = r_s + beta p                (into the next of npend direction buffers)
//   M       qq = M p, partial |qq|^2       (sloppy, all iterating systems in one launch)
//   alpha   alpha = rho / |qq|^2            ((p, M^dag M p) = |M p|^2, as in CGLS)
//   M^dag   r_s -= alpha M^dag qq, partial |r_s|^2 (sloppy, HOP_M_AXPY: the
//           update in the operator launch, q = M^dag qq is never stored;
//           x_s += alpha p is left pending)
//   decide  reliable update due?  then the system waits (active = 2), else
//           beta = |r_s|^2 / rho, rho = |r_s|^2
// and every `reliable_every` iterations (and at the end of each block) a
// reliable-update slot:
//   flush   x_s += the pending alpha_j p_j (all systems); then
//           x += x_s, x_s = 0                             (waiting systems)
//   M, M^dag (FP64) on x                                   (waiting systems)
//   relr    r = M^dag b - M^dag M x -> r_s, partial |r|^2 (waiting systems)
//   relfin  converged?  else beta = |r|^2 / rho, rho = |r|^2, iterate again
//   alldone stop flag when no system is iterating or waiting
// A waiting system skips the iterations up to the slot, so its iterates and
// iteration count are those of an update in the same iteration; the slots
// only save the launches (each a full grid of skipped workgroups) that an
// update check after every iteration would cost.  Groups longer than kPend
// iterations (reliable_every > kPend) also end with a flush.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

#include "wd_blas.h"
#include "wd_fields.h"
#include "wd_internal.h"
#include "wd_sloppy.h"

using namespace wdf;

namespace {

// Temporal-gauge frame of the sloppy links (WD_RECON_12T): v <- Omega v
// (dag = false: FP64 residual into the sloppy frame) or v <- Omega^dag v
// (dag = true: sloppy correction back), Omega of site i of parity par, stored
// complex [parity][9][stride].  omega == nullptr: the identity.
template <typename T>
__device__ __forceinline__ void omega_rotate(const double2 *__restrict__ omega, int par, int i,
                                             int stride, bool dag, cpx<T> (&v)[4][3])
{
    if (!omega)
        return;
    const double2 *o = omega + (size_t)par * 9 * stride + i;
    cpx<double> W[3][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double2 e = o[(size_t)k * stride];
        W[k / 3][k % 3] = {e.x, e.y};
    }
#pragma unroll
    for (int sp = 0; sp < 4; ++sp) {
        cpx<double> r[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            cpx<double> acc = {0.0, 0.0};
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const cpx<double> vb = {(double)v[sp][b].r, (double)v[sp][b].i};
                acc = dag ? cfma_conj(acc, W[b][a], vb) : cfma(acc, W[a][b], vb);
            }
            r[a] = acc;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
            v[sp][a] = {(T)r[a].r, (T)r[a].i};
    }
}

// The elementwise kernels stream whole fields with no reuse inside a launch:
// non-temporal loads and stores keep them from churning L2.
constexpr bool kStreamNT = true;

// Mixed CG keeps the search directions of up to kPend consecutive iterations
// in separate buffers and adds their steps alpha_j p_j to x_s in one pass at
// the end of each group (a reliable-update slot or kPend iterations): x_s is
// read and written once per group instead of once per iteration.
constexpr int kPend = 8;

struct MxState {
    double rho;      // |r|^2 carried by the recursion
    double alpha, beta;
    double maxr;     // |r| at the last reliable update
    double target;   // tol * |M^dag b|
    double rsd;      // last sloppy |r_s|
    double true_rsd; // |M^dag b - M^dag M x| at the last reliable update
    double b_norm;   // |M^dag b| (even-odd: of the Schur system, |A^dag bh|)
    double full_b;   // even-odd: |M^dag b| of the full system
    double full_rsd; // even-odd: |M^dag b - M^dag M x| at the last full check
    int iter, status, updates; // status: 0 running, 1 converged, 2 breakdown, 3 limit
    int pbuf;                  // direction buffer holding the latest p
    unsigned pmask;            // bit j: x_s += ahist[j] p_j still pending
    float ahist[kPend];        // alpha of the pending steps, by buffer
};

__device__ __forceinline__ double2 sum_rhs_partials(const double2 *partials, int count)
{
    return sum_partials_1024(partials + (size_t)blockIdx.x * count, count);
}

// ---- elementwise kernels, blockIdx.y = rhs, one thread per (parity, site) ----
// p_pos = r + beta p, p read from the buffer of the system's latest direction
// (st.pbuf; the same buffer when pbuf == pos: each thread reads its element
// before writing it).  Buffer j of the directions starts at pb + j * pbytes
// (scales at ps + j * pscale).
template <int SF>
__global__ __launch_bounds__(kBlock) void k_mx_xpay(char *pb, float *ps, size_t pbytes,
                                                    size_t pscale, int pos,
                                                    const char *__restrict__ r,
                                                    const float *__restrict__ rs,
                                                    size_t rhs_bytes, int stride, int nel,
                                                    const MxState *__restrict__ st,
                                                    const int *__restrict__ active)
{
    const int rhs = blockIdx.y;
    if (active[rhs] != 1)
        return;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nel)
        return;
    const int par = e / stride, i = e - par * stride;
    const size_t off = rhs * rhs_bytes + (size_t)par * spinor_chunks(SF) * 16 * stride;
    const size_t so = ((size_t)rhs * 2 + par) * stride;
    const int src = st[rhs].pbuf;
    const float beta = (float)st[rhs].beta;
    cpx<float> vp[4][3], vr[4][3];
    Spinor<SF>::template load<float, kStreamNT>(pb + src * pbytes + off,
                                                ps ? ps + src * pscale + so : nullptr, stride,
                                                (unsigned)i, vp);
    Spinor<SF>::template load<float, kStreamNT>(r + off, rs ? rs + so : nullptr, stride, (unsigned)i, vr);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            vp[s][c] = {fmaf(beta, vp[s][c].r, vr[s][c].r), fmaf(beta, vp[s][c].i, vr[s][c].i)};
    Spinor<SF>::template store<float, kStreamNT>(pb + pos * pbytes + off,
                                                 ps ? ps + pos * pscale + so : nullptr, stride,
                                                 (unsigned)i, vp);
}

// The packed words and scale of an F16 site (converted where they are used,
// so a group's pending directions can be in flight together in few registers)
struct Raw16 {
    unsigned w[12];
    float s;
};
__device__ __forceinline__ void load_raw16(const char *base, const float *scale, int stride,
                                           unsigned i, Raw16 &v)
{
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const u4v c = ldg<kStreamNT>(
            reinterpret_cast<const u4v *>(base + ((size_t)j * (unsigned)stride + i) * 16));
        v.w[4 * j] = c.x;
        v.w[4 * j + 1] = c.y;
        v.w[4 * j + 2] = c.z;
        v.w[4 * j + 3] = c.w;
    }
    v.s = ldg<kStreamNT>(scale + i);
}
// the values Spinor<WD_F16>::load returns for the same site
__device__ __forceinline__ void unpack_raw16(const Raw16 &v, cpx<float> (&out)[4][3])
{
#pragma unroll
    for (int k = 0; k < 12; ++k)
        out[k / 3][k % 3] = h2f(v.w[k], v.s);
}

// x (F64) of site i, parity par += Omega^dag v (v: x_s in the sloppy frame,
// omega == nullptr: the identity); x_s = 0
__device__ __forceinline__ void fold_into_x(double2 *__restrict__ xp, char *__restrict__ sp,
                                            const double2 *__restrict__ omega, int par, int i,
                                            int stride, cpx<float> (&vs)[4][3])
{
    cpx<double> vx[4][3];
    cpx<float> z[4][3];
    Spinor<WD_F64>::load(xp, nullptr, stride, (unsigned)i, vx);
    omega_rotate(omega, par, i, stride, true, vs);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            vx[s][c].r += (double)vs[s][c].r;
            vx[s][c].i += (double)vs[s][c].i;
            z[s][c] = {0.f, 0.f};
        }
    Spinor<WD_F64>::store(xp, nullptr, stride, (unsigned)i, vx);
    Spinor<WD_F32>::store(sp, nullptr, stride, (unsigned)i, z);
}

// Mixed CG, end of a group of iterations: x_s += alpha_j p_j for the pending
// steps of each system (st.pmask, in buffer order = iteration order, the same
// fmaf sequence as one update per iteration); then, for the systems due
// (reliable[rhs]; reliable == nullptr: none), x += Omega^dag x_s and x_s = 0.
template <int SF>
__global__ __launch_bounds__(kBlock) void k_mx_flush(double2 *__restrict__ x,
                                                     char *__restrict__ xs, size_t x_rhs_elems,
                                                     size_t xs_rhs_bytes, int stride, int nel,
                                                     const int *__restrict__ reliable,
                                                     const double2 *__restrict__ omega,
                                                     const char *__restrict__ pb,
                                                     const float *__restrict__ ps, size_t pbytes,
                                                     size_t pscale, size_t rhs_bytes,
                                                     const MxState *__restrict__ st)
{
    const int rhs = blockIdx.y;
    const bool due = reliable && reliable[rhs];
    const unsigned pm = st[rhs].pmask;
    if (!due && !pm)
        return;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nel)
        return;
    const int par = e / stride, i = e - par * stride;
    char *sp = xs + rhs * xs_rhs_bytes + (size_t)par * spinor_chunks(WD_F32) * 16 * stride;
    const size_t off = rhs * rhs_bytes + (size_t)par * spinor_chunks(SF) * 16 * stride;
    const size_t so = ((size_t)rhs * 2 + par) * stride;
    cpx<float> vs[4][3];
    Spinor<WD_F32>::template load<float, kStreamNT>(sp, nullptr, stride, (unsigned)i, vs);
    auto step = [&](int j, const cpx<float> (&vp)[4][3]) {
        const float alpha = st[rhs].ahist[j];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                vs[s][c] = {fmaf(alpha, vp[s][c].r, vs[s][c].r), fmaf(alpha, vp[s][c].i, vs[s][c].i)};
    };
    if constexpr (SF == WD_F16) {
        // F16: four pending directions at a time, their loads issued together
        // (one at a time, each load waited for the previous step)
#pragma unroll
        for (int j0 = 0; j0 < kPend; j0 += 4) {
            Raw16 raw[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if ((pm >> (j0 + u)) & 1u)
                    load_raw16(pb + (j0 + u) * pbytes + off, ps + (j0 + u) * pscale + so, stride,
                               (unsigned)i, raw[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if ((pm >> (j0 + u)) & 1u) {
                    cpx<float> vp[4][3];
                    unpack_raw16(raw[u], vp);
                    step(j0 + u, vp);
                }
        }
    } else {
        for (int j = 0; j < kPend; ++j) {
            if (!((pm >> j) & 1u))
                continue;
            cpx<float> vp[4][3];
            Spinor<SF>::template load<float, kStreamNT>(pb + j * pbytes + off, nullptr, stride,
                                                        (unsigned)i, vp);
            step(j, vp);
        }
    }
    if (due)
        fold_into_x(x + rhs * x_rhs_elems + (size_t)par * 12 * stride, sp, omega, par, i, stride,
                    vs);
    else
        Spinor<WD_F32>::template store<float, kStreamNT>(sp, nullptr, stride, (unsigned)i, vs);
}

// r = bp - q64 (F64) -> r_s (SF, rotated into the sloppy frame by omega when
// not null); partial |r|^2.  With reliable == nullptr every rhs runs
// (initialisation).
template <int SF>
__global__ __launch_bounds__(kBlock) void k_mx_relr(char *__restrict__ rsl, float *__restrict__ rss,
                                                    const double2 *__restrict__ bp,
                                                    const double2 *__restrict__ q64,
                                                    size_t x_rhs_elems, size_t rhs_bytes,
                                                    int stride, int nel,
                                                    const int *__restrict__ reliable,
                                                    double2 *__restrict__ partials,
                                                    const double2 *__restrict__ omega)
{
    const int rhs = blockIdx.y;
    if (reliable && !reliable[rhs])
        return;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    double acc = 0.0;
    if (e < nel) {
        const int par = e / stride, i = e - par * stride;
        const size_t xo = rhs * x_rhs_elems + (size_t)par * 12 * stride;
        cpx<double> vb[4][3], vq[4][3];
        Spinor<WD_F64>::load(bp + xo, nullptr, stride, (unsigned)i, vb);
        if (q64) {
            Spinor<WD_F64>::load(q64 + xo, nullptr, stride, (unsigned)i, vq);
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    vb[s][c] = {vb[s][c].r - vq[s][c].r, vb[s][c].i - vq[s][c].i};
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                acc = fma(vb[s][c].r, vb[s][c].r, acc);
                acc = fma(vb[s][c].i, vb[s][c].i, acc);
            }
        const size_t off = rhs * rhs_bytes + (size_t)par * spinor_chunks(SF) * 16 * stride;
        const size_t so = ((size_t)rhs * 2 + par) * stride;
        omega_rotate(omega, par, i, stride, false, vb);
        Spinor<SF>::store(rsl + off, rss ? rss + so : nullptr, stride, (unsigned)i, vb);
    }
    block_reduce_store(acc, 0.0, partials + (size_t)rhs * gridDim.x + blockIdx.x);
}

// ---- per-rhs scalar kernels: blockIdx.x = rhs, 1024 threads -------------------
__global__ __launch_bounds__(1024) void k_mx_init(const double2 *__restrict__ partials, int count,
                                                  MxState *__restrict__ st, int *__restrict__ active,
                                                  int *__restrict__ reliable, double tol)
{
    const double2 s = sum_rhs_partials(partials, count);
    if (threadIdx.x == 0) {
        const int r = blockIdx.x;
        MxState m = {};
        m.rho = s.x;
        m.b_norm = sqrt(s.x);
        m.maxr = m.b_norm;
        m.target = tol * m.b_norm;
        m.rsd = m.b_norm;
        m.true_rsd = m.b_norm;
        m.beta = 0.0;
        st[r] = m;
        reliable[r] = 0;
        // a zero right-hand side is solved by x = 0
        if (!(m.b_norm > 0.0)) {
            st[r].status = 1;
            active[r] = 0;
        } else {
            active[r] = 1;
        }
    }
}

// alpha of the iteration whose direction went into buffer pos; the step is
// recorded as pending for the group's k_mx_flush (pos == 0 opens a group:
// every system's pending set was flushed at the end of the last one)
__global__ __launch_bounds__(1024) void k_mx_alpha(const double2 *__restrict__ partials, int count,
                                                   MxState *__restrict__ st,
                                                   int *__restrict__ active, int pos,
                                                   float *__restrict__ alpha_f)
{
    const int r = blockIdx.x;
    if (pos == 0 && threadIdx.x == 0)
        st[r].pmask = 0u;
    if (active[r] != 1)
        return;
    const double2 pq = sum_rhs_partials(partials, count);
    if (threadIdx.x == 0) {
        MxState &m = st[r];
        if (!isfinite(pq.x) || !(pq.x > 0.0)) {
            m.status = 2;
            active[r] = 0;
            return;
        }
        m.alpha = m.rho / pq.x;
        alpha_f[r] = (float)m.alpha; // the operator's r -= alpha q
        m.ahist[pos] = (float)m.alpha;
        m.pmask |= 1u << pos;
        m.pbuf = pos;
    }
}

__global__ __launch_bounds__(1024) void k_mx_decide(const double2 *__restrict__ partials,
                                                    int count, MxState *__restrict__ st,
                                                    int *__restrict__ active,
                                                    int *__restrict__ reliable, int max_iter,
                                                    double delta)
{
    const int r = blockIdx.x;
    if (active[r] != 1)
        return;
    const double2 s = sum_rhs_partials(partials, count);
    if (threadIdx.x == 0) {
        MxState &m = st[r];
        const double rr = s.x, rsd = sqrt(rr);
        m.iter += 1;
        m.rsd = rsd;
        if (rsd < delta * m.maxr || rsd < m.target || m.iter >= max_iter || !isfinite(rr)) {
            reliable[r] = 1;
            active[r] = 2; // waits for the next reliable-update slot
        } else {
            m.beta = rr / m.rho;
            m.rho = rr;
        }
    }
}

__global__ __launch_bounds__(1024) void k_mx_relfin(const double2 *__restrict__ partials,
                                                    int count, MxState *__restrict__ st,
                                                    int *__restrict__ active,
                                                    int *__restrict__ reliable, int max_iter)
{
    const int r = blockIdx.x;
    if (!reliable[r])
        return;
    const double2 s = sum_rhs_partials(partials, count);
    if (threadIdx.x == 0) {
        MxState &m = st[r];
        const double rr = s.x, rsd = sqrt(rr);
        m.true_rsd = rsd;
        m.updates += 1;
        reliable[r] = 0;
        if (rsd < m.target) {
            m.status = 1;
            active[r] = 0;
        } else if (m.iter >= max_iter || !isfinite(rr)) {
            m.status = isfinite(rr) ? 3 : 2;
            active[r] = 0;
        } else {
            m.beta = rr / m.rho;
            m.rho = rr;
            m.maxr = rsd;
            active[r] = 1;
        }
    }
}

// ---- even-odd: checks on the full system ------------------------------------------
// partial |bp - q|^2 (|bp|^2 when q == nullptr) over both parities, systems
// with mask[r] (all when mask == nullptr)
__global__ __launch_bounds__(kBlock) void k_mx_fullres(const double2 *__restrict__ bp,
                                                       const double2 *__restrict__ q,
                                                       size_t x_rhs_elems, int stride,
                                                       const int *__restrict__ mask,
                                                       double2 *__restrict__ partials)
{
    const int rhs = blockIdx.y;
    if (mask && !mask[rhs])
        return;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    double acc = 0.0;
    if (e < 2 * stride) {
        const int par = e / stride, i = e - par * stride;
        const size_t xo = rhs * x_rhs_elems + (size_t)par * 12 * stride;
        cpx<double> vb[4][3], vq[4][3];
        Spinor<WD_F64>::load(bp + xo, nullptr, stride, (unsigned)i, vb);
        if (q)
            Spinor<WD_F64>::load(q + xo, nullptr, stride, (unsigned)i, vq);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double dr = q ? vb[s][c].r - vq[s][c].r : vb[s][c].r;
                const double di = q ? vb[s][c].i - vq[s][c].i : vb[s][c].i;
                acc = fma(dr, dr, acc);
                acc = fma(di, di, acc);
            }
    }
    block_reduce_store(acc, 0.0, partials + (size_t)rhs * gridDim.x + blockIdx.x);
}

// full_b = |M^dag b| per system
__global__ __launch_bounds__(1024) void k_mx_fullb(const double2 *__restrict__ partials, int count,
                                                   MxState *__restrict__ st)
{
    const double2 s = sum_rhs_partials(partials, count);
    if (threadIdx.x == 0)
        st[blockIdx.x].full_b = sqrt(s.x);
}

// Systems that met the Schur target (check[r]): accept when the full-system
// residual meets tol, else resume them with a Schur target tightened by the
// observed ratio (the state is that of a reliable update: r_s is the true
// Schur residual, x_s = 0).  resume = 0 (the final pass) only records the
// full-system residual.
__global__ __launch_bounds__(1024) void k_mx_fullfin(const double2 *__restrict__ partials,
                                                     int count, MxState *__restrict__ st,
                                                     int *__restrict__ active,
                                                     int *__restrict__ check, double tol,
                                                     int max_iter, int resume)
{
    const int r = blockIdx.x;
    if (!check[r])
        return;
    const double2 s = sum_rhs_partials(partials, count);
    if (threadIdx.x == 0) {
        MxState &m = st[r];
        const double full = sqrt(s.x);
        m.full_rsd = full;
        check[r] = 0;
        if (!resume || full <= tol * m.full_b || !(m.full_b > 0.0) || m.iter >= max_iter ||
            !isfinite(full))
            return;
        const double rr = m.true_rsd * m.true_rsd;
        m.target *= 0.5 * tol * m.full_b / full;
        m.beta = rr / m.rho;
        m.rho = rr;
        m.maxr = m.true_rsd;
        m.status = 0;
        active[r] = 1;
    }
}

// check[r] = 1 for the systems the Schur iteration declared converged
__global__ void k_mx_markconv(const MxState *__restrict__ st, int nrhs, int *__restrict__ check)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrhs)
        check[r] = st[r].status == 1 ? 1 : 0;
}

// Final pass of the Schur solvers: check[r] = 1 for every system whose x_o and
// full-system residual do not yet belong to its final x_e -- systems stopped
// at the iteration limit or by a breakdown (never checked), and systems that
// a failed check resumed and that stopped after the last check.  Accepted
// systems (converged and verified on the full system) keep their x_o.
template <class State>
__global__ void k_mark_final(const State *__restrict__ st, int nrhs, double tol,
                             int *__restrict__ check)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrhs) {
        const State &m = st[r];
        const bool accepted =
            m.status == 1 && (!(m.full_b > 0.0) || m.full_rsd <= tol * m.full_b);
        check[r] = accepted ? 0 : 1;
    }
}

__global__ void k_mx_alldone(const int *__restrict__ active, int nrhs, int *__restrict__ stop)
{
    if (threadIdx.x == 0) {
        int any = 0;
        for (int r = 0; r < nrhs; ++r)
            any |= active[r];
        *stop = any ? 0 : 1;
    }
}


// ---- mixed BiCGStab on the Schur system ---------------------------------------
// Per-system state of wd_bicgstab_mixed (complex recursion scalars as (re, im)).
struct BsState {
    double2 rho, alpha, omega, beta;
    double2 r0r;     // (r0, r) of the last reliable update (restarts after a full check)
    double maxr, target, rsd, true_rsd, b_norm, full_b, full_rsd;
    int iter, status, updates; // status: 0 running, 1 converged, 2 breakdown, 3 limit
    int pbuf;                  // direction buffer holding the latest p
    unsigned pmask;            // bit j: x_s += alpha_j p_j + omega_j s_j still pending
    float ahist[kPend][2], whist[kPend][2]; // (re, im) of alpha_j and omega_j, by buffer
};

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zdiv(double2 a, double2 b)
{
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ bool zfinite(double2 a) { return isfinite(a.x) && isfinite(a.y); }

// common prologue of the elementwise kernels: (parity, site) e < nel of system
// blockIdx.y; returns false for idle threads
#define BS_ELEM_PROLOGUE(SFMT)                                                                     \
    const int rhs = blockIdx.y;                                                                    \
    const int e = blockIdx.x * kBlock + threadIdx.x;                                              \
    const int par = e / stride, i = e - par * stride;                                              \
    const size_t off = rhs * rhs_bytes + (size_t)par * spinor_chunks(SFMT) * 16 * stride;         \
    const size_t so = ((size_t)rhs * 2 + par) * stride;

// p_pos = r + beta (p - omega v), p read from the buffer of the system's latest
// direction (st.pbuf; in place when pbuf == pos).  Direction buffer j starts
// at pb + j * pbytes (scales at ps + j * pscale).
template <int SF>
__global__ __launch_bounds__(kBlock) void k_bs_p(char *pb, float *ps, size_t pbytes,
                                                 size_t pscale, int pos,
                                                 const char *__restrict__ r,
                                                 const float *__restrict__ rs,
                                                 const char *__restrict__ v,
                                                 const float *__restrict__ vs, size_t rhs_bytes,
                                                 int stride, int nel, const BsState *__restrict__ st,
                                                 const int *__restrict__ active)
{
    if (active[blockIdx.y] != 1 || blockIdx.x * kBlock + threadIdx.x >= nel)
        return;
    BS_ELEM_PROLOGUE(SF)
    const int src = st[rhs].pbuf;
    const float br = (float)st[rhs].beta.x, bi = (float)st[rhs].beta.y;
    const float wr = (float)st[rhs].omega.x, wi = (float)st[rhs].omega.y;
    cpx<float> vp[4][3], vr[4][3], vv[4][3];
    Spinor<SF>::template load<float, kStreamNT>(pb + src * pbytes + off,
                                                ps ? ps + src * pscale + so : nullptr, stride,
                                                (unsigned)i, vp);
    Spinor<SF>::template load<float, kStreamNT>(r + off, rs ? rs + so : nullptr, stride, (unsigned)i, vr);
    Spinor<SF>::template load<float, kStreamNT>(v + off, vs ? vs + so : nullptr, stride, (unsigned)i, vv);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // u = p - omega v; p = r + beta u
            const float ur = vp[s][c].r - (wr * vv[s][c].r - wi * vv[s][c].i);
            const float ui = vp[s][c].i - (wr * vv[s][c].i + wi * vv[s][c].r);
            vp[s][c] = {vr[s][c].r + (br * ur - bi * ui), vr[s][c].i + (br * ui + bi * ur)};
        }
    Spinor<SF>::template store<float, kStreamNT>(pb + pos * pbytes + off,
                                                 ps ? ps + pos * pscale + so : nullptr, stride,
                                                 (unsigned)i, vp);
}

// s = r - alpha v
template <int SF>
__global__ __launch_bounds__(kBlock) void k_bs_s(char *__restrict__ sv, float *__restrict__ ss,
                                                 const char *__restrict__ r,
                                                 const float *__restrict__ rs,
                                                 const char *__restrict__ v,
                                                 const float *__restrict__ vs, size_t rhs_bytes,
                                                 int stride, int nel, const BsState *__restrict__ st,
                                                 const int *__restrict__ active)
{
    if (active[blockIdx.y] != 1 || blockIdx.x * kBlock + threadIdx.x >= nel)
        return;
    BS_ELEM_PROLOGUE(SF)
    const float ar = (float)st[rhs].alpha.x, ai = (float)st[rhs].alpha.y;
    cpx<float> vr[4][3], vv[4][3];
    Spinor<SF>::template load<float, kStreamNT>(r + off, rs ? rs + so : nullptr, stride, (unsigned)i, vr);
    Spinor<SF>::template load<float, kStreamNT>(v + off, vs ? vs + so : nullptr, stride, (unsigned)i, vv);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            vr[s][c] = {vr[s][c].r - (ar * vv[s][c].r - ai * vv[s][c].i),
                        vr[s][c].i - (ar * vv[s][c].i + ai * vv[s][c].r)};
    Spinor<SF>::template store<float, kStreamNT>(sv + off, ss ? ss + so : nullptr, stride, (unsigned)i, vr);
}

// x_s += alpha p + omega s (one step; the expression k_bs_flush applies per
// pending step)
__device__ __forceinline__ void bs_xstep(cpx<float> (&vx)[4][3], const cpx<float> (&vp)[4][3],
                                         const cpx<float> (&vs)[4][3], float ar, float ai,
                                         float wr, float wi)
{
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // explicit FMAs: the same rounding wherever the step is inlined
            const cpx<float> P = vp[s][c], S = vs[s][c];
            const float pr = fmaf(ar, P.r, -(ai * P.i)), pi = fmaf(ar, P.i, ai * P.r);
            const float sr = fmaf(wr, S.r, -(wi * S.i)), si = fmaf(wr, S.i, wi * S.r);
            vx[s][c] = {(vx[s][c].r + pr) + sr, (vx[s][c].i + pi) + si};
        }
}

// r = s - omega t; partials |r|^2 (pa) and (r0, r) (pb) of the stored r;
// x_s += alpha p + omega s is left pending for the group's k_bs_flush
template <int SF>
__global__ __launch_bounds__(kBlock) void k_bs_update(
    char *__restrict__ r, float *__restrict__ rs, const char *__restrict__ sv,
    const float *__restrict__ ss, const char *__restrict__ t, const float *__restrict__ ts,
    const char *__restrict__ r0, const float *__restrict__ r0s, size_t rhs_bytes, int stride,
    int nel, const BsState *__restrict__ st, const int *__restrict__ active,
    double2 *__restrict__ pa, double2 *__restrict__ pb)
{
    const int rhs = blockIdx.y;
    if (active[rhs] != 1)
        return;
    double rr = 0.0, dr = 0.0, di = 0.0;
    if (blockIdx.x * kBlock + threadIdx.x < nel) {
        BS_ELEM_PROLOGUE(SF)
        const float wr = (float)st[rhs].omega.x, wi = (float)st[rhs].omega.y;
        cpx<float> vs[4][3], vt[4][3];
        Spinor<SF>::template load<float, kStreamNT>(sv + off, ss ? ss + so : nullptr, stride,
                                               (unsigned)i, vs);
        Spinor<SF>::template load<float, kStreamNT>(t + off, ts ? ts + so : nullptr, stride, (unsigned)i,
                                               vt);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const cpx<float> S = vs[s][c], T = vt[s][c];
                vs[s][c] = {S.r - (wr * T.r - wi * T.i), S.i - (wr * T.i + wi * T.r)};
            }
        cpx<float> vq[4][3]; // r as stored
        if constexpr (SF == WD_F16) {
            unsigned wv[12];
            const float sc = Spinor<WD_F16>::template store_packed<float, kStreamNT>(
                r + off, rs + so, stride, (unsigned)i, vs, wv);
#pragma unroll
            for (int k = 0; k < 12; ++k)
                vq[k / 3][k % 3] = h2f(wv[k], sc);
        } else {
            Spinor<SF>::template store<float, kStreamNT>(r + off, nullptr, stride, (unsigned)i, vs);
#pragma unroll
            for (int k = 0; k < 12; ++k)
                vq[k / 3][k % 3] = vs[k / 3][k % 3];
        }
        cpx<float> v0[4][3];
        Spinor<SF>::template load<float, kStreamNT>(r0 + off, r0s ? r0s + so : nullptr, stride,
                                               (unsigned)i, v0);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double qr = vq[s][c].r, qi = vq[s][c].i, zr = v0[s][c].r, zi = v0[s][c].i;
                rr = fma(qr, qr, fma(qi, qi, rr));
                dr = fma(zr, qr, fma(zi, qi, dr));
                di = fma(zr, qi, fma(-zi, qr, di));
            }
    }
    block_reduce_store(rr, 0.0, pa + (size_t)rhs * gridDim.x + blockIdx.x);
    block_reduce_store(dr, di, pb + (size_t)rhs * gridDim.x + blockIdx.x);
}

// End of a group of BiCGStab iterations: x_s += alpha_j p_j + omega_j s_j for
// the pending steps of each system (st.pmask, in buffer order = iteration
// order: the arithmetic of one update per iteration); then, for the systems
// due (reliable[rhs]; reliable == nullptr: none), x += Omega^dag x_s and
// x_s = 0.  Buffer j of the directions p and of s starts at pb / sb + j * bytes
// (scales at ps / ss + j * scale).
template <int SF>
__global__ __launch_bounds__(kBlock) void k_bs_flush(
    double2 *__restrict__ x, char *__restrict__ xs, size_t x_rhs_elems, size_t xs_rhs_bytes,
    int stride, int nel, const int *__restrict__ reliable, const double2 *__restrict__ omega,
    const char *__restrict__ pb, const float *__restrict__ ps, const char *__restrict__ sb,
    const float *__restrict__ ss, size_t bytes, size_t scale, size_t rhs_bytes,
    const BsState *__restrict__ st)
{
    const bool due = reliable && reliable[blockIdx.y];
    const unsigned pm = st[blockIdx.y].pmask;
    if ((!due && !pm) || blockIdx.x * kBlock + threadIdx.x >= nel)
        return;
    BS_ELEM_PROLOGUE(SF)
    char *sp = xs + rhs * xs_rhs_bytes + (size_t)par * spinor_chunks(WD_F32) * 16 * stride;
    cpx<float> vx[4][3];
    Spinor<WD_F32>::template load<float, kStreamNT>(sp, nullptr, stride, (unsigned)i, vx);
    auto step = [&](int j, const cpx<float> (&vp)[4][3], const cpx<float> (&vs)[4][3]) {
        bs_xstep(vx, vp, vs, st[rhs].ahist[j][0], st[rhs].ahist[j][1], st[rhs].whist[j][0],
                 st[rhs].whist[j][1]);
    };
    if constexpr (SF == WD_F16) {
        // F16: two pending steps (p and s each) at a time, their loads issued
        // together
#pragma unroll
        for (int j0 = 0; j0 < kPend; j0 += 2) {
            Raw16 rp[2], rs[2];
#pragma unroll
            for (int u = 0; u < 2; ++u)
                if ((pm >> (j0 + u)) & 1u) {
                    load_raw16(pb + (j0 + u) * bytes + off, ps + (j0 + u) * scale + so, stride,
                               (unsigned)i, rp[u]);
                    load_raw16(sb + (j0 + u) * bytes + off, ss + (j0 + u) * scale + so, stride,
                               (unsigned)i, rs[u]);
                }
#pragma unroll
            for (int u = 0; u < 2; ++u)
                if ((pm >> (j0 + u)) & 1u) {
                    cpx<float> vp[4][3], vs[4][3];
                    unpack_raw16(rp[u], vp);
                    unpack_raw16(rs[u], vs);
                    step(j0 + u, vp, vs);
                }
        }
    } else {
        for (int j = 0; j < kPend; ++j) {
            if (!((pm >> j) & 1u))
                continue;
            cpx<float> vp[4][3], vs[4][3];
            Spinor<SF>::template load<float, kStreamNT>(pb + j * bytes + off, nullptr, stride,
                                                        (unsigned)i, vp);
            Spinor<SF>::template load<float, kStreamNT>(sb + j * bytes + off, nullptr, stride,
                                                        (unsigned)i, vs);
            step(j, vp, vs);
        }
    }
    if (due)
        fold_into_x(x + rhs * x_rhs_elems + (size_t)par * 12 * stride, sp, omega, par, i, stride,
                    vx);
    else
        Spinor<WD_F32>::template store<float, kStreamNT>(sp, nullptr, stride, (unsigned)i, vx);
}

// r = bp - q64 (F64; q64 == nullptr: r = bp) -> sloppy r (and r0 = r when r0w);
// partials |r|^2 (pa, FP64 values) and (r0, r) (pb, stored values); systems
// with mask[r] only when mask != nullptr
template <int SF>
__global__ __launch_bounds__(kBlock) void k_bs_relr(
    char *__restrict__ r, float *__restrict__ rs, char *__restrict__ r0, float *__restrict__ r0s,
    bool r0w, const double2 *__restrict__ bp, const double2 *__restrict__ q64, size_t x_rhs_elems,
    size_t rhs_bytes, int stride, int nel, const int *__restrict__ mask, double2 *__restrict__ pa,
    double2 *__restrict__ pb, const double2 *__restrict__ omega)
{
    const int rhs = blockIdx.y;
    if (mask && !mask[rhs])
        return;
    double rr = 0.0, dr = 0.0, di = 0.0;
    if (blockIdx.x * kBlock + threadIdx.x < nel) {
        BS_ELEM_PROLOGUE(SF)
        const size_t xo = rhs * x_rhs_elems + (size_t)par * 12 * stride;
        cpx<double> vb[4][3];
        Spinor<WD_F64>::load(bp + xo, nullptr, stride, (unsigned)i, vb);
        if (q64) {
            cpx<double> vq[4][3];
            Spinor<WD_F64>::load(q64 + xo, nullptr, stride, (unsigned)i, vq);
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    vb[s][c] = {vb[s][c].r - vq[s][c].r, vb[s][c].i - vq[s][c].i};
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                rr = fma(vb[s][c].r, vb[s][c].r, rr);
                rr = fma(vb[s][c].i, vb[s][c].i, rr);
            }
        omega_rotate(omega, par, i, stride, false, vb);
        Spinor<SF>::store(r + off, rs ? rs + so : nullptr, stride, (unsigned)i, vb);
        if (r0w)
            Spinor<SF>::store(r0 + off, r0s ? r0s + so : nullptr, stride, (unsigned)i, vb);
        cpx<float> vq[4][3], v0[4][3];
        Spinor<SF>::template load<float, kStreamNT>(r + off, rs ? rs + so : nullptr, stride, (unsigned)i,
                                               vq);
        Spinor<SF>::template load<float, kStreamNT>(r0 + off, r0s ? r0s + so : nullptr, stride,
                                               (unsigned)i, v0);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double qr = vq[s][c].r, qi = vq[s][c].i, zr = v0[s][c].r, zi = v0[s][c].i;
                dr = fma(zr, qr, fma(zi, qi, dr));
                di = fma(zr, qi, fma(-zi, qr, di));
            }
    }
    block_reduce_store(rr, 0.0, pa + (size_t)rhs * gridDim.x + blockIdx.x);
    block_reduce_store(dr, di, pb + (size_t)rhs * gridDim.x + blockIdx.x);
}
#undef BS_ELEM_PROLOGUE

// ---- per-system scalar kernels (blockIdx.x = system, 1024 threads) ----
__device__ __forceinline__ double2 sum_at(const double2 *partials, int count)
{
    return sum_partials_1024(partials + (size_t)blockIdx.x * count, count);
}

__global__ __launch_bounds__(1024) void k_bs_init(const double2 *__restrict__ pa,
                                                  const double2 *__restrict__ pb, int count,
                                                  BsState *__restrict__ st, int *__restrict__ active,
                                                  int *__restrict__ reliable, double tol)
{
    const double2 rr = sum_at(pa, count);
    const double2 r0r = sum_at(pb, count);
    if (threadIdx.x == 0) {
        const int r = blockIdx.x;
        BsState m = {};
        m.b_norm = sqrt(rr.x);
        m.maxr = m.b_norm;
        m.rsd = m.true_rsd = m.b_norm;
        m.target = tol * m.b_norm;
        m.rho = m.r0r = r0r;
        m.alpha = m.omega = make_double2(1.0, 0.0);
        m.beta = make_double2(0.0, 0.0);
        reliable[r] = 0;
        if (!(m.b_norm > 0.0)) { // x = 0 solves b = 0
            m.status = 1;
            active[r] = 0;
        } else {
            active[r] = 1;
        }
        st[r] = m;
    }
}

// alpha = rho / (r0, v) of the iteration whose direction went into buffer pos
// (pos == 0 opens a group: every pending step was flushed at the end of the
// last one)
__global__ __launch_bounds__(1024) void k_bs_alpha(const double2 *__restrict__ pa, int count,
                                                   BsState *__restrict__ st,
                                                   int *__restrict__ active, int pos)
{
    const int r = blockIdx.x;
    if (pos == 0 && threadIdx.x == 0)
        st[r].pmask = 0u;
    if (active[r] != 1)
        return;
    const double2 d = sum_at(pa, count);
    if (threadIdx.x == 0) {
        st[r].pbuf = pos;
        if (!zfinite(d) || !(d.x * d.x + d.y * d.y > 0.0)) {
            st[r].status = 2;
            active[r] = 0;
            return;
        }
        st[r].alpha = zdiv(st[r].rho, d);
    }
}

// omega = (t, s) / (t, t) = conj((s, t)) / |t|^2; t = 0 (s = 0: the system is
// solved) gives omega = 0.  The step x_s += alpha p + omega s of the iteration
// (buffer pos) is recorded as pending.
__global__ __launch_bounds__(1024) void k_bs_omega(const double2 *__restrict__ pst, int cst,
                                                   const double2 *__restrict__ ptt, int ctt,
                                                   BsState *__restrict__ st,
                                                   int *__restrict__ active, int pos)
{
    const int r = blockIdx.x;
    if (active[r] != 1)
        return;
    const double2 s_t = sum_at(pst, cst);
    const double2 tt = sum_at(ptt, ctt);
    if (threadIdx.x == 0) {
        if (!zfinite(s_t) || !isfinite(tt.x)) {
            st[r].status = 2;
            active[r] = 0;
            return;
        }
        BsState &m = st[r];
        m.omega = tt.x > 0.0 ? make_double2(s_t.x / tt.x, -s_t.y / tt.x) : make_double2(0.0, 0.0);
        m.ahist[pos][0] = (float)m.alpha.x;
        m.ahist[pos][1] = (float)m.alpha.y;
        m.whist[pos][0] = (float)m.omega.x;
        m.whist[pos][1] = (float)m.omega.y;
        m.pmask |= 1u << pos;
    }
}

// beta = (rho_new / rho) (alpha / omega); restart direction (beta = 0) when
// omega or rho vanished
__device__ __forceinline__ void bs_next(BsState &m, double2 r0r)
{
    const bool ok = (m.omega.x != 0.0 || m.omega.y != 0.0) && (m.rho.x != 0.0 || m.rho.y != 0.0);
    m.beta = ok ? zmul(zdiv(r0r, m.rho), zdiv(m.alpha, m.omega)) : make_double2(0.0, 0.0);
    m.rho = r0r;
}

__global__ __launch_bounds__(1024) void k_bs_decide(const double2 *__restrict__ pa,
                                                    const double2 *__restrict__ pb, int count,
                                                    BsState *__restrict__ st,
                                                    int *__restrict__ active,
                                                    int *__restrict__ reliable, int max_iter,
                                                    double delta)
{
    const int r = blockIdx.x;
    if (active[r] != 1)
        return;
    const double2 rr = sum_at(pa, count);
    const double2 r0r = sum_at(pb, count);
    if (threadIdx.x == 0) {
        BsState &m = st[r];
        const double rsd = sqrt(rr.x);
        m.iter += 1;
        m.rsd = rsd;
        if (rsd < delta * m.maxr || rsd < m.target || m.iter >= max_iter || !isfinite(rr.x) ||
            !zfinite(r0r)) {
            reliable[r] = 1;
            active[r] = 2; // waits for the next reliable-update slot
        } else {
            bs_next(m, r0r);
        }
    }
}

__global__ __launch_bounds__(1024) void k_bs_relfin(const double2 *__restrict__ pa,
                                                    const double2 *__restrict__ pb, int count,
                                                    BsState *__restrict__ st,
                                                    int *__restrict__ active,
                                                    int *__restrict__ reliable, int max_iter)
{
    const int r = blockIdx.x;
    if (!reliable[r])
        return;
    const double2 rr = sum_at(pa, count);
    const double2 r0r = sum_at(pb, count);
    if (threadIdx.x == 0) {
        BsState &m = st[r];
        const double rsd = sqrt(rr.x);
        m.true_rsd = rsd;
        m.updates += 1;
        m.r0r = r0r;
        reliable[r] = 0;
        if (rsd < m.target) {
            m.status = 1;
            active[r] = 0;
        } else if (m.iter >= max_iter || !isfinite(rr.x) || !zfinite(r0r)) {
            m.status = isfinite(rr.x) ? 3 : 2;
            active[r] = 0;
        } else {
            bs_next(m, r0r);
            m.maxr = rsd;
            active[r] = 1;
        }
    }
}

__global__ void k_bs_markconv(const BsState *__restrict__ st, int nrhs, int *__restrict__ check)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrhs)
        check[r] = st[r].status == 1 ? 1 : 0;
}

__global__ __launch_bounds__(1024) void k_bs_fullb(const double2 *__restrict__ partials, int count,
                                                   BsState *__restrict__ st)
{
    const double2 s = sum_at(partials, count);
    if (threadIdx.x == 0)
        st[blockIdx.x].full_b = sqrt(s.x);
}

// Systems that met the Schur target: accept when the full-system residual
// meets tol, else resume with a tighter target, restarting the direction
// (p = r) from the true Schur residual of the last reliable update.
__global__ __launch_bounds__(1024) void k_bs_fullfin(const double2 *__restrict__ partials,
                                                     int count, BsState *__restrict__ st,
                                                     int *__restrict__ active,
                                                     int *__restrict__ check, double tol,
                                                     int max_iter, int resume)
{
    const int r = blockIdx.x;
    if (!check[r])
        return;
    const double2 s = sum_at(partials, count);
    if (threadIdx.x == 0) {
        BsState &m = st[r];
        const double full = sqrt(s.x);
        m.full_rsd = full;
        check[r] = 0;
        if (!resume || full <= tol * m.full_b || !(m.full_b > 0.0) || m.iter >= max_iter ||
            !isfinite(full))
            return;
        m.target *= 0.5 * tol * m.full_b / full;
        m.rho = m.r0r;
        m.beta = make_double2(0.0, 0.0);
        m.maxr = m.true_rsd;
        m.status = 0;
        active[r] = 1;
    }
}

// device scratch and work fields of one solve, taken from and handed back to
// the context's work cache
struct DevBuf {
    wd_ctx *ctx = nullptr;
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf()
    {
        if (p)
            wd_work_put(ctx, p, bytes);
    }
    int alloc(wd_ctx *c, size_t n)
    {
        ctx = c;
        bytes = (n + 15) / 16 * 16;
        p = wd_work_get(c, bytes);
        if (!p)
            return wd_set_error(c, WD_ENOMEM, "mixed cg: allocation failed");
        return wd_launch_zero(c, p, bytes);
    }
};

struct FieldBuf {
    wd_field *f = nullptr;
    ~FieldBuf() { wd_field_destroy(f); }
};

// Reads the device stop flag after the work queued so far: returns 0 or 1,
// or an error status (< 0) when the copy or the synchronisation failed.
int poll_flag(wd_ctx *ctx, const int *stop, const char *who)
{
    hipError_t e = hipMemcpyAsync(ctx->hpinned, stop, sizeof(int), hipMemcpyDeviceToHost,
                                  ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
        return wd_set_error(ctx, WD_EHIP, std::string(who) + ": poll: " + hipGetErrorString(e));
    return *(const int *)ctx->hpinned ? 1 : 0;
}

} // namespace

// ---- autotuning of the sloppy operator ------------------------------------------
// Times HOP_M over `nrhs` systems for every (rhs per thread, site chunk)
// candidate and keeps the fastest per (nrhs, spinor format, link format,
// one or both parities).
// fixed_R > 0 restricts the search to that rhs-per-thread.  Every system of a
// site in one thread (WD_RHS_LOOP) is a candidate for F16 spinors with F16 or
// F32 12-real links from kLoopTuneMin systems on, where it shares each link
// load among many systems; with F16 links its products round the half spinor
// to F16, so below that the candidates stay the bit-identical work orders
// (with F32 links its operator is bit-identical to R = 1, 2, 4, but its dot
// products sum per wavefront).  For WD_RHS_LOOP the
// chunk candidates are systems per pass (0: the context's choice, 12 by
// default), so the systems-per-thread axis runs 1, 2, 4 and 12 .. 64; group
// sizes give bit-identical results.  The choice is cached per context, so the
// solves of one context repeat bit for bit.
static const int kTuneR[WD_TUNE_NR] = {1, 2, 4, WD_RHS_LOOP};
static const int kTuneC[WD_TUNE_NC] = {0, 16, 32, 64};
constexpr int kLoopTuneMin = 8;

int wd_tune_sloppy(wd_ctx *ctx, const SloppyArgs &a0, int sprec, int gprec, int recon, int cb,
                   int fixed_R, int *best_R, int *best_C, int *best_T, double *ms_out)
{
    // cb = -1: time the both-parity M (full-system solver); cb = 0: the
    // one-parity launch of the Schur solver, whose best work order differs
    const auto key = std::make_tuple(a0.nrhs, sprec, gprec * 100 + recon + (cb < 0 ? 0 : 1000));
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end() && !ms_out && (fixed_R <= 0 || std::get<0>(it->second) == fixed_R)) {
        *best_R = std::get<0>(it->second);
        *best_C = std::get<1>(it->second);
        if (best_T)
            *best_T = std::get<2>(it->second);
        return WD_OK;
    }
    if (ms_out)
        for (int k = 0; k < WD_TUNE_NR * WD_TUNE_NC; ++k)
            ms_out[k] = -1.0;
    double best = 1e30;
    int bestR = fixed_R > 0 ? fixed_R : 1, bestC = 0, bestT = 0;
    // work orders per candidate: the z-t tiled columns, and the natural site
    // order when the tiled one differs from it (one system at 32^3x64: the
    // one-parity launch 75 -> 72 us with F16 links, 118 -> 107 us with F32)
    const int ntiles = ctx->geom.tW > 0 ? 2 : 1;
    hipEvent_t e0, e1;
    WD_HIP_CHECK(ctx, hipEventCreate(&e0));
    WD_HIP_CHECK(ctx, hipEventCreate(&e1));
    int rc = WD_OK;
    for (int ir = 0; ir < WD_TUNE_NR && !rc; ++ir) {
        const int R = kTuneR[ir];
        if (fixed_R > 0 && R != fixed_R)
            continue;
        if (R == WD_RHS_LOOP && !(wd_sloppy_loop_supported(sprec, gprec, recon) &&
                                  (fixed_R == WD_RHS_LOOP || a0.nrhs >= kLoopTuneMin)))
            continue;
        const bool one_group = R != WD_RHS_LOOP && wd_sloppy_groups(a0.nrhs, R) == 1;
        int loop_groups_seen = 0; // bit g: a loop candidate with g + 1 groups was timed
        for (int ic = 0; ic < WD_TUNE_NC && !rc; ++ic) {
            if (one_group && ic > 0)
                break; // the work order cannot change with a single group
            if (R == WD_RHS_LOOP) {
                // the chunk candidate is the loop kernel's systems per pass;
                // candidates with the same number of passes are the same launch
                const int cap = wd_loop_group_cap(ctx, a0.nrhs, kTuneC[ic]);
                const int ng = (a0.nrhs + cap - 1) / cap;
                if (ng <= 31 && (loop_groups_seen >> (ng - 1) & 1))
                    continue;
                if (ng <= 31)
                    loop_groups_seen |= 1 << (ng - 1);
            }
            for (int itile = 0; itile < ntiles && !rc; ++itile) {
                SloppyArgs a = a0;
                a.chunk = kTuneC[ic];
                a.tile = -itile; // 0: tiled columns, -1: natural order
                for (int w = 0; w < 2 && !rc; ++w)
                    rc = wd_launch_hop_sloppy(ctx, a, sprec, gprec, recon, R, 0, cb, HOP_M);
                if (rc)
                    break;
                (void)hipEventRecord(e0, ctx->stream);
                const int reps = 5;
                for (int w = 0; w < reps && !rc; ++w)
                    rc = wd_launch_hop_sloppy(ctx, a, sprec, gprec, recon, R, 0, cb, HOP_M);
                (void)hipEventRecord(e1, ctx->stream);
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                const double t = ms / reps;
                if (ms_out && (itile == 0 || t < ms_out[ir * WD_TUNE_NC + ic]))
                    ms_out[ir * WD_TUNE_NC + ic] = t;
                if (t < best) {
                    best = t;
                    bestR = R;
                    bestC = kTuneC[ic];
                    bestT = a.tile;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc)
        return rc;
    if (fixed_R <= 0)
        ctx->tuned[key] = std::make_tuple(bestR, bestC, bestT);
    *best_R = bestR;
    *best_C = bestC;
    if (best_T)
        *best_T = bestT;
    return WD_OK;
}

extern "C" void wd_mixed_params_default(wd_mixed_params *p)
{
    if (!p)
        return;
    p->max_iter = 100000;
    p->tol = 1e-10;
    // measured on MI355X (32^3x64 Schur, F16 spinors): 162 iterations at 1e-2,
    // 152 at 2e-2, 151 at 3e-2 with one more update; 48 systems within 1 %
    p->delta = 2e-2;
    p->spinor_prec = WD_F16;
    p->rhs_per_thread = 0;
    p->site_chunk = -1;
    p->check_every = 16;
    p->use_graph = 1;
    p->even_odd = 0;
    p->reliable_every = 0;
}

extern "C" int wd_cg_mixed(wd_ctx *ctx, wd_field *x, const wd_field *b, const wd_gauge *U,
                           const wd_gauge *Us, const double *masses, const wd_mixed_params *prm,
                           wd_mixed_report *reports, double *seconds)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, x && b && U && Us && masses && prm, "mixed cg: null argument");
    WD_ARG_CHECK(ctx, x != b, "mixed cg: x and b must be distinct fields");
    WD_ARG_CHECK(ctx, x->prec == WD_F64 && b->prec == WD_F64, "mixed cg: x and b must be F64");
    WD_ARG_CHECK(ctx, x->nrhs == b->nrhs, "mixed cg: x and b nrhs differ");
    WD_ARG_CHECK(ctx, U->prec == WD_F64, "mixed cg: U must be F64 (reliable updates)");
    WD_ARG_CHECK(ctx, prm->spinor_prec == WD_F16 || prm->spinor_prec == WD_F32,
                 "mixed cg: spinor_prec must be F16 or F32");
    WD_ARG_CHECK(ctx, wd_sloppy_supported(prm->spinor_prec, Us->prec, Us->recon),
                 "mixed cg: unsupported sloppy link format");
    WD_ARG_CHECK(ctx, prm->rhs_per_thread == 0 || prm->rhs_per_thread == 1 ||
                          prm->rhs_per_thread == 2 || prm->rhs_per_thread == 4 ||
                          prm->rhs_per_thread == WD_RHS_LOOP,
                 "mixed cg: rhs_per_thread must be 0, 1, 2, 4 or WD_RHS_LOOP");
    WD_ARG_CHECK(ctx, prm->rhs_per_thread != WD_RHS_LOOP ||
                          wd_sloppy_loop_supported(prm->spinor_prec, Us->prec, Us->recon),
                 "mixed cg: WD_RHS_LOOP needs F16 spinors and F16 12-real links");
    WD_ARG_CHECK(ctx, prm->site_chunk >= -1, "mixed cg: site_chunk must be >= -1");
    WD_ARG_CHECK(ctx, prm->max_iter >= 1 && prm->check_every >= 1 && prm->tol > 0.0 &&
                          prm->delta > 0.0 && prm->delta < 1.0 && prm->reliable_every >= 0,
                 "mixed cg: bad iteration / tolerance parameters");
    const auto t0 = std::chrono::steady_clock::now();
    const int n = b->nrhs;
    const int sp = prm->spinor_prec;
    const int stride = (int)ctx->stride;
    int rc;

    const bool eo = prm->even_odd != 0;
    // automatic spacing of the reliable-update slots (measured on MI355X): 8
    // for one system, 4 for many, where updates of different systems share
    // the FP64 launches of a slot
    const int slot_every = prm->reliable_every > 0 ? prm->reliable_every : (n == 1 ? 8 : 4);
    // iterations per group of pending x_s steps: one search-direction buffer each
    const int npend = std::min(slot_every, kPend);
    // device work vectors (even-odd: the Schur system lives on the even half;
    // the odd half of an operator's output buffer holds its intermediate hop,
    // q only that of the A^dag launch); p holds npend buffers of n systems
    FieldBuf Mdb, Mfull, t64, q64, rs, p, q, qq, xs;
    if ((rc = wd_field_create_work(ctx, &Mdb.f, WD_F64, n)) || (rc = wd_field_create_work(ctx, &t64.f, WD_F64, n)) ||
        (rc = wd_field_create_work(ctx, &q64.f, WD_F64, n)) || (rc = wd_field_create_work(ctx, &rs.f, sp, n)) ||
        (rc = wd_field_create_work(ctx, &p.f, sp, npend * n)) ||
        (eo && (rc = wd_field_create_work(ctx, &q.f, sp, n))) ||
        (rc = wd_field_create_work(ctx, &qq.f, sp, n)) ||
        (rc = wd_field_create_work(ctx, &xs.f, WD_F32, n)) ||
        (eo && (rc = wd_field_create_work(ctx, &Mfull.f, WD_F64, n))))
        return rc;
    const int eg_full = (2 * stride + kBlock - 1) / kBlock; // elementwise workgroups per rhs
    // elementwise kernels cover (parity, site) e < nel: the even half only for
    // even-odd (the odd half of the work vectors holds intermediate hops)
    const int nel = eo ? stride : 2 * stride;
    const int eg = (nel + kBlock - 1) / kBlock;
    const int bpp = (int)((ctx->Vh + kBlock - 1) / kBlock);
    const int hop_nsb = eo ? bpp : 2 * bpp; // workgroups of the (last) operator launch
    // room for the operator's partials at kLoopWaves per workgroup (WD_RHS_LOOP)
    const int nsb_max = std::max(eg_full, 2 * bpp * kLoopWaves);
    DevBuf partials, state, flags, diag_d, diag_f, coef;
    if ((rc = partials.alloc(ctx, sizeof(double2) * (size_t)n * nsb_max)) ||
        (rc = state.alloc(ctx, sizeof(MxState) * n)) ||
        (rc = flags.alloc(ctx, sizeof(int) * (3 * (size_t)n + 4))) ||
        (rc = diag_d.alloc(ctx, sizeof(double) * n)) ||
        (rc = diag_f.alloc(ctx, sizeof(float) * 2 * n)) ||
        (rc = coef.alloc(ctx, sizeof(double) * 3 * n)))
        return rc;
    // diag_f: [n] the sloppy operator's diagonal, then [n] alpha of the current
    // iteration (the operator's r -= alpha q)
    float *alpha_f = (float *)diag_f.p + n;
    MxState *st = (MxState *)state.p;
    int *active = (int *)flags.p, *reliable = active + n, *check = reliable + n, *stop = check + n;
    double2 *part = (double2 *)partials.p;
    // per-system coefficients: a_r = 4 + m_r; even-odd also a_r^2 (diagonal of
    // A), 1/a_r and 1/(2 a_r) (reconstruction of x_o)
    double *a2_d = (double *)coef.p, *inva_d = a2_d + n, *inv2a_d = inva_d + n;
    {
        std::vector<double> ad(n), cd(3 * (size_t)n);
        std::vector<float> af(n);
        for (int r = 0; r < n; ++r) {
            ad[r] = 4.0 + masses[r];
            af[r] = eo ? (float)(ad[r] * ad[r]) : (float)ad[r];
            cd[r] = ad[r] * ad[r];
            cd[n + r] = 1.0 / ad[r];
            cd[2 * n + r] = 0.5 / ad[r];
        }
        WD_HIP_CHECK(ctx, hipMemcpyAsync(diag_d.p, ad.data(), sizeof(double) * n,
                                         hipMemcpyHostToDevice, ctx->stream));
        WD_HIP_CHECK(ctx, hipMemcpyAsync(diag_f.p, af.data(), sizeof(float) * n,
                                         hipMemcpyHostToDevice, ctx->stream));
        WD_HIP_CHECK(ctx, hipMemcpyAsync(coef.p, cd.data(), sizeof(double) * 3 * n,
                                         hipMemcpyHostToDevice, ctx->stream));
        WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    const F64Links Ud = f64_links(U);
    const size_t xe = x->rhs_elems;
    const size_t srb = rs.f->rhs_elems * prec_bytes_complex(sp);
    const size_t xsb = xs.f->rhs_elems * prec_bytes_complex(WD_F32);
    const double *ddiag = (const double *)diag_d.p;
    F64Batch all = {xe, nullptr, ddiag, nullptr};
    F64Batch rel = {xe, reliable, ddiag, nullptr};
    F64Batch all2 = {xe, nullptr, a2_d, nullptr};
    F64Batch rel2 = {xe, reliable, a2_d, nullptr};
    double2 *xd = (double2 *)x->data;
    const double2 *bd = (const double2 *)b->data;
    double2 *Mdbd = (double2 *)Mdb.f->data, *t64d = (double2 *)t64.f->data,
            *q64d = (double2 *)q64.f->data;

    // FP64 A v (dag 0) or A^dag v on the even half, A = a^2 - D_eo D_oe / 4;
    // the odd half of out holds the intermediate +D_oe v / 2
    auto apply_A64 = [&](int dag, const double2 *in, double2 *out, const F64Batch &b1,
                         const F64Batch &b2, const int *stp) -> int {
        int e = wd_launch_hop_f64_batch_c(ctx, out, in, Ud, dag, 1, HOP_ONLY, 0.5, nullptr,
                                          nullptr, nullptr, stp, b1, n);
        if (e)
            return e;
        return wd_launch_hop_f64_batch_c(ctx, out, out, Ud, dag, 0, HOP_M, -0.5, in, nullptr,
                                         nullptr, stp, b2, n);
    };

    // right-hand side of the CG system: M^dag b, or A^dag bh with
    // bh = a b_e + D_eo b_o / 2 (even-odd; then also M^dag b for the full check)
    if (!eo) {
        if ((rc = wd_launch_hop_f64_batch(ctx, Mdbd, bd, Ud, 1, -1, HOP_M, bd, nullptr, nullptr,
                                          nullptr, all, n)))
            return rc;
    } else {
        if ((rc = wd_launch_hop_f64_batch_c(ctx, t64d, bd, Ud, 0, 0, HOP_M, 0.5, bd, nullptr,
                                            nullptr, nullptr, all, n)) ||
            (rc = apply_A64(1, t64d, Mdbd, all, all2, nullptr)) ||
            (rc = wd_launch_hop_f64_batch(ctx, (double2 *)Mfull.f->data, bd, Ud, 1, -1, HOP_M,
                                          bd, nullptr, nullptr, nullptr, all, n)))
            return rc;
    }
    if ((rc = wd_launch_zero(ctx, x->data, x->bytes)))
        return rc;
    // r_s = rhs (even half for even-odd), rho = |rhs|^2
    if (sp == WD_F16)
        hipLaunchKernelGGL(k_mx_relr<WD_F16>, dim3(eg, n), dim3(kBlock), 0, ctx->stream,
                           (char *)rs.f->data, rs.f->scale, (const double2 *)Mdbd,
                           (const double2 *)nullptr, xe, srb, stride, nel, (const int *)nullptr, part,
                           (const double2 *)Us->omega);
    else
        hipLaunchKernelGGL(k_mx_relr<WD_F32>, dim3(eg, n), dim3(kBlock), 0, ctx->stream,
                           (char *)rs.f->data, rs.f->scale, (const double2 *)Mdbd,
                           (const double2 *)nullptr, xe, srb, stride, nel, (const int *)nullptr, part,
                           (const double2 *)Us->omega);
    if ((rc = wd_check_launch(ctx, "k_mx_relr")))
        return rc;
    // even-odd: the Schur target starts at half the full-system tolerance
    hipLaunchKernelGGL(k_mx_init, dim3(n), dim3(1024), 0, ctx->stream, part, eg, st, active,
                       reliable, eo ? 0.5 * prm->tol : prm->tol);
    if (eo) {
        hipLaunchKernelGGL(k_mx_fullres, dim3(eg_full, n), dim3(kBlock), 0, ctx->stream,
                           (const double2 *)Mfull.f->data, (const double2 *)nullptr, xe, stride,
                           (const int *)nullptr, part);
        hipLaunchKernelGGL(k_mx_fullb, dim3(n), dim3(1024), 0, ctx->stream,
                           (const double2 *)part, eg_full, st);
    }
    hipLaunchKernelGGL(k_mx_alldone, dim3(1), dim3(64), 0, ctx->stream, active, n, stop);
    if ((rc = wd_check_launch(ctx, "k_mx_init")))
        return rc;

    // sloppy operator arguments
    SloppyArgs A = {};
    A.rhs_bytes = srb;
    A.U = (const char *)Us->data;
    A.diag = (const float *)diag_f.p;
    A.active = active;
    A.partials = part;
    A.stop = stop;
    A.nrhs = n;
    A.hc = -0.5f;
    int R = prm->rhs_per_thread;
    int C = prm->site_chunk;
    if (R == 0 || C < 0) {
        SloppyArgs T = A;
        T.out = (char *)qq.f->data;
        T.out_scale = qq.f->scale;
        T.in = T.same = (const char *)p.f->data;
        T.in_scale = T.same_scale = p.f->scale;
        int tR = 0, tC = 0, tT = 0;
        if ((rc = wd_tune_sloppy(ctx, T, sp, Us->prec, Us->recon, eo ? 0 : -1, R, &tR, &tC,
                                 &tT, nullptr)))
            return rc;
        R = tR;
        if (C < 0) {
            C = tC;
            A.tile = tT; // the work order timed with that chunk
        }
    }
    A.chunk = C;
    const int nsb = hop_nsb * wd_sloppy_partials_per_block(R); // partials of the operator

    // sloppy out = M in (full) or A in (even-odd), with the partials of
    // (dotv, out) for HOP_M_DOT
    auto sloppy_op = [&](int dag, wd_field *out, const wd_field *in, int mode,
                         const wd_field *dotv) -> int {
        SloppyArgs m = A;
        m.out = (char *)out->data;
        m.out_scale = out->scale;
        m.in = m.same = (const char *)in->data;
        m.in_scale = m.same_scale = in->scale;
        if (dotv) {
            m.dotv = (const char *)dotv->data;
            m.dotv_scale = dotv->scale;
        }
        if (!eo)
            return wd_launch_hop_sloppy(ctx, m, sp, Us->prec, Us->recon, R, dag, -1, mode);
        SloppyArgs h = m; // +D_oe in / 2 into the odd half of out
        h.hc = 0.5f;
        int e = wd_launch_hop_sloppy(ctx, h, sp, Us->prec, Us->recon, R, dag, 1, HOP_ONLY);
        if (e)
            return e;
        m.in = (const char *)out->data; // a^2 in_e - D_eo (odd half of out) / 2
        m.in_scale = out->scale;
        return wd_launch_hop_sloppy(ctx, m, sp, Us->prec, Us->recon, R, dag, 0, mode);
    };
    // r_s -= alpha A^dag qq with the partials of |r_s|^2 (HOP_M_AXPY); even-odd:
    // the intermediate hop goes to the odd half of q
    auto sloppy_axpy = [&]() -> int {
        SloppyArgs m = A;
        m.out = (char *)rs.f->data;
        m.out_scale = rs.f->scale;
        m.dotv = (const char *)rs.f->data;
        m.dotv_scale = rs.f->scale;
        m.axpy = alpha_f;
        m.in = m.same = (const char *)qq.f->data;
        m.in_scale = m.same_scale = qq.f->scale;
        if (!eo)
            return wd_launch_hop_sloppy(ctx, m, sp, Us->prec, Us->recon, R, 1, -1, HOP_M_AXPY);
        SloppyArgs h = m;
        h.hc = 0.5f;
        h.out = (char *)q.f->data;
        h.out_scale = q.f->scale;
        int e = wd_launch_hop_sloppy(ctx, h, sp, Us->prec, Us->recon, R, 1, 1, HOP_ONLY);
        if (e)
            return e;
        m.in = (const char *)q.f->data;
        m.in_scale = q.f->scale;
        return wd_launch_hop_sloppy(ctx, m, sp, Us->prec, Us->recon, R, 1, 0, HOP_M_AXPY);
    };

    // search-direction buffer j: n systems at pbytes * j (scales at pscale * j)
    const size_t pbytes = (size_t)n * srb, pscale = (size_t)n * 2 * stride;
    auto pbuf = [&](int j) -> wd_field {
        wd_field v = *p.f;
        v.nrhs = n;
        v.data = (char *)p.f->data + j * pbytes;
        v.scale = p.f->scale ? p.f->scale + j * pscale : nullptr;
        return v;
    };

    // iteration whose new direction goes into buffer pos
    auto iteration = [&](int pos) -> int {
        int e;
        if (sp == WD_F16)
            hipLaunchKernelGGL(k_mx_xpay<WD_F16>, dim3(eg, n), dim3(kBlock), 0, ctx->stream,
                               (char *)p.f->data, p.f->scale, pbytes, pscale, pos,
                               (const char *)rs.f->data, (const float *)rs.f->scale, srb, stride,
                               nel, (const MxState *)st, (const int *)active);
        else
            hipLaunchKernelGGL(k_mx_xpay<WD_F32>, dim3(eg, n), dim3(kBlock), 0, ctx->stream,
                               (char *)p.f->data, p.f->scale, pbytes, pscale, pos,
                               (const char *)rs.f->data, (const float *)rs.f->scale, srb, stride,
                               nel, (const MxState *)st, (const int *)active);
        if ((e = wd_check_launch(ctx, "k_mx_xpay")))
            return e;
        // (p, q) = (p, A^dag A p) = |A p|^2, summed over the stored A p by the
        // launch that writes it; alpha; then r_s -= alpha A^dag (A p) and |r_s|^2
        // in the A^dag launch
        const wd_field pp = pbuf(pos);
        if ((e = sloppy_op(0, qq.f, &pp, HOP_M_NORM, nullptr)))
            return e;
        hipLaunchKernelGGL(k_mx_alpha, dim3(n), dim3(1024), 0, ctx->stream,
                           (const double2 *)part, nsb, st, active, pos, alpha_f);
        if ((e = wd_check_launch(ctx, "k_mx_alpha")) || (e = sloppy_axpy()))
            return e;
        hipLaunchKernelGGL(k_mx_decide, dim3(n), dim3(1024), 0, ctx->stream,
                           (const double2 *)part, nsb, st, active, reliable, prm->max_iter,
                           prm->delta);
        return wd_check_launch(ctx, "mixed cg iteration");
    };

    // x_s += the group's pending steps; x += x_s, x_s = 0 for the systems in
    // due (nullptr: none)
    auto flush = [&](const int *due) -> int {
        if (sp == WD_F16)
            hipLaunchKernelGGL(k_mx_flush<WD_F16>, dim3(eg, n), dim3(kBlock), 0, ctx->stream, xd,
                               (char *)xs.f->data, xe, xsb, stride, nel, due,
                               (const double2 *)Us->omega, (const char *)p.f->data,
                               (const float *)p.f->scale, pbytes, pscale, srb,
                               (const MxState *)st);
        else
            hipLaunchKernelGGL(k_mx_flush<WD_F32>, dim3(eg, n), dim3(kBlock), 0, ctx->stream, xd,
                               (char *)xs.f->data, xe, xsb, stride, nel, due,
                               (const double2 *)Us->omega, (const char *)p.f->data,
                               (const float *)p.f->scale, pbytes, pscale, srb,
                               (const MxState *)st);
        return wd_check_launch(ctx, "k_mx_flush");
    };

    // reliable-update slot for the waiting systems
    auto reliable_slot = [&]() -> int {
        int e;
        if ((e = flush(reliable)))
            return e;
        if (!eo) {
            if ((e = wd_launch_hop_f64_batch(ctx, t64d, xd, Ud, 0, -1, HOP_M, xd, nullptr,
                                             nullptr, stop, rel, n)) ||
                (e = wd_launch_hop_f64_batch(ctx, q64d, t64d, Ud, 1, -1, HOP_M, t64d, nullptr,
                                             nullptr, stop, rel, n)))
                return e;
        } else {
            if ((e = apply_A64(0, xd, t64d, rel, rel2, stop)) ||
                (e = apply_A64(1, t64d, q64d, rel, rel2, stop)))
                return e;
        }
        if (sp == WD_F16)
            hipLaunchKernelGGL(k_mx_relr<WD_F16>, dim3(eg, n), dim3(kBlock), 0, ctx->stream,
                               (char *)rs.f->data, rs.f->scale, (const double2 *)Mdbd,
                               (const double2 *)q64d, xe, srb, stride, nel, (const int *)reliable,
                               part, (const double2 *)Us->omega);
        else
            hipLaunchKernelGGL(k_mx_relr<WD_F32>, dim3(eg, n), dim3(kBlock), 0, ctx->stream,
                               (char *)rs.f->data, rs.f->scale, (const double2 *)Mdbd,
                               (const double2 *)q64d, xe, srb, stride, nel, (const int *)reliable,
                               part, (const double2 *)Us->omega);
        hipLaunchKernelGGL(k_mx_relfin, dim3(n), dim3(1024), 0, ctx->stream,
                           (const double2 *)part, eg, st, active, reliable, prm->max_iter);
        hipLaunchKernelGGL(k_mx_alldone, dim3(1), dim3(64), 0, ctx->stream,
                           (const int *)active, n, stop);
        return wd_check_launch(ctx, "mixed cg reliable update");
    };

    // a block of `chunk` iterations with a reliable-update slot after every
    // `slot_every` of them and after the last
    const int chunk = prm->check_every;
    // every group of iterations (at most npend) ends with a flush of the
    // pending x_s steps, fused into the reliable-update slot when one is due
    auto block = [&]() -> int {
        int e = 0;
        for (int k = 0, pos = 0; k < chunk && !e; ++k) {
            e = iteration(pos);
            const bool slot = (k + 1) % slot_every == 0 || k + 1 == chunk;
            if (!e && slot)
                e = reliable_slot();
            else if (!e && pos == npend - 1)
                e = flush(nullptr);
            pos = (slot || pos == npend - 1) ? 0 : pos + 1;
        }
        return e;
    };

    // even-odd: x_o = b_o / a + D_oe x_e / (2a) and |M^dag b - M^dag M x| for the
    // systems that met the Schur target; the ones that miss tol resume.  The
    // final pass (final_pass = 1, after the loop) does the same for every
    // system not yet accepted, without resuming any.
    auto full_check = [&](int final_pass) -> int {
        int e;
        F64Batch chk_rec = {xe, check, inva_d, inv2a_d};
        F64Batch chk = {xe, check, ddiag, nullptr};
        if (final_pass)
            hipLaunchKernelGGL(k_mark_final<MxState>, dim3((n + 255) / 256), dim3(256), 0,
                               ctx->stream, (const MxState *)st, n, prm->tol, check);
        else
            hipLaunchKernelGGL(k_mx_markconv, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                               (const MxState *)st, n, check);
        if ((e = wd_check_launch(ctx, "k_mx_markconv")) ||
            (e = wd_launch_hop_f64_batch_c(ctx, xd, xd, Ud, 0, 1, HOP_M, 0.0, bd, nullptr,
                                           nullptr, nullptr, chk_rec, n)) ||
            (e = wd_launch_hop_f64_batch(ctx, t64d, xd, Ud, 0, -1, HOP_M, xd, nullptr, nullptr,
                                         nullptr, chk, n)) ||
            (e = wd_launch_hop_f64_batch(ctx, q64d, t64d, Ud, 1, -1, HOP_M, t64d, nullptr,
                                         nullptr, nullptr, chk, n)))
            return e;
        hipLaunchKernelGGL(k_mx_fullres, dim3(eg_full, n), dim3(kBlock), 0, ctx->stream,
                           (const double2 *)Mfull.f->data, (const double2 *)q64d, xe, stride,
                           (const int *)check, part);
        hipLaunchKernelGGL(k_mx_fullfin, dim3(n), dim3(1024), 0, ctx->stream,
                           (const double2 *)part, eg_full, st, active, check, prm->tol,
                           prm->max_iter, final_pass ? 0 : 1);
        if (!final_pass)
            hipLaunchKernelGGL(k_mx_alldone, dim3(1), dim3(64), 0, ctx->stream,
                               (const int *)active, n, stop);
        return wd_check_launch(ctx, "mixed cg full check");
    };

    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    if (prm->use_graph) {
        WD_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int e = block();
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
                                std::string("mixed cg: graph instantiate: ") + hipGetErrorString(ce));
        }
    }
    int status = WD_OK;
    long launched = 0;
    const long limit = (long)prm->max_iter + chunk; // the device forces an exit at max_iter
    int checks = 0;
    bool checked = false; // a full check ran since the last iterations
    while (true) {
        // a failed poll ends the loop like a failed launch, so the graph is
        // always destroyed before the work blocks its nodes use are handed back
        if ((status = poll_flag(ctx, stop, "mixed cg")) < 0)
            break;
        const int hstop = status;
        status = WD_OK;
        if (hstop && eo && !checked && checks < 4) {
            // all systems stopped: verify the converged ones on the full system;
            // systems that miss tol are resumed and clear the stop flag
            ++checks;
            checked = true;
            if ((status = full_check(0)))
                break;
            continue;
        }
        if (hstop || launched > limit)
            break;
        checked = false;
        if (gexec) {
            hipError_t ge = hipGraphLaunch(gexec, ctx->stream);
            if (ge != hipSuccess) {
                status = wd_set_error(ctx, WD_EHIP,
                                      std::string("mixed cg: graph launch: ") + hipGetErrorString(ge));
                break;
            }
        } else {
            if ((status = block()))
                break;
        }
        launched += chunk;
    }
    if (gexec)
        (void)hipGraphExecDestroy(gexec);
    if (graph)
        (void)hipGraphDestroy(graph);
    // even-odd: x_o and the full-system residual of every system not accepted
    // by a check (stopped at the limit, broken down, or resumed after the last
    // check), so each report holds the measured residual of the returned x
    if (!status && eo)
        status = full_check(1);
    if (status)
        return status;
    std::vector<MxState> hs(n);
    WD_HIP_CHECK(ctx, hipMemcpyAsync(hs.data(), st, sizeof(MxState) * n, hipMemcpyDeviceToHost,
                                     ctx->stream));
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds)
        *seconds = std::chrono::duration<double>(t1 - t0).count();
    bool all_conv = true, breakdown = false;
    for (int r = 0; r < n; ++r) {
        const double bn = eo ? hs[r].full_b : hs[r].b_norm;
        const double tr = eo ? hs[r].full_rsd : hs[r].true_rsd;
        const bool conv = hs[r].status == 1 && (!eo || !(bn > 0.0) || tr <= prm->tol * bn);
        if (reports) {
            reports[r].iterations = hs[r].iter;
            reports[r].reliable_updates = hs[r].updates;
            reports[r].converged = conv;
            reports[r].true_rel_rsd = bn > 0 ? tr / bn : 0.0;
            reports[r].rhs_norm = bn;
            reports[r].rhs_per_thread = R;
            reports[r].site_chunk = C;
            reports[r].block_tile = A.tile;
        }
        all_conv = all_conv && conv;
        breakdown = breakdown || hs[r].status == 2;
    }
    if (breakdown)
        return wd_set_error(ctx, WD_EBREAKDOWN, "mixed cg: breakdown in at least one system");
    return all_conv ? WD_OK : WD_ENOTCONV;
}


// Mixed-precision BiCGStab (van der Vorst 1992) with reliable updates
// (Sleijpen & van der Vorst 1996) on the Schur system A x_e = bh of every
// system, A = a^2 - D_eo D_oe / 4, bh = a b_e + D_eo b_o / 2: two
// applications of A per iteration, as CGNR's A and A^dag, but A's condition
// number instead of A^dag A's.  Same sloppy operator, FP64 reliable updates
// (r = bh - A x) in slots, and the same full-system FP64 check of the
// reference's stop test as wd_cg_mixed with even_odd.
extern "C" int wd_bicgstab_mixed(wd_ctx *ctx, wd_field *x, const wd_field *b, const wd_gauge *U,
                                 const wd_gauge *Us, const double *masses,
                                 const wd_mixed_params *prm, wd_mixed_report *reports,
                                 double *seconds)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, x && b && U && Us && masses && prm, "mixed bicgstab: null argument");
    WD_ARG_CHECK(ctx, x != b, "mixed bicgstab: x and b must be distinct fields");
    WD_ARG_CHECK(ctx, x->prec == WD_F64 && b->prec == WD_F64, "mixed bicgstab: x and b must be F64");
    WD_ARG_CHECK(ctx, x->nrhs == b->nrhs, "mixed bicgstab: x and b nrhs differ");
    WD_ARG_CHECK(ctx, U->prec == WD_F64, "mixed bicgstab: U must be F64 (reliable updates)");
    WD_ARG_CHECK(ctx, prm->spinor_prec == WD_F16 || prm->spinor_prec == WD_F32,
                 "mixed bicgstab: spinor_prec must be F16 or F32");
    WD_ARG_CHECK(ctx, wd_sloppy_supported(prm->spinor_prec, Us->prec, Us->recon),
                 "mixed bicgstab: unsupported sloppy link format");
    WD_ARG_CHECK(ctx, prm->rhs_per_thread == 0 || prm->rhs_per_thread == 1 ||
                          prm->rhs_per_thread == 2 || prm->rhs_per_thread == 4 ||
                          prm->rhs_per_thread == WD_RHS_LOOP,
                 "mixed bicgstab: rhs_per_thread must be 0, 1, 2, 4 or WD_RHS_LOOP");
    WD_ARG_CHECK(ctx, prm->rhs_per_thread != WD_RHS_LOOP ||
                          wd_sloppy_loop_supported(prm->spinor_prec, Us->prec, Us->recon),
                 "mixed bicgstab: WD_RHS_LOOP needs F16 spinors and F16 12-real links");
    WD_ARG_CHECK(ctx, prm->site_chunk >= -1, "mixed bicgstab: site_chunk must be >= -1");
    WD_ARG_CHECK(ctx, prm->max_iter >= 1 && prm->check_every >= 1 && prm->tol > 0.0 &&
                          prm->delta > 0.0 && prm->delta < 1.0 && prm->reliable_every >= 0,
                 "mixed bicgstab: bad iteration / tolerance parameters");
    const auto t0 = std::chrono::steady_clock::now();
    const int n = b->nrhs;
    const int sp = prm->spinor_prec;
    const int stride = (int)ctx->stride;
    int rc;

    // reliable-update slots every slot_every iterations (8 for one system, 4 for
    // many, as the CG); groups of npend iterations keep their p and s in
    // separate buffers for one x_s pass per group
    const int slot_every = prm->reliable_every > 0 ? prm->reliable_every : (n == 1 ? 8 : 4);
    const int npend = std::min(slot_every, kPend);
    // F64: bh (even half), A x, M^dag b and the check's M^dag M x; sloppy: r, r0,
    // p and s (npend buffers of n systems each), v, t (the odd halves of v and t
    // hold the operator's intermediate hop)
    FieldBuf bh, Mfull, t64, q64, r, r0, pv, v, sv, t, xs;
    if ((rc = wd_field_create_work(ctx, &bh.f, WD_F64, n)) || (rc = wd_field_create_work(ctx, &t64.f, WD_F64, n)) ||
        (rc = wd_field_create_work(ctx, &q64.f, WD_F64, n)) || (rc = wd_field_create_work(ctx, &Mfull.f, WD_F64, n)) ||
        (rc = wd_field_create_work(ctx, &r.f, sp, n)) || (rc = wd_field_create_work(ctx, &r0.f, sp, n)) ||
        (rc = wd_field_create_work(ctx, &pv.f, sp, npend * n)) || (rc = wd_field_create_work(ctx, &v.f, sp, n)) ||
        (rc = wd_field_create_work(ctx, &sv.f, sp, npend * n)) || (rc = wd_field_create_work(ctx, &t.f, sp, n)) ||
        (rc = wd_field_create_work(ctx, &xs.f, WD_F32, n)))
        return rc;
    const int eg_full = (2 * stride + kBlock - 1) / kBlock;
    const int nel = stride; // the even half
    const int eg = (nel + kBlock - 1) / kBlock;
    const int bpp = (int)((ctx->Vh + kBlock - 1) / kBlock);
    const int hop_nsb = bpp; // workgroups of a one-parity operator launch
    const int nsb_max = std::max(eg_full, 2 * bpp * kLoopWaves);
    DevBuf pa_buf, pb_buf, state, flags, diag_d, diag_f, coef;
    if ((rc = pa_buf.alloc(ctx, sizeof(double2) * (size_t)n * nsb_max)) ||
        (rc = pb_buf.alloc(ctx, sizeof(double2) * (size_t)n * nsb_max)) ||
        (rc = state.alloc(ctx, sizeof(BsState) * n)) ||
        (rc = flags.alloc(ctx, sizeof(int) * (3 * (size_t)n + 4))) ||
        (rc = diag_d.alloc(ctx, sizeof(double) * n)) || (rc = diag_f.alloc(ctx, sizeof(float) * n)) ||
        (rc = coef.alloc(ctx, sizeof(double) * 3 * n)))
        return rc;
    BsState *st = (BsState *)state.p;
    int *active = (int *)flags.p, *reliable = active + n, *check = reliable + n, *stop = check + n;
    double2 *pa = (double2 *)pa_buf.p, *pb = (double2 *)pb_buf.p;
    double *a2_d = (double *)coef.p, *inva_d = a2_d + n, *inv2a_d = inva_d + n;
    {
        std::vector<double> ad(n), cd(3 * (size_t)n);
        std::vector<float> af(n);
        for (int k = 0; k < n; ++k) {
            ad[k] = 4.0 + masses[k];
            af[k] = (float)(ad[k] * ad[k]);
            cd[k] = ad[k] * ad[k];
            cd[n + k] = 1.0 / ad[k];
            cd[2 * n + k] = 0.5 / ad[k];
        }
        WD_HIP_CHECK(ctx, hipMemcpyAsync(diag_d.p, ad.data(), sizeof(double) * n,
                                         hipMemcpyHostToDevice, ctx->stream));
        WD_HIP_CHECK(ctx, hipMemcpyAsync(diag_f.p, af.data(), sizeof(float) * n,
                                         hipMemcpyHostToDevice, ctx->stream));
        WD_HIP_CHECK(ctx, hipMemcpyAsync(coef.p, cd.data(), sizeof(double) * 3 * n,
                                         hipMemcpyHostToDevice, ctx->stream));
        WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    const F64Links Ud = f64_links(U);
    const size_t xe = x->rhs_elems;
    const size_t srb = r.f->rhs_elems * prec_bytes_complex(sp);
    const size_t xsb = xs.f->rhs_elems * prec_bytes_complex(WD_F32);
    const double *ddiag = (const double *)diag_d.p;
    F64Batch all = {xe, nullptr, ddiag, nullptr};
    F64Batch rel = {xe, reliable, ddiag, nullptr};
    F64Batch rel2 = {xe, reliable, a2_d, nullptr};
    double2 *xd = (double2 *)x->data;
    const double2 *bd = (const double2 *)b->data;
    double2 *bhd = (double2 *)bh.f->data, *t64d = (double2 *)t64.f->data,
            *q64d = (double2 *)q64.f->data;

    // bh = a b_e + D_eo b_o / 2; M^dag b for the full-system check
    if ((rc = wd_launch_hop_f64_batch_c(ctx, bhd, bd, Ud, 0, 0, HOP_M, 0.5, bd, nullptr, nullptr,
                                        nullptr, all, n)) ||
        (rc = wd_launch_hop_f64_batch(ctx, (double2 *)Mfull.f->data, bd, Ud, 1, -1, HOP_M, bd,
                                      nullptr, nullptr, nullptr, all, n)) ||
        (rc = wd_launch_zero(ctx, x->data, x->bytes)))
        return rc;
    // r = r0 = bh, rho = (r0, r)
    if (sp == WD_F16)
        hipLaunchKernelGGL(k_bs_relr<WD_F16>, dim3(eg, n), dim3(kBlock), 0, ctx->stream,
                           (char *)r.f->data, r.f->scale, (char *)r0.f->data, r0.f->scale, true,
                           (const double2 *)bhd, (const double2 *)nullptr, xe, srb, stride, nel,
                           (const int *)nullptr, pa, pb, (const double2 *)Us->omega);
    else
        hipLaunchKernelGGL(k_bs_relr<WD_F32>, dim3(eg, n), dim3(kBlock), 0, ctx->stream,
                           (char *)r.f->data, r.f->scale, (char *)r0.f->data, r0.f->scale, true,
                           (const double2 *)bhd, (const double2 *)nullptr, xe, srb, stride, nel,
                           (const int *)nullptr, pa, pb, (const double2 *)Us->omega);
    // the Schur target starts at half the full-system tolerance (relative to |bh|)
    hipLaunchKernelGGL(k_bs_init, dim3(n), dim3(1024), 0, ctx->stream, (const double2 *)pa,
                       (const double2 *)pb, eg, st, active, reliable, 0.5 * prm->tol);
    hipLaunchKernelGGL(k_mx_fullres, dim3(eg_full, n), dim3(kBlock), 0, ctx->stream,
                       (const double2 *)Mfull.f->data, (const double2 *)nullptr, xe, stride,
                       (const int *)nullptr, pa);
    hipLaunchKernelGGL(k_bs_fullb, dim3(n), dim3(1024), 0, ctx->stream, (const double2 *)pa,
                       eg_full, st);
    hipLaunchKernelGGL(k_mx_alldone, dim3(1), dim3(64), 0, ctx->stream, active, n, stop);
    if ((rc = wd_check_launch(ctx, "mixed bicgstab init")))
        return rc;

    // sloppy A (two one-parity launches), with the (dotv, out) partials
    SloppyArgs As = {};
    As.rhs_bytes = srb;
    As.U = (const char *)Us->data;
    As.diag = (const float *)diag_f.p;
    As.active = active;
    As.partials = pa;
    As.stop = stop;
    As.nrhs = n;
    As.hc = -0.5f;
    int R = prm->rhs_per_thread;
    int C = prm->site_chunk;
    if (R == 0 || C < 0) {
        SloppyArgs T = As;
        T.out = (char *)v.f->data;
        T.out_scale = v.f->scale;
        T.in = T.same = (const char *)pv.f->data;
        T.in_scale = T.same_scale = pv.f->scale;
        int tR = 0, tC = 0, tT = 0;
        if ((rc = wd_tune_sloppy(ctx, T, sp, Us->prec, Us->recon, 0, R, &tR, &tC, &tT,
                                 nullptr)))
            return rc;
        R = tR;
        if (C < 0) {
            C = tC;
            As.tile = tT; // the work order timed with that chunk
        }
    }
    As.chunk = C;
    const int op_parts = hop_nsb * wd_sloppy_partials_per_block(R); // partials of the operator
    // norm: also the partials of |out|^2 into pb
    auto apply_A = [&](wd_field *out, const wd_field *in, const wd_field *dotv, bool norm) -> int {
        SloppyArgs m = As;
        m.out = (char *)out->data;
        m.out_scale = out->scale;
        m.in = m.same = (const char *)in->data;
        m.in_scale = m.same_scale = in->scale;
        m.dotv = (const char *)dotv->data;
        m.dotv_scale = dotv->scale;
        SloppyArgs h = m; // +D_oe in / 2 into the odd half of out
        h.hc = 0.5f;
        int e = wd_launch_hop_sloppy(ctx, h, sp, Us->prec, Us->recon, R, 0, 1, HOP_ONLY);
        if (e)
            return e;
        m.in = (const char *)out->data; // a^2 in_e - D_eo (odd half of out) / 2
        m.in_scale = out->scale;
        m.partials2 = pb;
        return wd_launch_hop_sloppy(ctx, m, sp, Us->prec, Us->recon, R, 0, 0,
                                    norm ? HOP_M_DOT_NORM : HOP_M_DOT);
    };
    // FP64 A v on the even half (the odd half of out holds +D_oe v / 2)
    auto apply_A64 = [&](const double2 *in, double2 *out, const F64Batch &b1, const F64Batch &b2,
                         const int *stp) -> int {
        int e = wd_launch_hop_f64_batch_c(ctx, out, in, Ud, 0, 1, HOP_ONLY, 0.5, nullptr, nullptr,
                                          nullptr, stp, b1, n);
        if (e)
            return e;
        return wd_launch_hop_f64_batch_c(ctx, out, out, Ud, 0, 0, HOP_M, -0.5, in, nullptr, nullptr,
                                         stp, b2, n);
    };

#define BS_SF(K, ...)                                                                              \
    do {                                                                                           \
        if (sp == WD_F16)                                                                          \
            hipLaunchKernelGGL(K<WD_F16>, dim3(eg, n), dim3(kBlock), 0, ctx->stream, __VA_ARGS__);  \
        else                                                                                       \
            hipLaunchKernelGGL(K<WD_F32>, dim3(eg, n), dim3(kBlock), 0, ctx->stream, __VA_ARGS__);  \
    } while (0)

    // buffer j of p or s: n systems at bbytes * j (scales at bscale * j)
    const size_t bbytes = (size_t)n * srb, bscale = (size_t)n * 2 * stride;
    auto buf = [&](const FieldBuf &f, int j) -> wd_field {
        wd_field b = *f.f;
        b.nrhs = n;
        b.data = (char *)f.f->data + j * bbytes;
        b.scale = f.f->scale ? f.f->scale + j * bscale : nullptr;
        return b;
    };

    // iteration whose p and s go into buffers pos
    auto iteration = [&](int pos) -> int {
        int e;
        const wd_field pp = buf(pv, pos), ss = buf(sv, pos);
        BS_SF(k_bs_p, (char *)pv.f->data, pv.f->scale, bbytes, bscale, pos,
              (const char *)r.f->data, (const float *)r.f->scale, (const char *)v.f->data,
              (const float *)v.f->scale, srb, stride, nel, (const BsState *)st,
              (const int *)active);
        if ((e = apply_A(v.f, &pp, r0.f, false)))
            return e;
        hipLaunchKernelGGL(k_bs_alpha, dim3(n), dim3(1024), 0, ctx->stream, (const double2 *)pa,
                           op_parts, st, active, pos);
        BS_SF(k_bs_s, (char *)ss.data, ss.scale, (const char *)r.f->data,
              (const float *)r.f->scale, (const char *)v.f->data, (const float *)v.f->scale, srb,
              stride, nel, (const BsState *)st, (const int *)active);
        if ((e = apply_A(t.f, &ss, &ss, true)))   // t = A s, (s, t) into pa, |t|^2 into pb
            return e;
        hipLaunchKernelGGL(k_bs_omega, dim3(n), dim3(1024), 0, ctx->stream, (const double2 *)pa,
                           op_parts, (const double2 *)pb, op_parts, st, active, pos);
        BS_SF(k_bs_update, (char *)r.f->data, r.f->scale, (const char *)ss.data,
              (const float *)ss.scale, (const char *)t.f->data, (const float *)t.f->scale,
              (const char *)r0.f->data, (const float *)r0.f->scale, srb, stride, nel,
              (const BsState *)st, (const int *)active, pa, pb);
        hipLaunchKernelGGL(k_bs_decide, dim3(n), dim3(1024), 0, ctx->stream, (const double2 *)pa,
                           (const double2 *)pb, eg, st, active, reliable, prm->max_iter,
                           prm->delta);
        return wd_check_launch(ctx, "mixed bicgstab iteration");
    };

    // x_s += the group's pending steps; x += x_s, x_s = 0 for the systems in due
    // (nullptr: none)
    auto flush = [&](const int *due) -> int {
        BS_SF(k_bs_flush, xd, (char *)xs.f->data, xe, xsb, stride, nel, due,
              (const double2 *)Us->omega, (const char *)pv.f->data, (const float *)pv.f->scale,
              (const char *)sv.f->data, (const float *)sv.f->scale, bbytes, bscale, srb,
              (const BsState *)st);
        return wd_check_launch(ctx, "k_bs_flush");
    };

    // reliable update of the waiting systems: x += x_s, r = bh - A x (FP64)
    auto reliable_slot = [&]() -> int {
        int e;
        if ((e = flush(reliable)) || (e = apply_A64(xd, t64d, rel, rel2, stop)))
            return e;
        BS_SF(k_bs_relr, (char *)r.f->data, r.f->scale, (char *)r0.f->data, r0.f->scale, false,
              (const double2 *)bhd, (const double2 *)t64d, xe, srb, stride, nel,
              (const int *)reliable, pa, pb, (const double2 *)Us->omega);
        hipLaunchKernelGGL(k_bs_relfin, dim3(n), dim3(1024), 0, ctx->stream, (const double2 *)pa,
                           (const double2 *)pb, eg, st, active, reliable, prm->max_iter);
        hipLaunchKernelGGL(k_mx_alldone, dim3(1), dim3(64), 0, ctx->stream, (const int *)active,
                           n, stop);
        return wd_check_launch(ctx, "mixed bicgstab reliable update");
    };
#undef BS_SF

    const int chunk = prm->check_every;
    // every group of iterations (at most npend) ends with a flush of the
    // pending x_s steps, fused into the reliable-update slot when one is due
    auto block = [&]() -> int {
        int e = 0;
        for (int k = 0, pos = 0; k < chunk && !e; ++k) {
            e = iteration(pos);
            const bool slot = (k + 1) % slot_every == 0 || k + 1 == chunk;
            if (!e && slot)
                e = reliable_slot();
            else if (!e && pos == npend - 1)
                e = flush(nullptr);
            pos = (slot || pos == npend - 1) ? 0 : pos + 1;
        }
        return e;
    };

    // x_o = b_o / a + D_oe x_e / (2a) and |M^dag b - M^dag M x| for the systems
    // that met the Schur target; the ones that miss tol resume.  The final
    // pass (after the loop) covers every system not yet accepted, resuming none.
    auto full_check = [&](int final_pass) -> int {
        int e;
        F64Batch chk_rec = {xe, check, inva_d, inv2a_d};
        F64Batch chk = {xe, check, ddiag, nullptr};
        if (final_pass)
            hipLaunchKernelGGL(k_mark_final<BsState>, dim3((n + 255) / 256), dim3(256), 0,
                               ctx->stream, (const BsState *)st, n, prm->tol, check);
        else
            hipLaunchKernelGGL(k_bs_markconv, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                               (const BsState *)st, n, check);
        if ((e = wd_check_launch(ctx, "k_bs_markconv")) ||
            (e = wd_launch_hop_f64_batch_c(ctx, xd, xd, Ud, 0, 1, HOP_M, 0.0, bd, nullptr,
                                           nullptr, nullptr, chk_rec, n)) ||
            (e = wd_launch_hop_f64_batch(ctx, t64d, xd, Ud, 0, -1, HOP_M, xd, nullptr, nullptr,
                                         nullptr, chk, n)) ||
            (e = wd_launch_hop_f64_batch(ctx, q64d, t64d, Ud, 1, -1, HOP_M, t64d, nullptr,
                                         nullptr, nullptr, chk, n)))
            return e;
        hipLaunchKernelGGL(k_mx_fullres, dim3(eg_full, n), dim3(kBlock), 0, ctx->stream,
                           (const double2 *)Mfull.f->data, (const double2 *)q64d, xe, stride,
                           (const int *)check, pa);
        hipLaunchKernelGGL(k_bs_fullfin, dim3(n), dim3(1024), 0, ctx->stream,
                           (const double2 *)pa, eg_full, st, active, check, prm->tol,
                           prm->max_iter, final_pass ? 0 : 1);
        if (!final_pass)
            hipLaunchKernelGGL(k_mx_alldone, dim3(1), dim3(64), 0, ctx->stream,
                               (const int *)active, n, stop);
        return wd_check_launch(ctx, "mixed bicgstab full check");
    };

    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    if (prm->use_graph) {
        WD_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int e = block();
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
            return wd_set_error(ctx, WD_EHIP, std::string("mixed bicgstab: graph instantiate: ") +
                                                  hipGetErrorString(ce));
        }
    }
    int status = WD_OK;
    long launched = 0;
    const long limit = (long)prm->max_iter + chunk;
    int checks = 0;
    bool checked = false;
    while (true) {
        if ((status = poll_flag(ctx, stop, "mixed bicgstab")) < 0)
            break;
        const int hstop = status;
        status = WD_OK;
        if (hstop && !checked && checks < 4) {
            ++checks;
            checked = true;
            if ((status = full_check(0)))
                break;
            continue;
        }
        if (hstop || launched > limit)
            break;
        checked = false;
        if (gexec) {
            hipError_t ge = hipGraphLaunch(gexec, ctx->stream);
            if (ge != hipSuccess) {
                status = wd_set_error(ctx, WD_EHIP, std::string("mixed bicgstab: graph launch: ") +
                                                        hipGetErrorString(ge));
                break;
            }
        } else if ((status = block())) {
            break;
        }
        launched += chunk;
    }
    if (gexec)
        (void)hipGraphExecDestroy(gexec);
    if (graph)
        (void)hipGraphDestroy(graph);
    // x_o and the full-system residual of every system not accepted by a check
    if (!status)
        status = full_check(1);
    if (status)
        return status;
    std::vector<BsState> hs(n);
    WD_HIP_CHECK(ctx, hipMemcpyAsync(hs.data(), st, sizeof(BsState) * n, hipMemcpyDeviceToHost,
                                     ctx->stream));
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds)
        *seconds = std::chrono::duration<double>(t1 - t0).count();
    bool all_conv = true, breakdown = false;
    for (int k = 0; k < n; ++k) {
        const double bn = hs[k].full_b, tr = hs[k].full_rsd;
        const bool conv = hs[k].status == 1 && (!(bn > 0.0) || tr <= prm->tol * bn);
        if (reports) {
            reports[k].iterations = hs[k].iter;
            reports[k].reliable_updates = hs[k].updates;
            reports[k].converged = conv;
            reports[k].true_rel_rsd = bn > 0 ? tr / bn : 0.0;
            reports[k].rhs_norm = bn;
            reports[k].rhs_per_thread = R;
            reports[k].site_chunk = C;
            reports[k].block_tile = As.tile;
        }
        all_conv = all_conv && conv;
        breakdown = breakdown || hs[k].status == 2;
    }
    if (breakdown)
        return wd_set_error(ctx, WD_EBREAKDOWN, "mixed bicgstab: breakdown in at least one system");
    return all_conv ? WD_OK : WD_ENOTCONV;
}

extern "C" int wd_tune_report(wd_ctx *ctx, const wd_field *f, const wd_gauge *Us, int *best_R,
                              int *best_chunk, double ms[WD_TUNE_NR * WD_TUNE_NC])
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, f && Us && best_R && best_chunk && ms, "tune: null argument");
    WD_ARG_CHECK(ctx, wd_sloppy_supported(f->prec, Us->prec, Us->recon),
                 "tune: unsupported precision combination");
    wd_field *out = nullptr;
    int rc = wd_field_create_work(ctx, &out, f->prec, f->nrhs);
    if (rc)
        return rc;
    DevBuf diag;
    if ((rc = diag.alloc(ctx, sizeof(float) * f->nrhs))) {
        wd_field_destroy(out);
        return rc;
    }
    SloppyArgs a = {};
    a.out = (char *)out->data;
    a.out_scale = out->scale;
    a.in = a.same = (const char *)f->data;
    a.in_scale = a.same_scale = f->scale;
    a.rhs_bytes = f->rhs_elems * prec_bytes_complex(f->prec);
    a.U = (const char *)Us->data;
    a.diag = (const float *)diag.p;
    a.nrhs = f->nrhs;
    a.hc = -0.5f;
    rc = wd_tune_sloppy(ctx, a, f->prec, Us->prec, Us->recon, -1, 0, best_R, best_chunk, nullptr,
                        ms);
    wd_field_destroy(out);
    return rc;
}

// Launches the sloppy M (cb = -1, both parities) or its one-parity part
// (cb = 0 or 1, the launch shape of the Schur solvers) `iters` times; with ms
// the average launch duration from HIP events on the context stream.
static int sloppy_apply(wd_ctx *ctx, wd_field *out, const wd_field *in, const wd_gauge *Us,
                        const double *masses, int dagger, int rhs_per_thread, int site_chunk,
                        int cb, int iters, double *ms)
{
    WD_ARG_CHECK(ctx, out && in && Us && masses, "apply_M_sloppy: null argument");
    WD_ARG_CHECK(ctx, out != in, "apply_M_sloppy: out and in must be distinct fields");
    WD_ARG_CHECK(ctx, out->prec == in->prec && out->nrhs == in->nrhs,
                 "apply_M_sloppy: out and in must have the same format and nrhs");
    WD_ARG_CHECK(ctx, wd_sloppy_supported(in->prec, Us->prec, Us->recon),
                 "apply_M_sloppy: unsupported precision combination");
    WD_ARG_CHECK(ctx, cb >= -1 && cb <= 1 && iters >= 1, "time_sloppy: bad cb or iteration count");
    const int n = in->nrhs;
    DevBuf diag;
    int rc = diag.alloc(ctx, sizeof(float) * n);
    if (rc)
        return rc;
    std::vector<float> af(n);
    for (int r = 0; r < n; ++r)
        af[r] = (float)(4.0 + masses[r]);
    WD_HIP_CHECK(ctx, hipMemcpyAsync(diag.p, af.data(), sizeof(float) * n, hipMemcpyHostToDevice,
                                     ctx->stream));
    SloppyArgs a = {};
    a.out = (char *)out->data;
    a.out_scale = out->scale;
    a.in = a.same = (const char *)in->data;
    a.in_scale = a.same_scale = in->scale;
    a.rhs_bytes = in->rhs_elems * prec_bytes_complex(in->prec);
    a.U = (const char *)Us->data;
    a.diag = (const float *)diag.p;
    a.nrhs = n;
    a.hc = -0.5f;
    a.chunk = site_chunk;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        // one untimed launch first (argument checks, first use of the kernel)
        if ((rc = wd_launch_hop_sloppy(ctx, a, in->prec, Us->prec, Us->recon, rhs_per_thread,
                                       dagger, cb, HOP_M)))
            return rc;
        WD_HIP_CHECK(ctx, hipEventCreate(&e0));
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return wd_set_error(ctx, WD_EHIP, "time_sloppy: event creation failed");
        }
        (void)hipEventRecord(e0, ctx->stream);
    }
    for (int k = 0; k < iters && !rc; ++k)
        rc = wd_launch_hop_sloppy(ctx, a, in->prec, Us->prec, Us->recon, rhs_per_thread, dagger,
                                  cb, HOP_M);
    if (ms) {
        hipError_t er = hipEventRecord(e1, ctx->stream);
        hipError_t es = er == hipSuccess ? hipEventSynchronize(e1) : er;
        float t = 0.f;
        hipError_t ee = es == hipSuccess ? hipEventElapsedTime(&t, e0, e1) : es;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (!rc && ee != hipSuccess)
            rc = wd_set_error(ctx, WD_EHIP, std::string("time_sloppy: ") + hipGetErrorString(ee));
        *ms = (double)t / iters;
    }
    if (rc)
        return rc;
    WD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WD_OK;
}

extern "C" int wd_apply_M_sloppy(wd_ctx *ctx, wd_field *out, const wd_field *in,
                                 const wd_gauge *Us, const double *masses, int dagger,
                                 int rhs_per_thread, int site_chunk)
{
    if (!ctx)
        return WD_EARG;
    return sloppy_apply(ctx, out, in, Us, masses, dagger, rhs_per_thread, site_chunk, -1, 1,
                        nullptr);
}

extern "C" int wd_time_sloppy(wd_ctx *ctx, wd_field *out, const wd_field *in, const wd_gauge *Us,
                              const double *masses, int rhs_per_thread, int site_chunk, int cb,
                              int iters, double *ms_per_launch)
{
    if (!ctx)
        return WD_EARG;
    WD_ARG_CHECK(ctx, ms_per_launch != nullptr, "time_sloppy: null result");
    return sloppy_apply(ctx, out, in, Us, masses, 0, rhs_per_thread, site_chunk, cb, iters,
                        ms_per_launch);
}
