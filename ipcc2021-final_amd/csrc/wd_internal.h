// wd_internal.h -- shared definitions of the HIP library behind include/wilson.h.
//
// Device storage (all fields even-odd checkerboarded, site index fastest;
// wd_fields.h has the exact formats):
//   F64 spinor : double2, per [rhs][parity] 64-site tiles [stride/64][12][64]
//   F32 spinor : float4   [rhs][parity][6][stride]  (2 complex per chunk)
//   F16 spinor : uint4    [rhs][parity][3][stride]  (4 complex as half2) + float scale
//   gauge      : per [mu][parity]: F64 tiles [stride/64][9 or 6][64], F32 float2[9][stride]
//                or float4[3][stride], F16 uint2[3][stride]  (row-major 3x3, 12-real = rows 0,1)
// with stride = Vh rounded up to a multiple of 64, so a wavefront reading one
// component of 64 consecutive sites issues one coalesced 1 KiB (F64) request.
// Padding sites are zero.
//
// Position i inside a parity block is the reference's eo index within that
// block: i = xh + Lxh*(y + Ly*(z + Lz*t)), x = 2*xh + ((parity+y+z+t) & 1)
// (include/operator.h:62-67).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <vector>
#include <string>
#include <tuple>

#include "../../include/wilson.h"

struct Geom {
    int Lxh, Ly, Lz, Lt; // half x extent and the other three extents
    int Vh;              // sites per parity
    int stride;          // component stride (>= Vh, multiple of 64)
    double bc[4];        // boundary phase applied to hops that cross the boundary in mu
    int tD;              // site blocks per time slice of one parity (0: slices not block aligned)
    int tW;              // block column width of the z-t tiled order (0: natural order)
};

struct wd_ctx {
    int L[4];
    int bc[4];
    int device;
    long V, Vh, stride;
    Geom geom;
    hipStream_t stream;
    // reduction partials and solver scratch
    void *partials;       // double2 [max_partials]
    int max_partials;
    void *dscalars;       // small device scratch for solver state
    size_t dscalars_bytes;
    void *hpinned;        // pinned host scratch (4 KiB)
    // host<->device staging for layout conversion
    void *staging;
    size_t staging_bytes;
    int xcd_remap;        // 1: map workgroups so each XCD owns a contiguous site range
    int parity_il;        // 1: both-parity FP64 launches alternate the parities per site block
    int ts_tb;            // t-run length of the t-streaming both-parity FP64 kernel, 0 = off
    int f64_pairs;        // 1: batched FP64 hops of several systems run two systems per thread
    int loop_prefetch;    // WD_RHS_LOOP kernel: neighbour loads 1 (3 wavefronts/SIMD) or 3 (2) ahead, 0 = by mode
    int loop_systems;     // WD_RHS_LOOP kernel: systems per pass (0 = automatic)
    // autotuned (rhs per thread, chunk, tile) of the sloppy operator, keyed by
    // (nrhs, spinor precision, 100 * gauge precision + recon + 1000 for one parity)
    std::map<std::tuple<int, int, int>, std::tuple<int, int, int>> tuned;
    // solver work blocks handed back by finished solves, (bytes, pointer),
    // reused by later solves that ask for the same size
    std::vector<std::pair<size_t, void *>> work;
    std::string last_error;
};

struct wd_field {
    wd_ctx *ctx;
    int prec;
    int nrhs;
    void *data;           // element type by prec
    float *scale;         // F16 only: per-site scale [rhs][parity][stride]
    size_t rhs_elems;     // elements (complex) per rhs = 24 * stride
    size_t bytes;
    bool work;            // data and scale belong to the context's work cache
};

struct wd_gauge {
    wd_ctx *ctx;
    int prec;
    int recon;
    void *data;
    size_t bytes;
    // WD_RECON_12T: the rotation Omega of the temporal gauge, FP64 complex
    // [parity][9][stride] (row-major 3x3 per site); nullptr otherwise
    double2 *omega;
};

// ---- device memory ----------------------------------------------------------
// hipMalloc that, when the device is full, releases the context's cached
// work blocks and tries once more.
hipError_t wd_malloc(wd_ctx *ctx, void **p, size_t bytes);
// Solver work memory.  A fresh hipMalloc of many GB after earlier frees costs
// the device seconds (measured on MI355X: 2.96 s for 140 GB, and 0.88 s to
// free it), so the blocks a solve hands back stay with the context and serve
// the next solve that asks for the same size.  wd_work_get returns nullptr
// when the memory cannot be had; the contents of a returned block are
// undefined.
void *wd_work_get(wd_ctx *ctx, size_t bytes);
void wd_work_put(wd_ctx *ctx, void *p, size_t bytes);
void wd_work_release(wd_ctx *ctx);
// A field whose memory comes from (and goes back to) the work cache;
// destroyed with wd_field_destroy.  Zeroed like wd_field_create's.
int wd_field_create_work(wd_ctx *ctx, wd_field **f, int prec, int nrhs);

// ---- error plumbing ---------------------------------------------------------
int wd_set_error(wd_ctx *ctx, int code, const std::string &msg);

#define WD_HIP_CHECK(ctx, expr)                                                                    \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return wd_set_error((ctx), WD_EHIP,                                                    \
                                std::string(#expr) + ": " + hipGetErrorString(_e));                \
    } while (0)

#define WD_ARG_CHECK(ctx, cond, msg)                                                               \
    do {                                                                                           \
        if (!(cond))                                                                               \
            return wd_set_error((ctx), WD_EARG, (msg));                                            \
    } while (0)

// Checks the launch status of the last kernel.
int wd_check_launch(wd_ctx *ctx, const char *what);

// ---- launch helpers ----------------------------------------------------------
constexpr int kBlock = 256;          // threads per workgroup for site kernels
constexpr int kRedBlocks = 1024;     // workgroups of the BLAS-1 / reduction kernels
constexpr int kStreamPartials = 16384; // largest grid of the fused streaming reductions
constexpr int kNumXCD = 8;           // MI355X: 8 accelerator complex dies

// Hardware hands workgroup b to XCD (b mod 8).  With remap on, logical block
// (b mod 8) * (n/8) + b/8 is processed instead, so each XCD (and its L2) owns
// one contiguous range of sites and neighbour spinors are reused within it.
__device__ __forceinline__ int logical_block(int remap)
{
    const int b = blockIdx.x, n = gridDim.x;
    if (remap && (n % kNumXCD) == 0)
        return (b % kNumXCD) * (n / kNumXCD) + b / kNumXCD;
    return b;
}

// Order of the site blocks of one parity.  Blocks of kBlock consecutive sites
// advance in x, then y, z, t; with D blocks per time slice the t neighbours of
// block b sit in blocks b -+ D, which the natural order reaches D blocks later,
// after the L2 has been refilled with links when D is large (64 blocks at
// 32^3 x 64).  In the tiled order the blocks run in columns of W: the first W
// blocks of slice 0, the same W blocks of slice 1, ..., of slice Lt-1, then
// the next W blocks of every slice.  A t neighbour is then W blocks away and
// only the z neighbours across a column edge leave the working set.  Returns
// the natural index of logical block lb; a bijection on [0, D * Lt).
__device__ __forceinline__ int block_order(int lb, const Geom &g)
{
    if (g.tW <= 0)
        return lb;
    const int per = g.tW * g.Lt;
    const int col = lb / per;
    const int w = min(g.tW, g.tD - col * g.tW); // the last column may be narrower
    const int rem = lb - col * per;
    const int t = rem / w;
    return col * g.tW + (rem - t * w) + g.tD * t;
}

// ---- precision helpers ---------------------------------------------------------
inline size_t prec_bytes_complex(int prec)
{
    return prec == WD_F64 ? 16 : prec == WD_F32 ? 8 : 4;
}

// ---- internal entry points shared between translation units --------------------
// FP64 links of a hop launch: 18-real (9 complex per link, the reference's
// format; what a bare pointer means) or 12-real (rows 0 and 1, row 2 rebuilt
// as conj(row0 x row1): SU(3) links only).
struct F64Links {
    const double2 *p;
    int recon;
    F64Links(const double2 *q, int r = WD_RECON_18) : p(q), recon(r) {}
};
inline F64Links f64_links(const wd_gauge *U)
{
    return F64Links(static_cast<const double2 *>(U->data), U->recon);
}
int wd_launch_hop_f64(wd_ctx *ctx, double2 *out, const double2 *in, F64Links U, int dagger,
                      int cb, int mode, double a, const double2 *same, const double2 *dotv,
                      double2 *partials, const int *stop);
int wd_launch_zero(wd_ctx *ctx, void *p, size_t bytes);
int wd_auto_tstream(const wd_ctx *ctx);
// device address of the allocation-free correlator's scratch, once per device
// (wd_contract.hip); called by wd_create with the device current
int wd_contract_setup(int device);
int wd_ensure_staging(wd_ctx *ctx, size_t bytes);

// layout / precision conversion (wd_convert.hip)
void *wd_rhs_ptr(const wd_field *f, int rhs);
float *wd_rhs_scale(const wd_field *f, int rhs);
int wd_launch_spinor_from_ref(wd_ctx *ctx, wd_field *f, int rhs, const double2 *src);
int wd_launch_spinor_to_ref(wd_ctx *ctx, const wd_field *f, int rhs, double2 *dst);
int wd_launch_gauge_from_ref(wd_ctx *ctx, wd_gauge *U, int mu, const double2 *src);
// Temporal gauge (WD_RECON_12T): Omega from the t links in ref layout
// (src), then the rotated links of direction mu from src into U.
int wd_launch_tg_omega(wd_ctx *ctx, wd_gauge *U, const double2 *src);
int wd_launch_tg_links(wd_ctx *ctx, wd_gauge *U, int mu, const double2 *src);
int wd_convert_rhs(wd_ctx *ctx, wd_field *dst, int rd, const wd_field *src, int rs);

// batch description for the FP64 kernel: rhs r = blockIdx.y at element offset
// r * rhs_elems; mask[r] == 0 skips it; a_rhs[r] / c_rhs[r] (when not null)
// replace the diagonal and the hop coefficient
struct F64Batch {
    size_t rhs_elems;
    const int *mask;
    const double *a_rhs;
    const double *c_rhs;
    const double *axpy; // HOP_M_AXPY: the (real) coefficient alpha, in device memory
};
int wd_launch_hop_f64_batch(wd_ctx *ctx, double2 *out, const double2 *in, F64Links U,
                            int dagger, int cb, int mode, const double2 *same,
                            const double2 *dotv, double2 *partials, const int *stop,
                            const F64Batch &b, int nrhs);
// general hop coefficient: out = a same + hc hop(in), hc = -1/2 for the Wilson M
int wd_launch_hop_f64c(wd_ctx *ctx, double2 *out, const double2 *in, F64Links U, int dagger,
                       int cb, int mode, double a, double hc, const double2 *same,
                       const double2 *dotv, double2 *partials, const int *stop);
int wd_launch_hop_f64_batch_c(wd_ctx *ctx, double2 *out, const double2 *in, F64Links U,
                              int dagger, int cb, int mode, double hc, const double2 *same,
                              const double2 *dotv, double2 *partials, const int *stop,
                              const F64Batch &b, int nrhs);

// hop modes for the fused FP64 kernel
enum HopMode {
    HOP_ONLY = 0,     // out[p] = -1/2 hop(in)
    HOP_M = 1,        // out[p] = a * same[p] - 1/2 hop(in)  (same == in for M)
    HOP_M_DOT = 2,    // as HOP_M, plus per-block partial of sum conj(dotv) * out
    HOP_M_DOT_NORM = 3, // as HOP_M_DOT, plus per-block partial of |out|^2 (sloppy kernel)
    HOP_M_NORM = 4,   // as HOP_M, plus per-block partial (|out|^2, 0) (sloppy kernels: no dagger)
    HOP_M_AXPY = 5,   // out = dotv - alpha (M in), partial (|out|^2, 0) (sloppy kernels: dagger only)
};
// r -= alpha (a same - 1/2 hop(in)) with the partials of |r|^2 (one system,
// alpha real, read from device memory)
int wd_launch_hop_f64_axpy(wd_ctx *ctx, double2 *r, const double2 *in, F64Links U, int dagger,
                           int cb, double a, const double2 *same, const double *alpha,
                           double2 *partials, const int *stop);
