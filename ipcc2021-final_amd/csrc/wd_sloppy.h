// wd_sloppy.h -- arguments and launchers of the FP32-arithmetic multi-rhs
// operator (wd_hop_sloppy.hip) used by the mixed-precision solvers.
#pragma once

#include "wd_internal.h"

struct SloppyArgs {
    char *out;               // field base (rhs 0), spinor format SF
    float *out_scale;        // F16 scales (rhs 0) or nullptr
    const char *in;
    const float *in_scale;
    const char *same;        // HOP_M / HOP_M_DOT: the diagonal term's input
    const float *same_scale;
    const char *dotv;        // HOP_M_DOT: partial of sum conj(dotv) * out; HOP_M_AXPY: the
    const float *dotv_scale; //   field the update is applied to (out = dotv - axpy_r M in)
    const float *axpy;       // [nrhs] HOP_M_AXPY coefficients
    size_t rhs_bytes;        // spinor bytes per rhs (all fields share the format)
    const char *U;           // gauge field
    const float *diag;       // [nrhs] a_r = 4 + m_r
    const int *active;       // [nrhs] 1 = apply to this rhs, else skip; nullptr = all
    double2 *partials;       // [nrhs][nsb] for HOP_M_DOT
    double2 *partials2;      // [nrhs][nsb] (|out|^2, 0) for HOP_M_DOT_NORM
    const int *stop;         // whole launch becomes a no-op when *stop != 0
    int nrhs;
    float hc;                // hop coefficient: out = diag in + hc hop(in); -1/2 for M
    int chunk;               // site blocks per L2 chunk (0: groups of one block consecutive);
                             // WD_RHS_LOOP: systems per pass (0: the context's choice)
    float one;               // 1.0f, set by the launcher (an operand the compiler cannot fold)
    int tile;                // work order: 0 = z-t tiled columns (the context's), -1 = natural order
    // filled by the launcher
    int groups, bpp, nsb;
    int tlinks;              // temporal link loads (links re-read within the launch)
    int tg;                  // links in temporal gauge (WD_RECON_12T): t links are the
                             // identity except on the last slice
};

bool wd_sloppy_supported(int spinor_prec, int gauge_prec, int recon);
// formats of the kernel that loops over every system of a site (WD_RHS_LOOP)
bool wd_sloppy_loop_supported(int spinor_prec, int gauge_prec, int recon);
int wd_sloppy_groups(int nrhs, int R);
// cb in {0, 1}: one output parity; cb = -1: both parities.  R in {1, 2, 4} or
// WD_RHS_LOOP (wd_hop_sloppy_loop.hip).
int wd_launch_hop_sloppy(wd_ctx *ctx, SloppyArgs a, int spinor_prec, int gauge_prec, int recon,
                         int R, int dagger, int cb, int mode);
int wd_launch_hop_sloppy_loop(wd_ctx *ctx, const SloppyArgs &a, const Geom &g, int dagger, int cb,
                              int mode, int gauge_prec);
// Largest system group of a WD_RHS_LOOP launch over nrhs systems: `chunk`
// (the site-chunk argument, which for this kernel counts systems per pass)
// when > 0, else WD_OPT_LOOP_SYSTEMS when set, else 12.
int wd_loop_group_cap(const wd_ctx *ctx, int nrhs, int chunk);
// partial sums per workgroup of the HOP_M_DOT / HOP_M_DOT_NORM launches: the
// WD_RHS_LOOP kernel writes one per wavefront, the others one per workgroup
constexpr int kLoopWaves = kBlock / 64;
inline int wd_sloppy_partials_per_block(int R) { return R == WD_RHS_LOOP ? kLoopWaves : 1; }
