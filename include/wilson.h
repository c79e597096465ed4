/*
 * wilson.h -- C-ABI of the MI355X Wilson-Dslash / CG library (libwilson.so).
 *
 * Plain C: opaque handles, pointers and sizes.  The C++ operator API that
 * mirrors the reference (lattice_fermion, lattice_gauge, DslashNew, CGinvert,
 * ...) is a thin layer over these entry points, and foreign callers (ctypes,
 * cgo, JNI) bind them directly.
 *
 * Conventions
 *  - Host arrays use the reference layout ("ref" layout): a spinor field is
 *    complex<double>[V*12] at eo(x)*12 + spin*3 + colour, a gauge field is four
 *    complex<double>[V*9] arrays (one per direction x,y,z,t) at
 *    eo(x)*9 + row*3 + col, with eo(x) = lex(x)/2 + parity(x)*V/2 and
 *    lex(x) = x0 + L0*(x1 + L1*(x2 + L2*x3))  (include/operator.h:56-67).
 *  - Device storage is owned by the library (even-odd, structure of arrays,
 *    site index fastest).  Host buffers are owned by the caller.
 *  - One wd_ctx = one device + one HIP stream; no internal locking (use one
 *    context per host thread).  Every call is stream-ordered and returns only
 *    when its results are complete, like the reference's blocking calls.
 *  - Return codes: 0 success, < 0 error (see WD_E*), > 0 "not converged".
 *    wd_strerror() names a code; wd_last_error() gives the context's last
 *    message.
 *  - `out` and `in` must be distinct fields (the reference does not support
 *    aliasing either: DslashNew zeroes dest before reading src).
 */
#ifndef WILSON_H
#define WILSON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define WD_OK 0
#define WD_EARG (-1)        /* invalid argument (shape, precision, null pointer ...) */
#define WD_EHIP (-2)        /* a HIP runtime call failed */
#define WD_ENOMEM (-3)      /* device or host allocation failed */
#define WD_EBREAKDOWN (-4)  /* CG breakdown: non-finite or non-positive (p, A p) */
#define WD_ENOTCONV 1       /* the solver stopped at its iteration limit */

/* ---- precisions and link formats ---------------------------------------- */
#define WD_F64 0   /* double */
#define WD_F32 1   /* float */
#define WD_F16 2   /* half (spinors: with one float scale per site) */

#define WD_RECON_18 18 /* links stored as all 9 complex entries */
#define WD_RECON_12 12 /* rows 0 and 1 stored; row 2 = conj(row0 x row1) (SU(3) only) */
/* Sloppy links (F32, F16) stored 12-real in temporal gauge: U'_mu(x) =
 * Omega(x) U_mu(x) Omega(x + mu)^dag with Omega(x, y, z, t) = U_t(t = 0) ...
 * U_t(t - 1) along the site's t column (Omega = 1 on the first slice), so
 * every t link but those of the last slice is the identity and the sloppy
 * operators load and multiply t links only there (SU(3) links only).  The
 * mixed solvers rotate their sloppy residuals by Omega and the corrections
 * back by Omega^dag at the FP64 reliable updates; wd_apply_M_sloppy and
 * wd_time_sloppy apply Omega M Omega^dag.  No counterpart in the reference. */
#define WD_RECON_12T 13

typedef struct wd_ctx wd_ctx;
typedef struct wd_field wd_field;
typedef struct wd_gauge wd_gauge;

/* ---- context ------------------------------------------------------------ */
/* L = {Lx, Ly, Lz, Lt}, every extent even and >= 2.  bc[mu] = +1 periodic,
 * -1 antiperiodic (NULL = periodic everywhere, the reference's only boundary
 * condition, src/dslash_new.cpp:1847).  device = HIP ordinal.
 * Replaces the reference's lattice set-up in src/main.cpp:17-66. */
int wd_create(wd_ctx **ctx, const int L[4], const int bc[4], int device);
void wd_destroy(wd_ctx *ctx);
const char *wd_strerror(int code);
const char *wd_last_error(const wd_ctx *ctx);
/* Blocks until all work on the context's stream is complete. */
int wd_sync(wd_ctx *ctx);
/* Volume V = Lx*Ly*Lz*Lt of the context. */
long wd_volume(const wd_ctx *ctx);
/* Solver work memory (no counterpart in the reference, whose CGinvert
 * constructs its work vectors on every call, src/invert.cpp:30-37).  The
 * work vectors of a finished solve stay with the context and serve the next
 * solve of the same shape: on the device a fresh allocation of many GB after
 * earlier frees costs seconds.  When the device runs out of memory for a
 * field, gauge or solve, the cached blocks are released and the allocation
 * is retried.  wd_release_workspace frees them now; wd_workspace_bytes
 * reports how many bytes are cached. */
int wd_release_workspace(wd_ctx *ctx);
int wd_workspace_bytes(const wd_ctx *ctx, size_t *bytes);

/* ---- gauge field (lattice_gauge, include/lattice_gauge.h:11-26) ---------- */
/* prec: WD_F64 | WD_F32 | WD_F16 ; recon: WD_RECON_18 | WD_RECON_12 (|
 * WD_RECON_12T for F32 and F16: the mixed solvers' sloppy links).  The
 * FP64 operators and solvers (wd_hop, wd_apply_M, wd_apply_MdagM, wd_cg,
 * wd_cg_eo, the reliable updates of the mixed solvers) take F64 links in
 * either form: 18-real is the reference's format and applies any links;
 * 12-real rebuilds row 2 and is for SU(3) links only, a third fewer link
 * bytes (FP64 M at 32^3x64 0.455 -> 0.378 ms on MI355X). */
int wd_gauge_create(wd_ctx *ctx, wd_gauge **U, int prec, int recon);
/* A[mu] = complex<double>[V*9] in ref layout (the reference's U.A[mu]). */
int wd_gauge_upload_ref(wd_ctx *ctx, wd_gauge *U, const double *const A[4]);
void wd_gauge_destroy(wd_gauge *U);

/* ---- spinor fields (lattice_fermion, include/lattice_fermion.h:12-35) ---- */
/* A field holds nrhs full-lattice spinors (both parities) in precision prec.
 * Fresh fields are zero. */
int wd_field_create(wd_ctx *ctx, wd_field **f, int prec, int nrhs);
void wd_field_destroy(wd_field *f);
/* A = complex<double>[V*12] in ref layout (the reference's lattice_fermion.A). */
int wd_field_upload_ref(wd_ctx *ctx, wd_field *f, const double *A, int rhs);
int wd_field_download_ref(wd_ctx *ctx, const wd_field *f, double *A, int rhs);
/* f = 0 (lattice_fermion::clean, src/lattice_fermion.cpp:44-49). */
int wd_field_zero(wd_ctx *ctx, wd_field *f);
/* dst = src, converting precision when the two differ (same nrhs). */
int wd_field_copy(wd_ctx *ctx, wd_field *dst, const wd_field *src);

/* ---- operator (F64 fields and an F64 gauge, 18- or 12-real) -------------- */
/* Hopping term: out[cb] = -1/2 sum_mu [(1 - f g_mu) U_mu(x) in(x+mu)
 *   + (1 + f g_mu) U_mu^dag(x-mu) in(x-mu)] for sites of parity cb, read from
 * parity 1-cb; out[1-cb] = 0.  f = -1 when dagger.
 * Replaces DslashoffdNew (src/dslash_new.cpp:1341-2677) / Dslashoffd
 * (src/dslash.cpp:38-1249). */
int wd_hop(wd_ctx *ctx, wd_field *out, const wd_field *in, const wd_gauge *U, int dagger, int cb);
/* out = (4 + mass) in over the full lattice.
 * Replaces DslashEEOONew (src/dslash_new.cpp:1301-1338); DslashEE / DslashOO
 * (src/dslash.cpp:11-35) are the two halves of it. */
int wd_diag(wd_ctx *ctx, wd_field *out, const wd_field *in, double mass);
/* out = M in (dagger = 0) or M^dagger in (dagger = 1), one fused pass.
 * Replaces DslashNew (src/invert.cpp:138-156) and Dslash (src/invert.cpp:123-136). */
int wd_apply_M(wd_ctx *ctx, wd_field *out, const wd_field *in, const wd_gauge *U, double mass,
               int dagger);
/* out = M^dagger M in; tmp is caller scratch of the same shape.
 * The normal operator of src/invert.cpp:79-80. */
int wd_apply_MdagM(wd_ctx *ctx, wd_field *out, const wd_field *in, wd_field *tmp,
                   const wd_gauge *U, double mass);

/* ---- reductions (include/operator_mpi.h:15-79) ---------------------------- */
/* *out = sum |f_i|^2 over all 12V complex entries (norm_2). */
int wd_norm2(wd_ctx *ctx, const wd_field *f, double *out);
/* out = sum conj(a_i) b_i (vector_p), out[0] real, out[1] imaginary part. */
int wd_dot(wd_ctx *ctx, const wd_field *a, const wd_field *b, double out[2]);

/* ---- solvers ------------------------------------------------------------ */
typedef struct wd_cg_params {
    int max_iter;        /* the reference's `max`: the loop runs f = 1 .. max_iter-1 */
    double tol;          /* stop threshold on |r| */
    int tol_relative;    /* 0: |r| < tol (the reference's absolute test, src/invert.cpp:108);
                            1: |r| < tol * |M^dagger b| */
    int use_x0;          /* 1: start from the incoming x (the reference seeds it with rand(),
                            src/invert.cpp:42-44); 0: start from x = 0 */
    int check_every;     /* iterations launched between host-side convergence checks (>= 1) */
    int use_graph;       /* 1: replay one captured hipGraph per check_every iterations */
    double *rsd_history; /* optional host array [max_iter]: |r| after each iteration */
} wd_cg_params;

typedef struct wd_cg_report {
    int iterations;      /* f at which the stop test passed (or the last f run) */
    int converged;       /* 1 when the stop test passed */
    double rsd;          /* |r| of the recursive residual at exit */
    double true_rsd;     /* |M^dagger b - M^dagger M x| recomputed in FP64 at exit */
    double rhs_norm;     /* |M^dagger b| */
    double seconds;      /* wall time of the solve on the device stream */
    long operator_applies; /* applications of M or M^dagger */
} wd_cg_report;

/* Fills params with the reference's settings: max 500, absolute tol 1e-9
 * (src/main.cpp:47-48), use_x0 = 1, check_every 8, use_graph 1. */
void wd_cg_params_default(wd_cg_params *p);

/* CGNR on M^dagger M x = M^dagger b, FP64 throughout, the reference's algorithm
 * (complex alpha and beta, same recurrences, same stop test).  x (F64) holds x0
 * on entry when params->use_x0, the solution on return; b is F64.
 * Replaces CGinvert (src/invert.cpp:15-121).  Returns 0, WD_ENOTCONV or < 0. */
int wd_cg(wd_ctx *ctx, wd_field *x, const wd_field *b, const wd_gauge *U, double mass,
          const wd_cg_params *params, wd_cg_report *report);

/* Schur even-odd preconditioned CGNR, FP64.  No counterpart in the reference's
 * code: the problem statement (doc/IPCC_2021.pdf p.11-12) describes the Schur
 * complement, the reference solves the full system only.  The even half solves
 * A x_e = a b_e + D_eo b_o / 2 with A = a^2 - D_eo D_oe / 4 (a = 4 + m) by CGNR
 * on A^dagger A; then x_o = (b_o + D_oe x_e / 2) / a.  Same parameters and
 * report as wd_cg; the stop test is the reference's, on the FP64 residual
 * |M^dagger b - M^dagger M x| of the full system (params->tol, relative or
 * absolute), checked after the reconstruction (CG resumes from x_e with a
 * tighter Schur threshold if it was not met, up to 3 times).  report->rsd is
 * the last Schur residual, iterations count Schur iterations, operator_applies
 * counts applications of A, A^dagger, M or M^dagger (each costs one
 * full-lattice hop).  rsd_history receives the Schur residual per iteration.
 * The solution equals wd_cg's to the tolerance, not bitwise; iteration counts
 * are lower. */
int wd_cg_eo(wd_ctx *ctx, wd_field *x, const wd_field *b, const wd_gauge *U, double mass,
             const wd_cg_params *params, wd_cg_report *report);

/* ---- mixed precision and many right-hand sides ----------------------------------
 * No counterpart in the reference (its only solver is the FP64 CGinvert,
 * src/invert.cpp:25-121).  The inner iterations use FP32 arithmetic on
 * F16 (per-site scaled) or F32 spinors and a sloppy gauge copy (F32 18/12-real
 * or F16 12-real); FP64 is kept for reductions, the solution and the true
 * residual of the reliable updates. */
/* rhs_per_thread value: every system of a site in one thread, one after the
 * other, with the site's links loaded once per pass and kept in registers.
 * F16 spinors with 12-real links.  F16 links are kept as packed halves and
 * the SU(3) products are packed F16 dot products accumulated in FP32 (the
 * half spinor and the rebuilt third link row rounded to F16): results differ
 * from R = 1, 2, 4 by those roundings.  F32 links are kept in FP32 with the
 * arithmetic of R = 1, 2, 4: the operator's results are bit-identical to
 * theirs. */
#define WD_RHS_LOOP 64

typedef struct wd_mixed_params {
    int max_iter;       /* inner iterations per system */
    double tol;         /* stop when the FP64 |M^dag b - M^dag M x| < tol |M^dag b| */
    double delta;       /* reliable update when the sloppy |r| < delta x |r| at the last one */
    int spinor_prec;    /* WD_F16 or WD_F32 storage of the inner vectors */
    int rhs_per_thread; /* 1, 2 or 4 systems per thread, WD_RHS_LOOP; 0 = autotune
                           (among the WD_TUNE_NR candidates) */
    int site_chunk;     /* site blocks (256 sites) per L2 chunk of the multi-system
                           operator; 0 = off, -1 = autotune (with rhs_per_thread).
                           With WD_RHS_LOOP: systems per pass of the kernel
                           (0 = WD_OPT_LOOP_SYSTEMS) */
    int check_every;    /* iterations between host polls of the convergence flags */
    int use_graph;      /* replay each block of check_every iterations as one hipGraph */
    int even_odd;       /* 1: iterate on the Schur complement of the even sites (as
                           wd_cg_eo; x_o reconstructed, tol checked on the full system
                           with the FP64 operator), 0: on the full M^dag M */
    int reliable_every; /* iterations between the FP64 reliable-update slots; a system
                           whose update falls due waits, without iterating, for the
                           next slot (iteration counts do not change); 0 = automatic */
} wd_mixed_params;

typedef struct wd_mixed_report {
    int iterations;       /* inner iterations */
    int reliable_updates; /* FP64 true-residual recomputations */
    int converged;        /* 1 when the true residual met tol */
    double true_rel_rsd;  /* |M^dag b - M^dag M x| / |M^dag b| at the last reliable update
                             (even_odd: at the last check of the full system) */
    double rhs_norm;      /* |M^dag b| */
    int rhs_per_thread;   /* the rhs-per-thread the operator ran with */
    int site_chunk;       /* the L2 chunk the operator ran with (WD_RHS_LOOP: systems
                             per pass, 0 = the context's choice) */
    int block_tile;       /* the operator's work order: 0 = z-t tiled columns
                             (WD_OPT_BLOCK_TILE), -1 = natural site order (chosen by
                             the autotuner) */
} wd_mixed_report;

/* max_iter 100000, tol 1e-10, delta 2e-2, F16 spinors, autotuned rhs per
 * thread and site chunk, check_every 16, graphs on, full system (even_odd 0),
 * automatic reliable-update slots (reliable_every 0). */
void wd_mixed_params_default(wd_mixed_params *p);

/* Solves M_r^dag M_r x_r = M_r^dag b_r for all nrhs systems of b at once
 * (x, b: F64 fields with the same nrhs; masses[r] = m_r, host array), x_r
 * starting from 0.  U: F64 18-real gauge (reliable updates); Us: the sloppy
 * copy of the same links.  reports (host, nrhs entries) and seconds may be
 * NULL.  Returns 0 when every system converged, WD_ENOTCONV, or < 0. */
int wd_cg_mixed(wd_ctx *ctx, wd_field *x, const wd_field *b, const wd_gauge *U,
                const wd_gauge *Us, const double *masses, const wd_mixed_params *params,
                wd_mixed_report *reports, double *seconds);

/* Mixed-precision BiCGStab (no counterpart in the reference, whose only
 * solver is CGinvert): every system is solved on its Schur complement,
 * A x_e = a b_e + D_eo b_o / 2 with A = a^2 - D_eo D_oe / 4, by BiCGStab with
 * reliable updates -- two sloppy applications of A per iteration, like the
 * CGNR's A and A^dag, but with A's condition number instead of A^dag A's --
 * then x_o = (b_o + D_oe x_e / 2) / a and the same full-system FP64 check of
 * the reference's stop test |M^dag b - M^dag M x| < tol |M^dag b| as
 * wd_cg_mixed with even_odd (systems that miss it resume).  Parameters,
 * reports and return values as wd_cg_mixed (even_odd ignored; iterations
 * counts BiCGStab iterations). */
int wd_bicgstab_mixed(wd_ctx *ctx, wd_field *x, const wd_field *b, const wd_gauge *U,
                      const wd_gauge *Us, const double *masses, const wd_mixed_params *params,
                      wd_mixed_report *reports, double *seconds);

/* out_r = M_r in_r (dagger 0) or M_r^dag in_r with the sloppy operator for all
 * nrhs systems (fields F16 or F32 of the same format, masses[r] host array),
 * rhs_per_thread 1, 2, 4 or WD_RHS_LOOP, site_chunk >= 0 (work order only:
 * every chunk gives bit-identical results; with WD_RHS_LOOP it is the largest
 * number of systems per pass, 0 = WD_OPT_LOOP_SYSTEMS). */
int wd_apply_M_sloppy(wd_ctx *ctx, wd_field *out, const wd_field *in, const wd_gauge *Us,
                      const double *masses, int dagger, int rhs_per_thread, int site_chunk);

/* Average duration of the sloppy M over in's systems (cb = -1, both parities)
 * or of its one-parity launch (cb = 0 or 1, the shape of the Schur solvers'
 * operator), iters launches timed with HIP events on the context stream after
 * one untimed launch; rhs_per_thread 1, 2, 4 or WD_RHS_LOOP.  No counterpart in
 * the reference (measurement). */
int wd_time_sloppy(wd_ctx *ctx, wd_field *out, const wd_field *in, const wd_gauge *Us,
                   const double *masses, int rhs_per_thread, int site_chunk, int cb, int iters,
                   double *ms_per_launch);

/* Candidates of the autotuner: rhs per thread {1, 2, 4, WD_RHS_LOOP} x site
 * chunk {0, 16, 32, 64}.  WD_RHS_LOOP is timed for F16 spinors with F16 or
 * F32 12-real links from 8 systems on, its chunk values counting systems per
 * pass (0 = WD_OPT_LOOP_SYSTEMS, 12 by default; values that give the same
 * number of passes are timed once).  The operator gives the same bits for
 * every candidate except WD_RHS_LOOP with F16 links (its F16 roundings; all
 * its pass sizes give the same bits).  The choice is cached per context. */
#define WD_TUNE_NR 4
#define WD_TUNE_NC 4

/* Autotuner: times the sloppy M over f's nrhs systems for every candidate
 * (ms[ir * WD_TUNE_NC + ic], HIP events on the context stream; candidates that
 * cannot differ from chunk 0 because f has one thread group are not timed and
 * read -1) and records the fastest (*best_R, *best_chunk) for later solves of
 * that shape.  Each candidate runs in two work orders, the z-t tiled columns
 * and the natural site order (when they differ); ms holds the faster, and the
 * solvers use the faster (wd_mixed_report.block_tile).  Work orders give
 * bit-identical results. */
int wd_tune_report(wd_ctx *ctx, const wd_field *f, const wd_gauge *Us, int *best_R,
                   int *best_chunk, double ms[WD_TUNE_NR * WD_TUNE_NC]);

/* ---- propagators and contractions ----------------------------------------------
 * The consumer of point-source solves in the reference's Python model
 * (refs/python/main.py:50-65, refs/python/contraction.py:25-45). */

/* Pion correlator of the point-source propagators S_r, r = rhs_first ..
 * rhs_first + rhs_count - 1 (the 12 spin x colour solutions of one source,
 * F64 field): Ct[dt] = sum_r sum_{x : t(x) = (t0 + dt) mod Lt} |S_r(x)|^2 for
 * dt = 0 .. Lt-1 (host array, Lt entries), t0 the source time slice.
 * Deterministic. */
int wd_pion_correlator(wd_ctx *ctx, const wd_field *S, int rhs_first, int rhs_count, int t0,
                       double *Ct);

/* wd_pion_correlator without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device or host memory (tight serving
 * loops after each batched solve, code that forbids hipMalloc).  Same
 * arguments and the same result, bit for bit; the per-slice partial sums live
 * in a fixed device array of the library (16384 + 512 doubles, one per GPU,
 * held until the process ends), so rhs_count * 2 * Lt must not exceed 16384
 * and Lt not 512 (WD_EARG otherwise).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (contexts used from several host
 * threads) race and must be serialised by the caller. */
int wd_pion_correlator_noalloc(wd_ctx *ctx, const wd_field *S, int rhs_first, int rhs_count,
                               int t0, double *Ct);

/* ---- measurement and tuning --------------------------------------------------
 * No counterpart in the reference, which times whole phases with MPI_Wtime
 * (src/main.cpp:88-110). */
#define WD_OP_HOP 0 /* one-parity hop (wd_hop without zeroing the other half) */
#define WD_OP_M 1   /* fused M / M^dagger over both parities (wd_apply_M) */
/* Launches the operator `iters` times back to back on the context stream and
 * returns the average launch duration measured with HIP events on that
 * stream. */
int wd_time_op(wd_ctx *ctx, int op, wd_field *out, const wd_field *in, const wd_gauge *U,
               double mass, int dagger, int cb, int iters, double *ms_per_launch);

/* MFMA-vs-VALU experiment for the multi-system operator's link product
 * ("3x3 by 3x2N": one F16 SU(3) link times the F16 half spinors of N systems at
 * its site, FP32 accumulation).  No counterpart in the reference (AVX2 CPU
 * code).  Inputs are host arrays: U [nsites][3][3] and h [nsites][nsys][2][3]
 * of packed (re, im) F16 pairs (real part in the low 16 bits); out [nsites]
 * [nsys][2][3] (re, im) floats receives the products summed over `inner`
 * repetitions (inner x U h up to FP32 rounding).  The kernel runs once
 * untimed and `iters` times timed (HIP events on the context stream);
 * ms_per_launch is the average.  WD_LINKPROD_VALU: one site per lane, packed
 * F16 dot products (the loop sloppy kernel's arithmetic); WD_LINKPROD_MFMA:
 * v_mfma_f32_16x16x16f16 with two sites' links block-diagonal in one tile and
 * 8 systems x 2 spins as its columns.  nsites even, nsys in 1 .. 64. */
#define WD_LINKPROD_VALU 0
#define WD_LINKPROD_MFMA 1
int wd_link_product(wd_ctx *ctx, int impl, int nsites, int nsys, const uint32_t *U,
                    const uint32_t *h, int inner, int iters, float *out, double *ms_per_launch);

#define WD_OPT_XCD_REMAP 1 /* 1 (default): each XCD processes a contiguous site range */
/* Work order of the Dslash kernels: 0 = natural site order; W > 0 = z-t tiled
 * order with columns of W site blocks (256 sites each), so the t neighbours of
 * a block are read W blocks later instead of one time slice later.  Takes
 * effect when a time slice of one parity (L0/2 * L1 * L2 sites) is a multiple
 * of 256 sites and W is below the blocks per slice; else the natural order.
 * Default: columns of 4 z-planes, W = round(4 * L0/2 * L1 / 256).  Every W
 * gives bit-identical results. */
#define WD_OPT_BLOCK_TILE 2
/* Parity order of the both-parity FP64 launches (M, M^dagger) and of the
 * WD_RHS_LOOP sloppy kernel: 0 = all blocks of parity 0, then parity 1; 1
 * (default) = the two parities of each site block side by side, sharing each
 * link's forward and backward use through L2.  The sloppy kernel with R
 * systems per thread always runs the parities side by side; this option
 * makes its links load through L2 (1) or non-temporal (0).  Results are
 * bit-identical. */
#define WD_OPT_PARITY_ORDER 3
/* t-streaming of the both-parity FP64 launches (M, M^dagger and the CG's
 * M^dagger with its dot product, one system): each thread owns a column
 * (x, y, z) for a run of T consecutive t, keeping the t link and the t
 * neighbour of a site in registers for the next site.  0 = off (site-block
 * kernel); 1 (default) = automatic: T = 16 when Lt is a multiple of 16 and the
 * launch still has at least 512 workgroups, else off; T >= 2 forces runs of T
 * (Lt must be a multiple of T).  M and the hop are bit-identical to the
 * site-block kernel; the dot product's partial sums are grouped differently. */
#define WD_OPT_TSTREAM 4
/* FP64 hops over several systems (the mixed solvers' reliable updates and
 * full-system checks): 1 (default) = two systems per thread, each link loaded
 * once for both (links are three quarters of an FP64 hop's bytes); 0 = one
 * system per thread.  Results are bit-identical. */
#define WD_OPT_F64_PAIRS 5
/* The WD_RHS_LOOP sloppy kernel's neighbour prefetch: 1 = the loads of the
 * next hop direction are issued before each direction's arithmetic, at 3
 * wavefronts per SIMD; 3 = the third next direction's, at 2 wavefronts per
 * SIMD; 0 (default) = 1 for the hop and M launches, 3 for the launches fused
 * with dot products (which would spill registers at 3 wavefronts per SIMD).
 * With F32 links the kernel always loads one direction ahead (2 wavefronts
 * per SIMD).  Results are bit-identical. */
#define WD_OPT_LOOP_PREFETCH 6
/* Systems per workgroup of the WD_RHS_LOOP sloppy kernel: the systems run in
 * groups, each group a separate run of workgroups over all sites (group 0 for
 * every site block, then group 1, ...), so neighbouring workgroups reach the
 * same system at nearly the same time and share its spinors through L2, at
 * the cost of loading the links once per group.  Groups are formed from the
 * systems a launch applies (a solver's masked systems are left out).  0
 * (default) = automatic: balanced groups of at most 12 systems; S > 0 =
 * balanced groups of at most S (S >= the system count: one group).  Results
 * are bit-identical. */
#define WD_OPT_LOOP_SYSTEMS 7
int wd_set_option(wd_ctx *ctx, int option, int value);
/* The value an option has in effect on this context: 0/1 for
 * WD_OPT_XCD_REMAP, WD_OPT_PARITY_ORDER and WD_OPT_F64_PAIRS, the column width W (0 = natural
 * order) for WD_OPT_BLOCK_TILE, the t-run length T (0 = site-block kernel)
 * for WD_OPT_TSTREAM. */
int wd_get_option(wd_ctx *ctx, int option, int *value);

#ifdef __cplusplus
}
#endif
#endif
