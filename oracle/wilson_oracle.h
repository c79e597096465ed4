/*
 * wilson_oracle.h -- CPU restatement of the reference Wilson-Dslash hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: tests/, the smoke entry
 * point and bench.py's cpu_baseline leg may call it; the product library never
 * links, loads or calls it.
 *
 * Layout is the reference's (include/operator.h:56-75, src/dslash.cpp:544-551):
 *   spinor  complex<double>[V*12] at eo(x)*12 + spin*3 + colour
 *   links   complex<double>[V*9] per direction mu at eo(x)*9 + row*3 + col
 *   eo(x) = lex(x)/2 + ((x0+x1+x2+x3) mod 2) * V/2,
 *   lex(x) = x0 + L0*(x1 + L1*(x2 + L2*x3)).
 * Complex numbers are stored as interleaved (re, im) doubles.
 *
 * Parity pins (how each function is checked, see tests/test_oracle_*.py):
 *   or_hop / or_dslash   golden vectors written by the compiled reference,
 *                        dense gamma-matrix operator built from the problem
 *                        statement (PDF eq. 2.4-2.7), plane-wave closed form,
 *                        gamma5-hermiticity, gauge covariance, zero links.
 *   or_diag              zero-link case and golden vectors.
 *   or_norm2 / or_dot    closed forms on structured vectors, golden values.
 *   or_cg                golden iteration counts and residual histories written
 *                        by the compiled reference, dense solve on tiny lattices.
 */
#ifndef WILSON_ORACLE_H
#define WILSON_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Hopping term for output parity cb from input parity 1-cb; the other half of
 * dest is zeroed.  dag != 0 applies the M^dagger hop.  Periodic in all four
 * directions.  Follows DslashoffdNew / Dslashoffd single-rank semantics
 * (src/dslash_new.cpp:1341-2677, src/dslash.cpp:38-1249). */
void or_hop(const double *src, double *dest, const double *const U[4], const int L[4], int dag,
            int cb);

/* dest = (4+mass) * src over the full lattice (src/dslash_new.cpp:1301-1338). */
void or_diag(const double *src, double *dest, const int L[4], double mass);

/* dest = M src (dag == 0) or M^dagger src (dag != 0), full lattice
 * (src/invert.cpp:138-156 DslashNew, :123-136 Dslash). */
void or_dslash(const double *src, double *dest, const double *const U[4], const int L[4],
               double mass, int dag);

/* sum |v_i|^2 over n complex entries (include/operator_mpi.h:15-46). */
double or_norm2(const double *v, long n);

/* out = sum conj(a_i) * b_i over n complex entries (include/operator_mpi.h:48-79). */
void or_dot(const double *a, const double *b, long n, double out[2]);

/* CGNR on M^dagger M x = M^dagger b (src/invert.cpp:25-121).  x holds the
 * starting vector x0 on entry (the reference draws it from rand()) and the
 * solution on return.  The loop runs f = 1 .. max-1 and stops when
 * sqrt(|r|^2) < accuracy (absolute).  rsd_hist (may be NULL, length >= max)
 * receives |r| after each iteration.  Returns the iteration count f at which
 * the stop test passed, or -(max-1) when it never passed. */
int or_cg(const double *b, double *x, const double *const U[4], const int L[4], double mass,
          int max, double accuracy, double *rsd_hist);

/* The reference's CG start vector (src/invert.cpp:42-44): srand(seed), then
 * x_i = rand()/RAND_MAX + 0i, i = 0 .. n-1 (n complex entries). */
void or_rand_x0(double *x, long n, unsigned seed);

/* Site index helpers (include/operator.h:56-75). */
long or_site2index(const int x[4], const int L[4]);
long or_site2eoindex(const int x[4], const int L[4]);

#ifdef __cplusplus
}
#endif
#endif
