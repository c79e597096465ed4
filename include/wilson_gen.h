/*
 * wilson_gen.h -- deterministic synthetic inputs (part of libwilson.so, host code).
 *
 * No counterpart in the reference, which reads gauge files or sets unit links
 * (src/load_gauge.cpp:29-181) and draws its source from rand()
 * (src/main.cpp:77-94).  These generators produce the synthetic workloads the
 * benchmark and tests run on, reproducibly for any volume: every value is a
 * function of (seed, stream, global lexicographic site, component) through a
 * counter-based Philox4x32-10 generator (Salmon, Moraes, Dror, Shaw, "Parallel
 * random numbers: as easy as 1, 2, 3", SC'11) and Box-Muller Gaussians.
 * Output arrays use the reference layout (see include/wilson.h).
 */
#ifndef WILSON_GEN_H
#define WILSON_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Random SU(3) links: Gaussian complex 3x3, Gram-Schmidt rows 0 and 1,
 * row 2 = conj(row0 x row1).  A[mu] = complex<double>[V*9]. */
void wdg_gauge_random_su3(const int L[4], uint64_t seed, double *const A[4]);

/* Near-identity SU(3) links U = exp(i * eps * H), H traceless Hermitian built
 * from Gaussian entries, reunitarised as above. */
void wdg_gauge_near_identity(const int L[4], uint64_t seed, double eps, double *const A[4]);

/* Unit links (the reference's "UNIT" gauge, src/load_gauge.cpp:40-52). */
void wdg_gauge_unit(const int L[4], double *const A[4]);

/* Multiplies the mu-links on the last slice x_mu = L_mu - 1 by -1: the
 * periodic operator on the result equals the antiperiodic-in-mu operator on
 * the input. */
void wdg_gauge_flip_boundary(const int L[4], int mu, double *const A[4]);

/* Gaussian spinor, every real component N(0,1); `stream` selects independent
 * fields for one seed.  A = complex<double>[V*12]. */
void wdg_spinor_gaussian(const int L[4], uint64_t seed, uint32_t stream, double *A);

/* Point source: 1 at (site x[4], spin, colour), 0 elsewhere. */
void wdg_spinor_point(const int L[4], const int x[4], int spin, int colour, double *A);

/* Philox4x32-10 block for counter c and key k (exposed for tests). */
void wdg_philox4x32(const uint32_t c[4], const uint32_t k[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
