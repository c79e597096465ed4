/*
 * wilson_oracle.c -- CPU restatement of the reference Wilson-Dslash hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, the smoke entry point and
 * bench.py's cpu_baseline leg.  The product library never links or calls it.
 *
 * Plain C99, strict IEEE double (no -ffast-math).  Loops over output sites are
 * parallelised with OpenMP when the compiler provides it; every output element
 * is computed by exactly one thread in a fixed order, so results do not depend
 * on the thread count.  Reductions are sequential.
 */
#include "wilson_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- complex helpers on interleaved doubles ------------------------------ */
typedef struct {
    double re, im;
} cplx;

static inline cplx cx(double re, double im)
{
    cplx r = {re, im};
    return r;
}
static inline cplx cadd(cplx a, cplx b) { return cx(a.re + b.re, a.im + b.im); }
static inline cplx cmul(cplx a, cplx b)
{
    return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx cconj(cplx a) { return cx(a.re, -a.im); }
static inline cplx cld(const double *p, long k) { return cx(p[2 * k], p[2 * k + 1]); }

/* ---- index maps (include/operator.h:56-75) --------------------------------- */
long or_site2index(const int x[4], const int L[4])
{
    return (long)(x[0] % L[0]) +
           (long)L[0] * ((x[1] % L[1]) + (long)L[1] * ((x[2] % L[2]) + (long)L[2] * (x[3] % L[3])));
}

long or_site2eoindex(const int x[4], const int L[4])
{
    long Vh = (long)L[0] * L[1] * L[2] * L[3] / 2;
    long lex = or_site2index(x, L);
    return lex / 2 + (long)((x[0] + x[1] + x[2] + x[3]) % 2) * Vh;
}

/*
 * One directional contribution of the hopping term.
 *
 * With the half-spinor pair
 *     h0 = psi[0] + p0 * psi[s0],   h1 = psi[1] + p1 * psi[s1]
 * and r_k = -1/2 * W h_k (W = U, or U^dagger when `dag_link`), the
 * contribution is
 *     d[0] += r0, d[1] += r1, d[q0] += c0 * r0, d[q1] += c1 * r1.
 * The coefficients for each direction are read off the reference's scalar
 * loops, cited at the call sites below; they are the rows of (1 -/+ f*gamma_mu)
 * in the DeGrand-Rossi basis.
 */
static void hop_dir(double *d, const double *W, int dag_link, const double *psi, int s0, cplx p0,
                    int s1, cplx p1, int q0, cplx c0, int q1, cplx c1)
{
    cplx h0[3], h1[3];
    for (int c = 0; c < 3; ++c) {
        h0[c] = cadd(cld(psi, 0 * 3 + c), cmul(p0, cld(psi, s0 * 3 + c)));
        h1[c] = cadd(cld(psi, 1 * 3 + c), cmul(p1, cld(psi, s1 * 3 + c)));
    }
    for (int a = 0; a < 3; ++a) {
        cplx r0 = cx(0.0, 0.0), r1 = cx(0.0, 0.0);
        for (int b = 0; b < 3; ++b) {
            /* U.v uses A[a*3+b]; U^dagger.v uses conj(A[b*3+a]) (src/dslash.cpp:545, :596) */
            cplx w = dag_link ? cconj(cld(W, b * 3 + a)) : cld(W, a * 3 + b);
            r0 = cadd(r0, cmul(w, h0[b]));
            r1 = cadd(r1, cmul(w, h1[b]));
        }
        r0 = cx(-0.5 * r0.re, -0.5 * r0.im);
        r1 = cx(-0.5 * r1.re, -0.5 * r1.im);
        cplx t0 = cmul(c0, r0), t1 = cmul(c1, r1);
        d[2 * (0 * 3 + a)] += r0.re;
        d[2 * (0 * 3 + a) + 1] += r0.im;
        d[2 * (1 * 3 + a)] += r1.re;
        d[2 * (1 * 3 + a) + 1] += r1.im;
        d[2 * (q0 * 3 + a)] += t0.re;
        d[2 * (q0 * 3 + a) + 1] += t0.im;
        d[2 * (q1 * 3 + a)] += t1.re;
        d[2 * (q1 * 3 + a) + 1] += t1.im;
    }
}

/* Hopping term, output parity cb (src/dslash.cpp:38-1249 single-rank branches;
 * src/dslash_new.cpp:1341-2677 computes the same sums with AVX2). */
void or_hop(const double *src, double *dest, const double *const U[4], const int L[4], int dag,
            int cb)
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    const long Vh = V / 2;
    const double f = dag ? -1.0 : 1.0; /* src/dslash.cpp:95 */
    const cplx I = cx(0.0, 1.0);
    const cplx fI = cx(0.0, f), mfI = cx(0.0, -f), F = cx(f, 0.0), mF = cx(-f, 0.0);

    /* the other half is zeroed (src/dslash_new.cpp:1398 memsets the whole field) */
    memset(dest + 24 * (long)(1 - cb) * Vh, 0, sizeof(double) * 24 * Vh);
    (void)I;

#pragma omp parallel for schedule(static)
    for (long lex = 0; lex < V; ++lex) {
        int x[4];
        x[0] = (int)(lex % L[0]);
        x[1] = (int)((lex / L[0]) % L[1]);
        x[2] = (int)((lex / ((long)L[0] * L[1])) % L[2]);
        x[3] = (int)(lex / ((long)L[0] * L[1] * L[2]));
        if (((x[0] + x[1] + x[2] + x[3]) & 1) != cb)
            continue;
        const long e = or_site2eoindex(x, L);
        double *d = dest + 24 * e;
        for (int k = 0; k < 24; ++k)
            d[k] = 0.0;
        for (int mu = 0; mu < 4; ++mu) {
            int xf[4] = {x[0], x[1], x[2], x[3]}, xb[4] = {x[0], x[1], x[2], x[3]};
            xf[mu] = (x[mu] + 1) % L[mu];
            xb[mu] = (x[mu] - 1 + L[mu]) % L[mu];
            const long ef = or_site2eoindex(xf, L), eb = or_site2eoindex(xb, L);
            const double *Uf = U[mu] + 18 * e;  /* U_mu(x) */
            const double *Ub = U[mu] + 18 * eb; /* U_mu(x - mu), daggered */
            const double *pf = src + 24 * ef, *pb = src + 24 * eb;
            switch (mu) {
            case 0:
                /* x fwd: (psi0 - f i psi3), (psi1 - f i psi2); d3 += f i r0, d2 += f i r1
                 * (src/dslash.cpp:541-553) */
                hop_dir(d, Uf, 0, pf, 3, mfI, 2, mfI, 3, fI, 2, fI);
                /* x bwd: (psi0 + f i psi3), (psi1 + f i psi2); d3 -= f i r0, d2 -= f i r1
                 * (src/dslash.cpp:588-600) */
                hop_dir(d, Ub, 1, pb, 3, fI, 2, fI, 3, mfI, 2, mfI);
                break;
            case 1:
                /* y fwd: (psi0 + f psi3), (psi1 - f psi2); d3 += f r0, d2 -= f r1
                 * (src/dslash.cpp:632-642) */
                hop_dir(d, Uf, 0, pf, 3, F, 2, mF, 3, F, 2, mF);
                /* y bwd: (psi0 - f psi3), (psi1 + f psi2); d3 -= f r0, d2 += f r1
                 * (src/dslash.cpp:678-688) */
                hop_dir(d, Ub, 1, pb, 3, mF, 2, F, 3, mF, 2, F);
                break;
            case 2:
                /* z fwd: (psi0 - f i psi2), (psi1 + f i psi3); d2 += f i r0, d3 -= f i r1
                 * (src/dslash.cpp:725-735) */
                hop_dir(d, Uf, 0, pf, 2, mfI, 3, fI, 2, fI, 3, mfI);
                /* z bwd: (psi0 + f i psi2), (psi1 - f i psi3); d2 -= f i r0, d3 += f i r1
                 * (src/dslash.cpp:770-780) */
                hop_dir(d, Ub, 1, pb, 2, fI, 3, mfI, 2, mfI, 3, fI);
                break;
            default:
                /* t fwd: (psi0 - f psi2), (psi1 - f psi3); d2 -= f r0, d3 -= f r1
                 * (src/dslash.cpp:815-825) */
                hop_dir(d, Uf, 0, pf, 2, mF, 3, mF, 2, mF, 3, mF);
                /* t bwd: (psi0 + f psi2), (psi1 + f psi3); d2 += f r0, d3 += f r1
                 * (src/dslash.cpp:862-872) */
                hop_dir(d, Ub, 1, pb, 2, F, 3, F, 2, F, 3, F);
                break;
            }
        }
    }
}

/* dest = (4 + mass) src over all 12V complex entries (src/dslash_new.cpp:1301-1338) */
void or_diag(const double *src, double *dest, const int L[4], double mass)
{
    const long n = 24L * L[0] * L[1] * L[2] * L[3];
    const double a = 4.0 + mass;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        dest[i] = a * src[i];
}

/* dest = (4+m) src + hop(cb=0) + hop(cb=1) (src/invert.cpp:138-156) */
void or_dslash(const double *src, double *dest, const double *const U[4], const int L[4],
               double mass, int dag)
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    const long n = 24 * V;
    double *tmp = (double *)malloc(sizeof(double) * n);
    or_diag(src, dest, L, mass);
    for (int cb = 0; cb < 2; ++cb) {
        or_hop(src, tmp, U, L, dag, cb);
        double *d = dest + 24 * (long)cb * (V / 2);
        const double *t = tmp + 24 * (long)cb * (V / 2);
        for (long i = 0; i < 24 * (V / 2); ++i)
            d[i] += t[i];
    }
    free(tmp);
}

/* include/operator_mpi.h:15-46: sum of |v_i|^2 */
double or_norm2(const double *v, long n)
{
    double s = 0.0;
    for (long i = 0; i < 2 * n; ++i)
        s += v[i] * v[i];
    return s;
}

/* include/operator_mpi.h:48-79: sum of conj(a_i) b_i */
void or_dot(const double *a, const double *b, long n, double out[2])
{
    double re = 0.0, im = 0.0;
    for (long i = 0; i < n; ++i) {
        re += a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
        im += a[2 * i] * b[2 * i + 1] - a[2 * i + 1] * b[2 * i];
    }
    out[0] = re;
    out[1] = im;
}

/* The reference's CG start vector (src/invert.cpp:42-44): srand(seed), then
 * x_i = rand()/RAND_MAX + 0i for the n complex entries, from this C library. */
void or_rand_x0(double *x, long n, unsigned seed)
{
    srand(seed);
    for (long i = 0; i < n; ++i) {
        x[2 * i] = 1.0 * rand() / RAND_MAX;
        x[2 * i + 1] = 0.0;
    }
}

static cplx cdiv(cplx a, cplx b)
{
    const double den = b.re * b.re + b.im * b.im;
    return cx((a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den);
}

/* CGNR (src/invert.cpp:25-121) */
int or_cg(const double *b, double *x, const double *const U[4], const int L[4], double mass,
          int max, double accuracy, double *rsd_hist)
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    const long nc = 12 * V, n = 2 * nc;
    double *r = (double *)malloc(sizeof(double) * n);
    double *q = (double *)malloc(sizeof(double) * n);
    double *qq = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n);
    double *Mdb = (double *)malloc(sizeof(double) * n);
    double *tmp = (double *)malloc(sizeof(double) * n);
    double d[2];
    cplx r1 = cx(0.0, 0.0), alpha, beta;
    int result = -(max - 1);

    or_dslash(b, Mdb, U, L, mass, 1);   /* Mdb = M^dagger b       :41    */
    or_dslash(x, tmp, U, L, mass, 0);   /* tmp = M x0             :53    */
    or_dslash(tmp, r, U, L, mass, 1);   /* r = M^dagger M x0      :54    */
    for (long i = 0; i < n; ++i)        /* r = Mdb - r            :56-58 */
        r[i] = Mdb[i] - r[i];

    for (int f = 1; f < max; ++f) {
        if (f == 1) {
            memcpy(p, r, sizeof(double) * n); /* :64-66 */
        } else {
            or_dot(r, r, nc, d); /* beta = (r,r) / r1     :68 */
            beta = cdiv(cx(d[0], d[1]), r1);
            for (long i = 0; i < nc; ++i) { /* p = r + beta p :71-76 */
                cplx pi = cx(p[2 * i], p[2 * i + 1]);
                cplx t = cmul(beta, pi);
                p[2 * i] = r[2 * i] + t.re;
                p[2 * i + 1] = r[2 * i + 1] + t.im;
            }
        }
        or_dslash(p, qq, U, L, mass, 0); /* qq = M p          :79 */
        or_dslash(qq, q, U, L, mass, 1); /* q = M^dagger qq   :80 */
        or_dot(r, r, nc, d);             /* r1 = (r,r)        :82 */
        r1 = cx(d[0], d[1]);
        or_dot(p, q, nc, d);             /* alpha = r1/(p,q)  :84 */
        alpha = cdiv(r1, cx(d[0], d[1]));
        for (long i = 0; i < nc; ++i) { /* x += alpha p ; r -= alpha q   :89-105 */
            cplx t = cmul(alpha, cx(p[2 * i], p[2 * i + 1]));
            x[2 * i] += t.re;
            x[2 * i + 1] += t.im;
            t = cmul(alpha, cx(q[2 * i], q[2 * i + 1]));
            r[2 * i] -= t.re;
            r[2 * i + 1] -= t.im;
        }
        const double rsd = sqrt(or_norm2(r, nc)); /* :106-107 */
        if (rsd_hist)
            rsd_hist[f - 1] = rsd;
        if (rsd < accuracy) { /* :108-113 */
            result = f;
            break;
        }
    }
    free(r);
    free(q);
    free(qq);
    free(p);
    free(Mdb);
    free(tmp);
    return result;
}
