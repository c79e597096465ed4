// ref_harness.cpp -- extern "C" driver around the compiled reference sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference's own src/*.cpp (compiled where they lie, never copied) into
// oracle/_ref/libref.so.  It lets tests/golden/make_golden.py and the CPU
// baseline call the reference's DslashNew / Dslash / DslashoffdNew /
// Dslashoffd / DslashEEOONew / CGinvert / norm_2 / vector_p / check on one MPI
// rank (singleton MPI_Init), with the reference's 32-byte alignment contract
// honoured by copying through _mm_malloc'd buffers.
#include <mpi.h>
#include <immintrin.h>

#include <complex>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "check.h"
#include "dslash.h"
#include "dslash_new.h"
#include "invert.h"
#include "lattice_fermion.h"
#include "lattice_gauge.h"
#include "load_gauge.h"
#include "operator_mpi.h"

typedef std::complex<double> cd;

namespace {

int ensure_mpi()
{
    int inited = 0;
    MPI_Initialized(&inited);
    if (!inited) {
        int argc = 0;
        char **argv = nullptr;
        MPI_Init(&argc, &argv);
    }
    int size = 0;
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    return size;
}

struct Geometry {
    int subgs[4];
    int site_vec[4];
    long V;
    explicit Geometry(const int L[4])
    {
        V = 1;
        for (int i = 0; i < 4; ++i) {
            subgs[i] = L[i];
            site_vec[i] = L[i];
            V *= L[i];
        }
    }
};

cd *aligned_copy(const double *src, long ncomplex)
{
    cd *p = (cd *)_mm_malloc(sizeof(cd) * ncomplex, 64);
    if (src)
        std::memcpy(p, src, sizeof(cd) * ncomplex);
    else
        std::memset(p, 0, sizeof(cd) * ncomplex);
    return p;
}

struct GaugeCopy {
    cd *A[4];
    GaugeCopy(const double *const U[4], long V)
    {
        for (int mu = 0; mu < 4; ++mu)
            A[mu] = aligned_copy(U[mu], 9 * V);
    }
    ~GaugeCopy()
    {
        for (int mu = 0; mu < 4; ++mu)
            _mm_free(A[mu]);
    }
};

// Runs fn with std::cout redirected; the captured text is written to `log`.
template <class F>
void captured(char *log, long log_len, F fn)
{
    std::ostringstream os;
    std::streambuf *old = std::cout.rdbuf(os.rdbuf());
    fn();
    std::cout.rdbuf(old);
    if (log && log_len > 0) {
        std::string s = os.str();
        long n = (long)s.size() < log_len - 1 ? (long)s.size() : log_len - 1;
        std::memcpy(log, s.data(), n);
        log[n] = '\0';
    }
}

} // namespace

extern "C" {

int ref_mpi_size() { return ensure_mpi(); }

// which: 0 DslashNew (AVX2, src/invert.cpp:138), 1 Dslash (scalar, src/invert.cpp:123)
void ref_dslash(int which, const double *src, double *dest, const double *const U[4],
                const int L[4], double mass, int dag)
{
    ensure_mpi();
    Geometry g(L);
    GaugeCopy gc(U, g.V);
    cd *s = aligned_copy(src, 12 * g.V), *d = aligned_copy(nullptr, 12 * g.V);
    {
        lattice_fermion fs(s, g.subgs, g.site_vec), fd(d, g.subgs, g.site_vec);
        lattice_gauge gu(gc.A, g.subgs, g.site_vec);
        if (which == 0)
            DslashNew(fs, fd, gu, mass, dag != 0);
        else
            Dslash(fs, fd, gu, mass, dag != 0);
    }
    std::memcpy(dest, d, sizeof(cd) * 12 * g.V);
    _mm_free(s);
    _mm_free(d);
}

// which: 0 DslashoffdNew (src/dslash_new.cpp:1341), 1 Dslashoffd (src/dslash.cpp:38)
void ref_hop(int which, const double *src, double *dest, const double *const U[4], const int L[4],
             int dag, int cb)
{
    ensure_mpi();
    Geometry g(L);
    GaugeCopy gc(U, g.V);
    cd *s = aligned_copy(src, 12 * g.V), *d = aligned_copy(nullptr, 12 * g.V);
    {
        lattice_fermion fs(s, g.subgs, g.site_vec), fd(d, g.subgs, g.site_vec);
        lattice_gauge gu(gc.A, g.subgs, g.site_vec);
        if (which == 0)
            DslashoffdNew(fs, fd, gu, dag != 0, cb);
        else
            Dslashoffd(fs, fd, gu, dag != 0, cb);
    }
    std::memcpy(dest, d, sizeof(cd) * 12 * g.V);
    _mm_free(s);
    _mm_free(d);
}

// DslashEEOONew (src/dslash_new.cpp:1301)
void ref_diag(const double *src, double *dest, const int L[4], double mass)
{
    ensure_mpi();
    Geometry g(L);
    cd *s = aligned_copy(src, 12 * g.V), *d = aligned_copy(nullptr, 12 * g.V);
    {
        lattice_fermion fs(s, g.subgs, g.site_vec), fd(d, g.subgs, g.site_vec);
        DslashEEOONew(fs, fd, mass);
    }
    std::memcpy(dest, d, sizeof(cd) * 12 * g.V);
    _mm_free(s);
    _mm_free(d);
}

// norm_2 and vector_p (include/operator_mpi.h:15-79)
double ref_norm2(const double *v, const int L[4])
{
    ensure_mpi();
    Geometry g(L);
    cd *s = aligned_copy(v, 12 * g.V);
    double r;
    {
        lattice_fermion f(s, g.subgs, g.site_vec);
        r = norm_2(f);
    }
    _mm_free(s);
    return r;
}

void ref_dot(const double *a, const double *b, const int L[4], double out[2])
{
    ensure_mpi();
    Geometry g(L);
    cd *pa = aligned_copy(a, 12 * g.V), *pb = aligned_copy(b, 12 * g.V);
    {
        lattice_fermion fa(pa, g.subgs, g.site_vec), fb(pb, g.subgs, g.site_vec);
        cd r = vector_p(fa, fb);
        out[0] = r.real();
        out[1] = r.imag();
    }
    _mm_free(pa);
    _mm_free(pb);
}

// CGinvert (src/invert.cpp:25-121) after srand(seed): the reference draws its
// starting vector from rand().  The per-iteration printout is captured into log.
void ref_cg(const double *src, double *dest, const double *const U[4], const int L[4],
            double mass, int max, double accuracy, unsigned seed, char *log, long log_len)
{
    ensure_mpi();
    Geometry g(L);
    GaugeCopy gc(U, g.V);
    cd *s = aligned_copy(src, 12 * g.V), *d = aligned_copy(nullptr, 12 * g.V);
    {
        lattice_fermion fs(s, g.subgs, g.site_vec), fd(d, g.subgs, g.site_vec);
        lattice_gauge gu(gc.A, g.subgs, g.site_vec);
        srand(seed);
        captured(log, log_len, [&] { CGinvert(fs, fd, gu, mass, max, accuracy); });
    }
    std::memcpy(dest, d, sizeof(cd) * 12 * g.V);
    _mm_free(s);
    _mm_free(d);
}

// check (src/check.cpp:28-110); the report is captured into log.
void ref_check(const double *src, const double *x, const double *const U[4], const int L[4],
               double mass, char *log, long log_len)
{
    ensure_mpi();
    Geometry g(L);
    GaugeCopy gc(U, g.V);
    cd *s = aligned_copy(src, 12 * g.V), *px = aligned_copy(x, 12 * g.V);
    {
        lattice_fermion fs(s, g.subgs, g.site_vec), fx(px, g.subgs, g.site_vec);
        lattice_gauge gu(gc.A, g.subgs, g.site_vec);
        captured(log, log_len, [&] { check(fs, fx, gu, mass); });
    }
    _mm_free(s);
    _mm_free(px);
}

// The steps of the reference's driver (src/main.cpp:77-117) on one rank:
// srand(seed); source from rand(); load_gauge(gaugefile) ("UNIT", "RANDOM" or a
// big-endian file); CGinvert(MaxCG = 500, Accuracy = 1e-9); check().  All the
// text the reference prints is captured into log.
void ref_main_flow(const int L[4], double mass, const char *gaugefile, unsigned seed, char *log,
                   long log_len)
{
    ensure_mpi();
    Geometry g(L);
    captured(log, log_len, [&] {
        srand(seed);
        std::vector<double> source(g.V * 24);
        for (auto &v : source)
            v = (double)rand() / RAND_MAX;
        lattice_fermion src(g.subgs, g.site_vec);
        for (long i = 0; i < g.V * 12; i++)
            src.A[i] = cd(source[2 * i], source[2 * i + 1]);
        lattice_gauge U(g.subgs, g.site_vec);
        std::string fn(gaugefile);
        load_gauge(U.A, fn, g.subgs, g.site_vec);
        lattice_fermion dest(g.subgs, g.site_vec);
        CGinvert(src, dest, U, mass, 500, 1.0e-9);
        check(src, dest, U, mass);
    });
}

} // extern "C"
