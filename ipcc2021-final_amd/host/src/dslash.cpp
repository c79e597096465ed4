// dslash.cpp -- the reference's operator entry points (include/dslash.h,
// include/dslash_new.h, include/invert.h: Dslash / DslashNew) on the GPU.
//
// Each call uploads the caller's host arrays, runs one stream-ordered kernel
// sequence in libwilson.so and downloads dest, so the blocking, fully
// overwriting semantics of the reference hold.  The links are taken from a
// live gauge_resident of U when there is one (no upload), and the device
// spinor fields are kept per context between calls.  The scalar family (Dslash,
// Dslashoffd, DslashEE/OO) and the AVX2 family (DslashNew, DslashoffdNew,
// DslashEEOONew) compute the same operator and map to the same kernels.
#include <cstring>
#include <stdexcept>

#include "dslash.h"
#include "dslash_new.h"
#include "invert.h"
#include "wd_runtime.h"

namespace {

void hop(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, bool dag, int cb)
{
    if (src.A == dest.A)
        throw std::runtime_error("Dslashoffd: src and dest must not alias");
    wd_ctx *ctx = wdrt::context(src.subgs, src.site_vec);
    wdrt::GaugeRef g(ctx, U);
    wd_field *in = wdrt::scratch_field(ctx, 0), *out = wdrt::scratch_field(ctx, 1);
    wdrt::upload(ctx, in, src);
    wdrt::check(ctx, wd_hop(ctx, out, in, g.g, dag ? 1 : 0, cb), "wd_hop");
    wdrt::download(ctx, out, dest);
}

void full(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, double mass, bool dag)
{
    if (src.A == dest.A)
        throw std::runtime_error("Dslash: src and dest must not alias");
    wd_ctx *ctx = wdrt::context(src.subgs, src.site_vec);
    wdrt::GaugeRef g(ctx, U);
    wd_field *in = wdrt::scratch_field(ctx, 0), *out = wdrt::scratch_field(ctx, 1);
    wdrt::upload(ctx, in, src);
    wdrt::check(ctx, wd_apply_M(ctx, out, in, g.g, mass, dag ? 1 : 0), "wd_apply_M");
    wdrt::download(ctx, out, dest);
}

// dest = (4+m) src over the full lattice (half = -1) or one half (0 even, 1 odd)
void diag(lattice_fermion &src, lattice_fermion &dest, double mass, int half)
{
    wd_ctx *ctx = wdrt::context(src.subgs, src.site_vec);
    wd_field *in = wdrt::scratch_field(ctx, 0), *out = wdrt::scratch_field(ctx, 1);
    wdrt::upload(ctx, in, src);
    wdrt::check(ctx, wd_diag(ctx, out, in, mass), "wd_diag");
    wdrt::download(ctx, out, dest);
    if (half >= 0) {
        const int nh = dest.size / 2;
        std::memset(static_cast<void *>(dest.A + (1 - half) * nh), 0,
                    sizeof(std::complex<double>) * nh);
    }
}

} // namespace

// src/dslash.cpp:11-22 / :24-35
void DslashEE(lattice_fermion &src, lattice_fermion &dest, const double mass)
{
    diag(src, dest, mass, 0);
}
void DslashOO(lattice_fermion &src, lattice_fermion &dest, const double mass)
{
    diag(src, dest, mass, 1);
}

// src/dslash.cpp:38-1249 (cb = 0: even output from odd input)
void Dslashoffd(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag,
                int cb)
{
    hop(src, dest, U, dag, cb);
}
void DslashEO(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag)
{
    hop(src, dest, U, dag, 0);
}
void DslashOE(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag)
{
    hop(src, dest, U, dag, 1);
}

// src/dslash_new.cpp:1301-1338
void DslashEEOONew(lattice_fermion &src, lattice_fermion &dest, const double mass)
{
    diag(src, dest, mass, -1);
}

// src/dslash_new.cpp:1341-2677
void DslashoffdNew(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag,
                   int cb)
{
    hop(src, dest, U, dag, cb);
}
void DslashEONew(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag)
{
    hop(src, dest, U, dag, 0);
}
void DslashOENew(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const bool dag)
{
    hop(src, dest, U, dag, 1);
}

// src/invert.cpp:123-136 and :138-156
void Dslash(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const double mass,
            const bool dagger)
{
    full(src, dest, U, mass, dagger);
}
void DslashNew(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const double mass,
               const bool dagger)
{
    full(src, dest, U, mass, dagger);
}
