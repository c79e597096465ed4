// invert.cpp -- CGinvert (src/invert.cpp:15-121) on the GPU.
//
// The start vector is drawn from the C library's rand() exactly like the
// reference (x0_i = rand()/RAND_MAX + 0i, src/invert.cpp:42-44), so a caller
// that seeds with srand() as src/main.cpp:77-78 does gets the same x0, and
// the device CG (wd_cg) then runs the reference's recurrences with its
// absolute stop test |r| < accuracy over f = 1 .. max-1.
#include <cstdlib>
#include <iostream>
#include <vector>

#include "invert.h"
#include "wd_runtime.h"

using std::complex;

int CGinvert(complex<double> *src_p, complex<double> *dest_p, complex<double> *gauge[4],
             const double mass, const int max, const double accuracy, int *subgs, int *site_vec)
{
    lattice_gauge U(gauge, subgs, site_vec);
    lattice_fermion src(src_p, subgs, site_vec);
    lattice_fermion dest(dest_p, subgs, site_vec);
    CGinvert(src, dest, U, mass, max, accuracy);
    return 0;
}

int CGinvert(lattice_fermion &src, lattice_fermion &dest, lattice_gauge &U, const double mass,
             const int max, const double accuracy)
{
    // the start vector comes first: nothing may draw from rand() in between
    // (creating the GPU context on first use is not guaranteed not to)
    for (int i = 0; i < dest.size; i++)
        dest.A[i] = 1.0 * rand() / RAND_MAX;
    wd_ctx *ctx = wdrt::context(src.subgs, src.site_vec);
    wdrt::GaugeRef g(ctx, U);
    wdrt::Field b(ctx, src), x(ctx, dest);
    wd_cg_params p;
    wd_cg_params_default(&p);
    p.max_iter = max;
    p.tol = accuracy;
    p.tol_relative = 0;
    p.use_x0 = 1;
    std::vector<double> hist(max > 0 ? max : 1);
    p.rsd_history = hist.data();
    wd_cg_report rep{};
    const int rc = wd_cg(ctx, x.f, b.f, g.g, mass, &p, &rep);
    wdrt::check(ctx, rc, "wd_cg");
    x.download(dest);
#ifndef VERBOSE_SIMPLE
    const int printed = rep.converged ? rep.iterations - 1 : rep.iterations;
    for (int f = 1; f <= printed; ++f)
        std::cout << "CG: " << f << " iter, rsd |r| = " << hist[f - 1] << std::endl;
#endif
    if (rep.converged)
        std::cout << "CG: " << rep.iterations << " iterations, convergence residual |r| = "
                  << rep.rsd << std::endl;
    return 0;
}
