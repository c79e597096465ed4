"""Every solver on lattices of assorted shapes (odd x/2, unequal extents,
half volumes that are not a multiple of the 256-site workgroup, t extents
with and without t-streaming runs): FP64 CG, Schur FP64 CG, mixed CG on the
full and on the Schur system, three systems at three kappas in one mixed
solve.  Each solution must meet the reference's stop test (checked by the
library with the FP64 operator), agree with the FP64 CG solution, and the
mixed solvers must take the FP64 iteration counts within a few."""
import numpy as np
import pytest

from oracle_py import kappa_to_mass

pytestmark = pytest.mark.gpu

SHAPES = [[12, 12, 12, 12], [6, 6, 6, 12], [10, 12, 14, 16], [20, 20, 20, 40]]


@pytest.mark.parametrize("L", SHAPES)
def test_solvers_agree_on_shape(w, L):
    A = w.gauge_near_identity(L, 11, 0.3)
    n = 3
    bs = [w.spinor_gaussian(L, 12, r) for r in range(n)]
    masses = [kappa_to_mass(k) for k in (0.12, 0.125, 0.13)]
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        fb1 = ctx.field(data=bs[2])
        x = ctx.field()
        fp = ctx.cg(x, fb1, U, masses[2], max_iter=5000, tol=1e-10, tol_relative=True,
                    use_x0=False)
        ref = x.download()
        assert fp["converged"] and fp["true_rsd"] <= 1.0001e-10 * fp["rhs_norm"]
        xe = ctx.field()
        eo = ctx.cg(xe, fb1, U, masses[2], max_iter=5000, tol=1e-10, tol_relative=True,
                    use_x0=False, even_odd=True)
        assert eo["converged"] and eo["true_rsd"] <= 1.0001e-10 * eo["rhs_norm"]
        sols = {"eo": xe.download()}
        its = {}
        for even_odd, gp in ((False, w.F32), (True, w.F16)):
            Us = ctx.gauge(A, prec=gp, recon=12)
            fb = ctx.field(nrhs=n, data=bs)
            X = ctx.field(nrhs=n)
            mx = ctx.cg_mixed(X, fb, U, Us, masses, tol=1e-10, spinor_prec=w.F16,
                              even_odd=even_odd)
            assert mx["status"] == 0
            for s in mx["systems"]:
                assert s["converged"] and s["true_rel_rsd"] <= 1e-10
            sols[f"mixed{int(even_odd)}"] = X.download(2)
            its[even_odd] = mx["systems"][2]["iterations"]
    print(L, "fp64", fp["iterations"], "eo", eo["iterations"], "mixed", its, flush=True)
    for k, v in sols.items():
        d = np.linalg.norm(v - ref) / np.linalg.norm(ref)
        assert d < 1e-7, (k, d)
    assert abs(its[False] - fp["iterations"]) <= 3
    assert abs(its[True] - eo["iterations"]) <= 6
