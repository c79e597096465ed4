"""FP64 operators and solvers on 12-real FP64 links (rows 0 and 1 stored, row
2 rebuilt as conj(row0 x row1) in registers): the opt-in format for SU(3)
links.  Against the oracle on the full 18-real links: 1e-13 relative L2 for
the hop and M (the rebuilt row differs from the stored one by the links'
unitarity rounding, about 1e-16), CG iteration counts equal to the 18-real
run's, mixed solvers whose reliable updates run on the 12-real links meeting
the oracle's FP64 residual, and the t-streaming kernel forced at small
shapes."""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-13
SHAPES = [[4, 4, 4, 8], [6, 4, 2, 10], [8, 8, 8, 16]]


@pytest.mark.parametrize("L", SHAPES, ids=lambda L: "x".join(map(str, L)))
@pytest.mark.parametrize("dag", [0, 1])
def test_recon12_operators_match_oracle(w, oracle, L, dag):
    A = w.gauge_random_su3(L, 20211025)
    b = w.spinor_gaussian(L, 20211026)
    mass = kappa_to_mass(0.13)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U12 = ctx.gauge(A, prec=w.F64, recon=12)
        fin, fout = ctx.field(data=b), ctx.field()
        hops = []
        for cb in (0, 1):
            ctx.hop(fout, fin, U12, bool(dag), cb)
            hops.append(fout.download())
        Ms = []
        for ts in (0, 2):   # site-block kernel, then t-streaming with runs of 2
            ctx.set_option(w.OPT_TSTREAM, ts)
            ctx.apply_M(fout, fin, U12, mass, bool(dag))
            Ms.append(fout.download())
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for cb in (0, 1):
        e = rel_l2(hops[cb], oracle.hop(b, Af, L, dag, cb))
        print(f"L={L} dag={dag} hop cb={cb}: {e:.2e}", flush=True)
        assert e < TOL
    ref = oracle.dslash(b, Af, L, mass, dag)
    for M in Ms:
        assert rel_l2(M, ref) < TOL


def test_recon12_cg_same_iterations(w):
    """FP64 CG and Schur CG: the same iteration counts on 12-real and 18-real
    links, solutions within the conditioning of the rounding difference."""
    L = [8, 8, 8, 8]
    A = w.gauge_random_su3(L, 7)
    b = w.spinor_gaussian(L, 8)
    mass = kappa_to_mass(0.12)
    out = {}
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        fb = ctx.field(data=b)
        for recon in (18, 12):
            U = ctx.gauge(A, prec=w.F64, recon=recon)
            for eo in (False, True):
                x = ctx.field()
                res = ctx.cg(x, fb, U, mass, max_iter=1000, tol=1e-10, tol_relative=True,
                             use_x0=False, even_odd=eo)
                out[(recon, eo)] = (res["iterations"], x.download(), res)
    for eo in (False, True):
        it18, x18, r18 = out[(18, eo)]
        it12, x12, r12 = out[(12, eo)]
        print(f"eo={eo}: iterations 18-real {it18}, 12-real {it12}", flush=True)
        assert r12["converged"] and abs(it12 - it18) <= 1
        assert rel_l2(x12, x18) < 1e-8


@pytest.mark.parametrize("solver", ["cg_mixed", "bicgstab_mixed"])
def test_recon12_reliable_updates(w, oracle, solver):
    """Mixed solvers whose FP64 reliable updates and full-system checks run
    on 12-real links reach the oracle's FP64 residual on the 18-real links."""
    L = [6, 4, 6, 10]
    n = 4
    A = w.gauge_near_identity(L, 11, 0.3)
    bs = [w.spinor_gaussian(L, 12, r) for r in range(n)]
    masses = [kappa_to_mass(k) for k in (0.11, 0.12, 0.125, 0.13)]
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U12 = ctx.gauge(A, prec=w.F64, recon=12)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=n, data=bs)
        x = ctx.field(nrhs=n)
        kw = {"even_odd": True} if solver == "cg_mixed" else {}
        res = getattr(ctx, solver)(x, fb, U12, Us, masses, tol=1e-10, **kw)
        xs = [x.download(r) for r in range(n)]
    assert res["status"] == 0
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in range(n):
        Mdb = oracle.dslash(bs[r], Af, L, masses[r], 1)
        rr = Mdb - oracle.dslash(oracle.dslash(xs[r], Af, L, masses[r], 0), Af, L, masses[r], 1)
        assert np.linalg.norm(rr) / np.linalg.norm(Mdb) < 1.05e-10
