"""Sloppy links in temporal gauge (WD_RECON_12T) through the C-ABI.

The operator on such links is Omega M Omega^dag: checked as M'(Omega x) =
Omega M x against the FP64 oracle, with Omega computed independently in NumPy
(tests/temporal_gauge.py, pinned on the CPU by tests/test_temporal_gauge_cpu.py).
Tolerances as for the other sloppy formats: 2e-3 relative L2 with F16
spinors, 1e-6 with F32 spinors and F32 links.  The kernels with R systems per
thread and with every system of a site in one thread both take the unit t
links off the boundary slices; lattices with antiperiodic t and half volumes
that are not a multiple of the workgroup.  The mixed solvers on such links
rotate residuals and corrections at the FP64 reliable updates and must meet
the FP64 stop test 1e-10 (re-checked with the oracle) and agree with the
solutions on the plain 12-real links.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, rel_l2
import temporal_gauge as tgm

pytestmark = pytest.mark.gpu

SHAPES = [[4, 4, 6, 8], [6, 4, 6, 10]]


@pytest.mark.parametrize("L", SHAPES, ids=lambda L: "x".join(map(str, L)))
@pytest.mark.parametrize("fmt", ["F16-F16", "F16-F32", "F32-F32"])
@pytest.mark.parametrize("R", [1, 2, "loop"])
@pytest.mark.parametrize("dag", [0, 1])
def test_operator_is_rotated_M(w, oracle, L, fmt, R, dag):
    sp, gp = {"F16": w.F16, "F32": w.F32}[fmt[:3]], {"F16": w.F16, "F32": w.F32}[fmt[4:]]
    if R == "loop" and sp != w.F16:
        pytest.skip("the every-system-in-one-thread kernel takes F16 spinors")
    R = w.RHS_LOOP if R == "loop" else R
    n = 3
    A = w.gauge_random_su3(L, 20211103)
    xs = [w.spinor_gaussian(L, 20211104, r) for r in range(n)]
    masses = [kappa_to_mass(k) for k in (0.12, 0.13, 0.135)]
    Om = tgm.omega(A, L)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        Ut = ctx.gauge(A, prec=gp, recon=w.RECON_12T)
        fin = ctx.field(prec=sp, nrhs=n, data=[tgm.rotate_spinor(x, Om) for x in xs])
        fout = ctx.field(prec=sp, nrhs=n)
        ctx.apply_M_sloppy(fout, fin, Ut, masses, bool(dag), R)
        got = [fout.download(r) for r in range(n)]
    Af = w.gauge_flip_boundary(L, 3, [a.copy() for a in A])
    tol = 1e-6 if sp == w.F32 else 2e-3
    for r in range(n):
        ref = tgm.rotate_spinor(oracle.dslash(xs[r], Af, L, masses[r], dag), Om)
        e = rel_l2(got[r], ref)
        print(f"L={L} {fmt} R={R} dag={dag} system {r}: {e:.2e}", flush=True)
        assert e < tol


def test_plain_and_rotated_hops_differ_only_by_the_rotation(w, oracle):
    """One-parity launches (the Schur solvers' operator): out = Omega hop(Omega^dag in)."""
    L = [6, 4, 6, 10]
    A = w.gauge_random_su3(L, 20211105)
    x = w.spinor_gaussian(L, 20211106)
    Om = tgm.omega(A, L)
    mass = kappa_to_mass(0.13)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        Ut = ctx.gauge(A, prec=w.F32, recon=w.RECON_12T)
        fin = ctx.field(prec=w.F32, data=tgm.rotate_spinor(x, Om))
        fout = ctx.field(prec=w.F32)
        # full M, then compare both parities: the one-parity launches are its halves
        for dag in (0, 1):
            ctx.apply_M_sloppy(fout, fin, Ut, [mass], bool(dag), 1)
            got = fout.download()
            Af = w.gauge_flip_boundary(L, 3, [a.copy() for a in A])
            ref = tgm.rotate_spinor(oracle.dslash(x, Af, L, mass, dag), Om)
            assert rel_l2(got, ref) < 1e-6


@pytest.mark.parametrize("solver", ["cg_full", "cg_schur", "bicgstab"])
@pytest.mark.parametrize("links", ["F16", "F32"])
@pytest.mark.parametrize("R", [1, "loop"])
def test_solvers_on_rotated_links(w, oracle, solver, links, R):
    L = [6, 4, 6, 10]
    n = 4
    A = w.gauge_near_identity(L, 20211107, 0.3)
    bs = [w.spinor_gaussian(L, 20211108, r) for r in range(n)]
    masses = [kappa_to_mass(k) for k in (0.11, 0.12, 0.13, 0.135)]
    gp = w.F16 if links == "F16" else w.F32
    R = w.RHS_LOOP if R == "loop" else R
    kw = {"cg_full": {}, "cg_schur": {"even_odd": True}, "bicgstab": {}}[solver]
    fn = "bicgstab_mixed" if solver == "bicgstab" else "cg_mixed"
    out = {}
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        fb = ctx.field(nrhs=n, data=bs)
        for recon in (12, w.RECON_12T):
            Us = ctx.gauge(A, prec=gp, recon=recon)
            x = ctx.field(nrhs=n)
            res = getattr(ctx, fn)(x, fb, U, Us, masses, tol=1e-10, delta=1e-2,
                                   rhs_per_thread=R, site_chunk=0, check_every=4, **kw)
            out[recon] = (res, [x.download(r) for r in range(n)])
    res, xs = out[w.RECON_12T]
    its = [s["iterations"] for s in res["systems"]]
    its12 = [s["iterations"] for s in out[12][0]["systems"]]
    print(f"{solver} {links} R={R}: iterations temporal gauge {its}, 12-real {its12}", flush=True)
    assert res["status"] == 0
    Af = w.gauge_flip_boundary(L, 3, [a.copy() for a in A])
    for r in range(n):
        Mdb = oracle.dslash(bs[r], Af, L, masses[r], 1)
        rr = Mdb - oracle.dslash(oracle.dslash(xs[r], Af, L, masses[r], 0), Af, L, masses[r], 1)
        assert res["systems"][r]["converged"]
        assert np.linalg.norm(rr) / np.linalg.norm(Mdb) < 1.05e-10
        assert rel_l2(xs[r], out[12][1][r]) < 1e-7
        assert abs(its[r] - its12[r]) <= max(3, its12[r] // 10)


def test_recon_12t_argument_errors(w):
    L = [4, 4, 4, 4]
    with w.Context(L) as ctx:
        with pytest.raises(w.WilsonError):
            ctx.gauge(None, prec=w.F64, recon=w.RECON_12T)
