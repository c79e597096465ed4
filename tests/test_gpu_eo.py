"""Schur even-odd preconditioned CGNR (wd_cg_eo) against the oracle.

A new algorithm beside the reference's full-lattice CG, so the parity target
is the solution, not the iteration count: the solution must satisfy the
reference's stop test (FP64 |M^dag b - M^dag M x| < tol |M^dag b|, recomputed
here with the oracle's M, which is pinned to the compiled reference), agree
with wd_cg's solution, and take fewer iterations.  Closed forms: all-zero
links make M = (4+m) 1, so x = b / (4+m) after one iteration; b = 0 gives x = 0.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, libc_rand_x0, rel_l2

pytestmark = pytest.mark.gpu


def _normal_residual(oracle, x, b, A, L, mass):
    Mdb = oracle.dslash(b, A, L, mass, 1)
    r = Mdb - oracle.dslash(oracle.dslash(x, A, L, mass, 0), A, L, mass, 1)
    return np.linalg.norm(r) / np.linalg.norm(Mdb)


@pytest.mark.parametrize("kind,kappa,bc", [("su3", 0.12, [1, 1, 1, -1]),
                                            ("near", 0.135, [1, 1, 1, -1]),
                                            ("su3", 0.10, [1, 1, 1, 1])])
def test_cg_eo_solution(w, oracle, kind, kappa, bc):
    L = [8, 8, 8, 8]
    A = w.gauge_random_su3(L, 31) if kind == "su3" else w.gauge_near_identity(L, 31, 0.3)
    b = w.spinor_gaussian(L, 32)
    mass = kappa_to_mass(kappa)
    print(f"cg_eo {kind} kappa={kappa} bc={bc}", flush=True)
    with w.Context(L, bc=bc) as ctx:
        U = ctx.gauge(A)
        fb = ctx.field(data=b)
        xe, xf = ctx.field(), ctx.field()
        eo = ctx.cg(xe, fb, U, mass, max_iter=5000, tol=1e-10, tol_relative=True, use_x0=False,
                    even_odd=True)
        full = ctx.cg(xf, fb, U, mass, max_iter=5000, tol=1e-10, tol_relative=True,
                      use_x0=False)
        x_eo, x_full = xe.download(), xf.download()
    print(f"iterations eo={eo['iterations']} full={full['iterations']} "
          f"true_rsd={eo['true_rsd'] / eo['rhs_norm']:.3e}", flush=True)
    assert eo["converged"] and full["converged"]
    assert eo["true_rsd"] <= 1e-10 * eo["rhs_norm"]
    assert eo["iterations"] < full["iterations"]
    Af = [a.copy() for a in A]
    if bc[3] == -1:
        w.gauge_flip_boundary(L, 3, Af)
    assert _normal_residual(oracle, x_eo, b, Af, L, mass) < 1.05e-10
    assert rel_l2(x_eo, x_full) < 1e-7


def test_cg_eo_start_vector_and_history(w, oracle):
    L = [4, 4, 4, 8]
    A = w.gauge_near_identity(L, 41, 0.3)
    b = w.spinor_gaussian(L, 42)
    x0 = libc_rand_x0(77, w.volume(L) * 12)
    mass = 0.05
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        fb, x = ctx.field(data=b), ctx.field(data=x0)
        res = ctx.cg(x, fb, U, mass, max_iter=2000, tol=1e-11, tol_relative=True, use_x0=True,
                     even_odd=True, check_every=3)
        xs = x.download()
    assert res["converged"]
    assert _normal_residual(oracle, xs, b, A, L, mass) < 1.05e-11
    h = res["history"]
    assert len(h) == res["iterations"] and np.all(np.isfinite(h)) and h[-1] < h[0]


def test_cg_eo_zero_links_closed_form(w):
    L = [4, 4, 4, 4]
    A = [np.zeros(w.volume(L) * 9, dtype=np.complex128) for _ in range(4)]
    b = w.spinor_gaussian(L, 5)
    mass = 0.3
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        fb, x = ctx.field(data=b), ctx.field()
        res = ctx.cg(x, fb, U, mass, max_iter=10, tol=1e-12, tol_relative=True, use_x0=False,
                     even_odd=True)
        xs = x.download()
    assert res["converged"] and res["iterations"] == 1
    assert rel_l2(xs, b / (4 + mass)) < 1e-14


def test_cg_eo_zero_source(w):
    L = [4, 4, 4, 4]
    with w.Context(L) as ctx:
        U = ctx.gauge(w.gauge_random_su3(L, 3))
        fb = ctx.field(data=np.zeros(w.volume(L) * 12, dtype=np.complex128))
        x = ctx.field(data=w.spinor_gaussian(L, 4))
        res = ctx.cg(x, fb, U, 0.1, tol=1e-10, tol_relative=True, use_x0=False, even_odd=True)
        assert res["converged"] and res["iterations"] == 0
        assert np.all(x.download() == 0)


def test_cg_eo_matches_graph_and_eager(w):
    L = [8, 8, 8, 8]
    A = w.gauge_random_su3(L, 51)
    b = w.spinor_gaussian(L, 52)
    out = {}
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        fb = ctx.field(data=b)
        for g in (True, False):
            x = ctx.field()
            r = ctx.cg(x, fb, U, kappa_to_mass(0.12), max_iter=3000, tol=1e-10,
                       tol_relative=True, use_x0=False, even_odd=True, use_graph=g, check_every=7)
            out[g] = (r["iterations"], x.download(), r["history"])
    assert out[True][0] == out[False][0]
    assert np.array_equal(out[True][1], out[False][1])
    assert np.array_equal(out[True][2], out[False][2])


# ---- mixed precision on the Schur system -------------------------------------------
@pytest.mark.parametrize("sprec,gprec,recon", [("F16", "F32", 12), ("F32", "F32", 18),
                                                ("F16", "F16", 12)])
def test_mixed_eo_reaches_fp64_true_residual(w, oracle, sprec, gprec, recon):
    """Three systems at three kappas, antiperiodic t: every solution meets the
    reference's FP64 stop test on the full system, re-checked with the oracle."""
    L = [8, 8, 8, 8]
    n = 3
    A = w.gauge_near_identity(L, 61, 0.3)
    bs = [w.spinor_gaussian(L, 62, r) for r in range(n)]
    masses = [kappa_to_mass(k) for k in (0.12, 0.125, 0.13)]
    print(f"mixed eo {sprec}/{gprec}/{recon}", flush=True)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=getattr(w, gprec), recon=recon)
        fb = ctx.field(nrhs=n, data=bs)
        xe, xf = ctx.field(nrhs=n), ctx.field(nrhs=n)
        eo = ctx.cg_mixed(xe, fb, U, Us, masses, tol=1e-10, spinor_prec=getattr(w, sprec),
                          even_odd=True, check_every=4)
        full = ctx.cg_mixed(xf, fb, U, Us, masses, tol=1e-10, spinor_prec=getattr(w, sprec),
                            check_every=4)
        xs = [xe.download(r) for r in range(n)]
    assert eo["status"] == 0
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in range(n):
        s, sf = eo["systems"][r], full["systems"][r]
        print(r, s["iterations"], sf["iterations"], s["true_rel_rsd"], flush=True)
        assert s["converged"] and s["true_rel_rsd"] <= 1e-10
        assert s["iterations"] < sf["iterations"]
        assert _normal_residual(oracle, xs[r], bs[r], Af, L, masses[r]) < 1.05e-10


@pytest.mark.parametrize("L", [[4, 4, 6, 8], [6, 6, 6, 6]])
@pytest.mark.parametrize("sprec", ["F16", "F32"])
def test_mixed_eo_half_volume_not_a_multiple_of_the_block(w, oracle, L, sprec):
    """Half volumes of 384 and 648 sites: the last workgroup of a parity runs
    past the even half (regression: the elementwise kernels once spilled into
    the odd half there, corrupting the Schur residual norms)."""
    n = 2
    A = w.gauge_near_identity(L, 81, 0.3)
    bs = [w.spinor_gaussian(L, 82, r) for r in range(n)]
    masses = [kappa_to_mass(k) for k in (0.11, 0.125)]
    print(f"mixed eo L={L} {sprec}", flush=True)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16 if sprec == "F16" else w.F32, recon=12)
        fb = ctx.field(nrhs=n, data=bs)
        x = ctx.field(nrhs=n)
        res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, spinor_prec=getattr(w, sprec),
                           even_odd=True, check_every=4)
        xs = [x.download(r) for r in range(n)]
    assert res["status"] == 0
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in range(n):
        s = res["systems"][r]
        print(r, s["iterations"], s["true_rel_rsd"], flush=True)
        assert s["converged"] and s["iterations"] < 120
        assert _normal_residual(oracle, xs[r], bs[r], Af, L, masses[r]) < 1.05e-10


def test_mixed_eo_masks_and_zero_source(w, oracle):
    L = [4, 4, 4, 8]
    n = 4
    A = w.gauge_near_identity(L, 71, 0.3)
    bs = [w.spinor_gaussian(L, 72, r) for r in range(n)]
    bs[2] = np.zeros_like(bs[2])
    masses = [2.0, 0.5, 0.3, kappa_to_mass(0.13)]
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F32, recon=12)
        fb = ctx.field(nrhs=n, data=bs)
        x = ctx.field(nrhs=n)
        res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=2, check_every=3,
                           even_odd=True)
        xs = [x.download(r) for r in range(n)]
    its = [s["iterations"] for s in res["systems"]]
    print(its, flush=True)
    assert res["status"] == 0
    assert its[0] < its[1] < its[3] and its[2] == 0
    assert np.all(xs[2] == 0)
    for r in (0, 1, 3):
        assert _normal_residual(oracle, xs[r], bs[r], A, L, masses[r]) < 1.05e-10


def test_mixed_eo_graph_and_direct_launch_agree(w):
    L = [4, 4, 4, 8]
    A = w.gauge_near_identity(L, 81, 0.3)
    bs = [w.spinor_gaussian(L, 82, r) for r in range(2)]
    out = []
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=2, data=bs)
        for g in (True, False):
            x = ctx.field(nrhs=2)
            res = ctx.cg_mixed(x, fb, U, Us, [0.1, 0.2], use_graph=g, check_every=3,
                               rhs_per_thread=1, even_odd=True)
            out.append(([x.download(r) for r in range(2)], res))
    for r in range(2):
        assert np.array_equal(out[0][0][r], out[1][0][r])
        assert out[0][1]["systems"][r]["iterations"] == out[1][1]["systems"][r]["iterations"]


@pytest.mark.parametrize("solver", ["cg_mixed", "bicgstab_mixed"])
def test_schur_stopped_systems_report_measured_residual(w, oracle, solver):
    """Systems that stop at the iteration limit come back with x_o rebuilt from
    their final x_e, x_o = (b_o + D_oe x_e / 2) / a, and with the measured
    full-system residual |M^dag b - M^dag M x| / |M^dag b| of the returned x
    (here far above tol), not a placeholder."""
    L = [4, 4, 4, 8]
    A = w.gauge_near_identity(L, 91, 0.3)
    bs = [w.spinor_gaussian(L, 92, r) for r in range(2)]
    masses = [kappa_to_mass(0.13), kappa_to_mass(0.135)]
    V = w.volume(L)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=2, data=bs)
        X = ctx.field(nrhs=2)
        kw = {"even_odd": True} if solver == "cg_mixed" else {}
        res = getattr(ctx, solver)(X, fb, U, Us, masses, max_iter=4, check_every=2, **kw)
        xs = [X.download(r) for r in range(2)]
    assert res["status"] == w.WD_ENOTCONV
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in range(2):
        s = res["systems"][r]
        x, b, mass = xs[r], bs[r], masses[r]
        assert not s["converged"] and 0 < s["iterations"] <= 4
        # x_o from x_e: or_hop(x, cb=1) is -D_oe x_e / 2 on the odd half
        xe_only = x.copy()
        xe_only[V // 2 * 12:] = 0
        hop_o = oracle.hop(xe_only, Af, L, 0, 1)[V // 2 * 12:]
        want_o = (b[V // 2 * 12:] - hop_o) / (4.0 + mass)
        assert rel_l2(x[V // 2 * 12:], want_o) < 1e-13
        meas = _normal_residual(oracle, x, b, Af, L, mass)
        print(f"{solver} system {r}: reported {s['true_rel_rsd']:.6e} measured {meas:.6e}",
              flush=True)
        assert s["true_rel_rsd"] > 1e-10
        assert abs(s["true_rel_rsd"] - meas) <= 1e-8 * meas
