"""Mixed-precision BiCGStab on the Schur system (wd_bicgstab_mixed).

A new algorithm beside the reference's CGNR, so its parity target is the
solution: every system must meet the reference's stop test on the full
system, FP64 |M^dag b - M^dag M x| < 1e-10 |M^dag b|, re-checked here with the
oracle's M (pinned to the compiled reference), and agree with wd_cg's
solution.  Closed forms: zero links make A = a^2, so x = b / a; b = 0 gives
x = 0.  Work orders (graph or direct launches, reliable-update slot spacing)
must not change iterates.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, rel_l2

pytestmark = pytest.mark.gpu


def _normal_residual(oracle, x, b, A, L, mass):
    Mdb = oracle.dslash(b, A, L, mass, 1)
    r = Mdb - oracle.dslash(oracle.dslash(x, A, L, mass, 0), A, L, mass, 1)
    return np.linalg.norm(r) / np.linalg.norm(Mdb)


@pytest.mark.parametrize("sprec,gprec,recon", [("F16", "F16", 12), ("F16", "F32", 12),
                                                ("F32", "F32", 18)])
@pytest.mark.parametrize("L", [[8, 8, 8, 8], [4, 4, 6, 8]])
def test_bicgstab_reaches_fp64_true_residual(w, oracle, sprec, gprec, recon, L):
    """Three systems at three kappas, antiperiodic t (a half volume of 384
    sites included: not a multiple of the workgroup)."""
    n = 3
    A = w.gauge_near_identity(L, 61, 0.3)
    bs = [w.spinor_gaussian(L, 62, r) for r in range(n)]
    masses = [kappa_to_mass(k) for k in (0.12, 0.125, 0.13)]
    print(f"bicgstab L={L} {sprec}/{gprec}/{recon}", flush=True)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=getattr(w, gprec), recon=recon)
        fb = ctx.field(nrhs=n, data=bs)
        X, Xe = ctx.field(nrhs=n), ctx.field(nrhs=n)
        bi = ctx.bicgstab_mixed(X, fb, U, Us, masses, tol=1e-10, spinor_prec=getattr(w, sprec),
                                check_every=4)
        cg = ctx.cg_mixed(Xe, fb, U, Us, masses, tol=1e-10, spinor_prec=getattr(w, sprec),
                          even_odd=True, check_every=4)
        xs = [X.download(r) for r in range(n)]
        ref = []
        for r in range(n):
            x1 = ctx.field()
            ctx.cg(x1, ctx.field(data=bs[r]), U, masses[r], max_iter=5000, tol=1e-10,
                   tol_relative=True, use_x0=False)
            ref.append(x1.download())
    assert bi["status"] == 0
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in range(n):
        s, sc = bi["systems"][r], cg["systems"][r]
        print(r, "bicgstab", s["iterations"], s["reliable_updates"], "cg", sc["iterations"],
              s["true_rel_rsd"], flush=True)
        assert s["converged"] and s["true_rel_rsd"] <= 1e-10
        assert s["iterations"] < sc["iterations"]
        assert _normal_residual(oracle, xs[r], bs[r], Af, L, masses[r]) < 1.05e-10
        assert rel_l2(xs[r], ref[r]) < 1e-7


def test_bicgstab_closed_forms(w):
    """Zero links: A = a^2, the solution is b / a, found within a few
    iterations (the first reliable update corrects the F32 accumulation of
    x).  A zero right-hand side is solved by x = 0 without iterating."""
    L = [4, 4, 4, 8]
    zero = [np.zeros(w.volume(L) * 9, dtype=np.complex128) for _ in range(4)]
    b = w.spinor_gaussian(L, 71)
    mass = 0.3
    with w.Context(L) as ctx:
        U = ctx.gauge(zero)
        Us = ctx.gauge(zero, prec=w.F32, recon=18)
        fb = ctx.field(nrhs=2, data=[b, np.zeros_like(b)])
        X = ctx.field(nrhs=2)
        res = ctx.bicgstab_mixed(X, fb, U, Us, [mass, mass], spinor_prec=w.F32)
        x0, x1 = X.download(0), X.download(1)
    assert res["status"] == 0
    assert res["systems"][0]["iterations"] <= 3
    assert rel_l2(x0, b / (4 + mass)) < 1e-10
    assert res["systems"][1]["iterations"] == 0 and np.all(x1 == 0)


def test_bicgstab_work_orders_bitwise(w):
    """Graph replay or direct launches, reliable-update slots after every
    iteration, every 3rd, or every 10th in blocks of 20 (groups of 8
    iterations, the most whose p and s are kept for one x_s pass, end between
    the slots): identical iterates and counts."""
    L = [4, 4, 4, 8]
    A = w.gauge_near_identity(L, 81, 0.3)
    bs = [w.spinor_gaussian(L, 82, r) for r in range(2)]
    masses = [kappa_to_mass(0.12), kappa_to_mass(0.13)]
    out = []
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=2, data=bs)
        for g, K, chunk in ((True, 1, 3), (False, 1, 3), (True, 3, 3), (True, 10, 20)):
            X = ctx.field(nrhs=2)
            res = ctx.bicgstab_mixed(X, fb, U, Us, masses, use_graph=g, check_every=chunk,
                                     rhs_per_thread=2, reliable_every=K)
            out.append(([X.download(r) for r in range(2)], res))
            X.close()
    for k in (1, 2, 3):
        for r in range(2):
            assert np.array_equal(out[k][0][r], out[0][0][r])
            assert out[k][1]["systems"][r]["iterations"] == out[0][1]["systems"][r]["iterations"]


def test_bicgstab_iteration_limit_and_argument_errors(w):
    """max_iter below what the systems need: status WD_ENOTCONV, no system
    reported converged, at most max_iter iterations, a finite partial
    solution.  Argument errors come back as WD_EARG before any launch."""
    L = [4, 4, 4, 8]
    A = w.gauge_near_identity(L, 91, 0.3)
    bs = [w.spinor_gaussian(L, 92, r) for r in range(2)]
    masses = [kappa_to_mass(0.13), kappa_to_mass(0.135)]
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=2, data=bs)
        X = ctx.field(nrhs=2)
        res = ctx.bicgstab_mixed(X, fb, U, Us, masses, max_iter=4, check_every=2)
        assert res["status"] == w.WD_ENOTCONV
        for r in range(2):
            s = res["systems"][r]
            assert not s["converged"] and 0 < s["iterations"] <= 4
            assert np.all(np.isfinite(X.download(r)))
        for kwargs, what in (({"tol": 0.0}, "tol"), ({"delta": 1.0}, "delta"),
                             ({"rhs_per_thread": 3}, "rhs_per_thread"),
                             ({"spinor_prec": w.F64}, "spinor_prec")):
            with pytest.raises(w.WilsonError) as e:
                ctx.bicgstab_mixed(X, fb, U, Us, masses, **kwargs)
            assert e.value.code == w.WD_EARG, what
        with pytest.raises(w.WilsonError) as e:
            ctx.bicgstab_mixed(fb, fb, U, Us, masses)
        assert e.value.code == w.WD_EARG
        with pytest.raises(w.WilsonError) as e:   # sloppy links where U must be F64
            ctx.bicgstab_mixed(X, fb, Us, Us, masses)
        assert e.value.code == w.WD_EARG


def test_bicgstab_rhs_per_thread_and_chunk_bitwise(w):
    """The autotuner's choices (systems per thread R, site-chunk order C) move
    work between threads and workgroups only: five systems (groups of R = 4
    leave one short), identical solutions and counts for every (R, C)."""
    L = [4, 4, 8, 8]
    n = 5
    A = w.gauge_near_identity(L, 111, 0.3)
    bs = [w.spinor_gaussian(L, 112, r) for r in range(n)]
    masses = [kappa_to_mass(k) for k in (0.12, 0.122, 0.125, 0.128, 0.13)]
    out = []
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=n, data=bs)
        for R, C in ((1, 0), (2, 0), (4, 0), (2, 16), (4, 32)):
            X = ctx.field(nrhs=n)
            res = ctx.bicgstab_mixed(X, fb, U, Us, masses, rhs_per_thread=R, site_chunk=C)
            assert res["status"] == 0 and res["systems"][0]["rhs_per_thread"] == R
            out.append(([s["iterations"] for s in res["systems"]],
                        [X.download(r) for r in range(n)]))
            X.close()
    for its, xs in out[1:]:
        assert its == out[0][0]
        for r in range(n):
            assert np.array_equal(xs[r], out[0][1][r])


def test_bicgstab_repeats_bit_for_bit(w):
    """Regression: the scalar and elementwise BiCGStab kernels take two block
    sums each, and the sums share one LDS array per kernel; without a barrier
    after thread 0 has read the first sum, the second could overwrite it, and
    repeated solves of one system differed in their last bits (and by an
    iteration or two).  24^3x48, one system, five solves in one context after
    a mixed CG solve: identical iterates, iteration counts and residuals."""
    L = [24, 24, 24, 48]
    A = w.gauge_near_identity(L, 20211025, 0.3)
    b = w.spinor_gaussian(L, 20211026, 0)
    masses = [kappa_to_mass(0.135)]
    print(f"bicgstab repeat L={L}", flush=True)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=w.RECON_12T)
        fb = ctx.field(nrhs=1, data=[b])
        x = ctx.field(nrhs=1)
        ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, delta=1e-2, rhs_per_thread=1,
                     site_chunk=0, even_odd=True)
        runs = []
        for rep in range(5):
            res = ctx.bicgstab_mixed(x, fb, U, Us, masses, tol=1e-10, delta=1e-2,
                                     rhs_per_thread=1, site_chunk=0)
            s = res["systems"][0]
            runs.append((x.download(0), s["iterations"], s["true_rel_rsd"]))
            print(rep, s["iterations"], s["true_rel_rsd"], flush=True)
    assert all(r[1] == runs[0][1] and r[2] == runs[0][2] for r in runs)
    for r in runs[1:]:
        assert np.array_equal(r[0], runs[0][0])
