"""FP64 HIP path vs the CPU oracle, through the C-ABI (libwilson.so).

Tolerance: 1e-13 relative L2 for the operator (the bar for the drop-in
Dslash), equal iteration counts for CG.  Inputs come from the seeded
generators in libwilson.so and are fed identically to both sides.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, libc_rand_x0, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-13
SHAPES = [[4, 4, 4, 4], [4, 4, 4, 8], [2, 4, 6, 8], [8, 8, 8, 8], [6, 4, 2, 10]]


def _inputs(w, L, seed=20211025, kind="su3"):
    if kind == "su3":
        A = w.gauge_random_su3(L, seed)
    elif kind == "near":
        A = w.gauge_near_identity(L, seed, 0.3)
    else:
        A = w.gauge_unit(L)
    b = w.spinor_gaussian(L, seed + 1)
    return A, b


@pytest.mark.parametrize("L", SHAPES, ids=lambda L: "x".join(map(str, L)))
@pytest.mark.parametrize("dag", [0, 1])
@pytest.mark.parametrize("cb", [0, 1])
def test_hop_matches_oracle(w, oracle, L, dag, cb):
    A, b = _inputs(w, L)
    print(f"hop L={L} dag={dag} cb={cb}", flush=True)
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        fin, fout = ctx.field(data=b), ctx.field()
        ctx.hop(fout, fin, U, bool(dag), cb)
        got = fout.download()
    ref = oracle.hop(b, A, L, dag, cb)
    assert rel_l2(got, ref) < TOL
    # the other parity half is exactly zero (src/dslash_new.cpp:1398)
    Vh = w.volume(L) // 2
    other = got.reshape(2, Vh * 12)[1 - cb]
    assert np.all(other == 0)


@pytest.mark.parametrize("L", SHAPES, ids=lambda L: "x".join(map(str, L)))
@pytest.mark.parametrize("dag", [0, 1])
def test_apply_M_matches_oracle(w, oracle, L, dag):
    A, b = _inputs(w, L, kind="near")
    mass = kappa_to_mass(0.13)
    print(f"apply_M L={L} dag={dag}", flush=True)
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        fin, fout = ctx.field(data=b), ctx.field()
        ctx.apply_M(fout, fin, U, mass, bool(dag))
        got = fout.download()
    ref = oracle.dslash(b, A, L, mass, dag)
    assert rel_l2(got, ref) < TOL


def test_diag_matches_oracle(w, oracle):
    L = [4, 4, 4, 8]
    _, b = _inputs(w, L)
    with w.Context(L) as ctx:
        fin, fout = ctx.field(data=b), ctx.field()
        ctx.diag(fout, fin, 0.1)
        got = fout.download()
    assert np.array_equal(got, oracle.diag(b, L, 0.1))


@pytest.mark.parametrize("mu", [0, 1, 2, 3])
def test_antiperiodic_equals_flipped_links(w, oracle, mu):
    """bc[mu] = -1 on the GPU equals the periodic oracle on links whose last
    mu-slice is negated (the reference has periodic boundaries only)."""
    L = [4, 6, 4, 8]
    A, b = _inputs(w, L)
    bc = [1, 1, 1, 1]
    bc[mu] = -1
    mass = kappa_to_mass(0.12)
    with w.Context(L, bc=bc) as ctx:
        U = ctx.gauge(A)
        fin, fout = ctx.field(data=b), ctx.field()
        ctx.apply_M(fout, fin, U, mass, False)
        got = fout.download()
        ctx.hop(fout, fin, U, True, 1)
        got_hop = fout.download()
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, mu, Af)
    assert rel_l2(got, oracle.dslash(b, Af, L, mass, 0)) < TOL
    assert rel_l2(got_hop, oracle.hop(b, Af, L, 1, 1)) < TOL


def test_mdagm_matches_two_applications(w, oracle):
    L = [4, 4, 4, 8]
    A, b = _inputs(w, L, kind="near")
    mass = kappa_to_mass(0.135)
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        fin, fout, tmp = ctx.field(data=b), ctx.field(), ctx.field()
        ctx.apply_MdagM(fout, fin, tmp, U, mass)
        got = fout.download()
    ref = oracle.dslash(oracle.dslash(b, A, L, mass, 0), A, L, mass, 1)
    assert rel_l2(got, ref) < TOL


def test_norm_and_dot(w, oracle):
    L = [4, 4, 4, 8]
    a = w.spinor_gaussian(L, 7, 0)
    b = w.spinor_gaussian(L, 7, 1)
    with w.Context(L) as ctx:
        fa, fb = ctx.field(data=a), ctx.field(data=b)
        n2 = ctx.norm2(fa)
        d = ctx.dot(fa, fb)
    assert abs(n2 - oracle.norm2(a)) <= 1e-13 * n2
    ref = oracle.dot(a, b)
    assert abs(d - ref) <= 1e-13 * abs(ref) + 1e-12


def test_upload_download_roundtrip(w):
    L = [4, 4, 6, 8]
    a = w.spinor_gaussian(L, 3)
    with w.Context(L) as ctx:
        f64 = ctx.field(data=a)
        assert np.array_equal(f64.download(), a)
        f32 = ctx.field(prec=w.F32, data=a)
        assert rel_l2(f32.download(), a) < 1e-7
        f16 = ctx.field(prec=w.F16, data=a)
        assert rel_l2(f16.download(), a) < 1e-3
        back = ctx.field()
        back.copy_from(f16)
        assert rel_l2(back.download(), a) < 1e-3


@pytest.mark.parametrize("use_graph", [True, False])
def test_cg_matches_oracle_relative(w, oracle, use_graph):
    """8^4, random SU(3), kappa 0.12, antiperiodic t (link flip on the
    oracle side), CG to 1e-10 relative: same iteration count and history."""
    L = [8, 8, 8, 8]
    A, b = _inputs(w, L)
    mass = kappa_to_mass(0.12)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        fb, x = ctx.field(data=b), ctx.field()
        res = ctx.cg(x, fb, U, mass, max_iter=1000, tol=1e-10, tol_relative=True, use_x0=False,
                     use_graph=use_graph, check_every=5)
        xs = x.download()
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    acc = 1e-10 * res["rhs_norm"]
    it, xo, hist = oracle.cg(b, np.zeros_like(b), Af, L, mass, 1000, acc)
    print(f"cg iterations gpu={res['iterations']} oracle={it}", flush=True)
    assert res["converged"] and it == res["iterations"]
    assert np.allclose(res["history"], hist, rtol=1e-8, atol=0)
    assert rel_l2(xs, xo) < 1e-9
    assert res["true_rsd"] < 1.5e-10 * res["rhs_norm"]


def test_cg_reference_start_vector(w, oracle):
    """The reference seeds x0 with rand()/RAND_MAX (src/invert.cpp:42-44) and
    stops on the absolute |r| < acc."""
    L = [4, 4, 4, 8]
    A, b = _inputs(w, L, kind="near")
    mass = 0.05
    x0 = libc_rand_x0(1234, w.volume(L) * 12)
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        fb, x = ctx.field(data=b), ctx.field(data=x0)
        res = ctx.cg(x, fb, U, mass, max_iter=500, tol=1e-9, use_x0=True)
        xs = x.download()
    it, xo, hist = oracle.cg(b, x0, A, L, mass, 500, 1e-9)
    assert res["converged"] and it == res["iterations"]
    assert rel_l2(xs, xo) < 1e-9


def test_cg_iteration_limit(w):
    L = [4, 4, 4, 4]
    A, b = _inputs(w, L)
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        fb, x = ctx.field(data=b), ctx.field()
        res = ctx.cg(x, fb, U, 0.0, max_iter=4, tol=1e-30, use_x0=False)
    assert res["status"] == w.WD_ENOTCONV
    assert res["iterations"] == 3 and not res["converged"]


def test_zero_links_closed_form(w):
    """All-zero links (the reference's "RANDOM" gauge stub) make M = (4+m) 1:
    CG converges in one iteration to x = b / (4+m)."""
    L = [4, 4, 4, 4]
    A = [np.zeros(w.volume(L) * 9, dtype=np.complex128) for _ in range(4)]
    b = w.spinor_gaussian(L, 5)
    mass = 0.3
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        fb, x = ctx.field(data=b), ctx.field()
        res = ctx.cg(x, fb, U, mass, max_iter=10, tol=1e-12, tol_relative=True, use_x0=False)
        xs = x.download()
    assert res["iterations"] == 1
    assert rel_l2(xs, b / (4 + mass)) < 1e-14


def test_argument_errors(w):
    with pytest.raises(w.WilsonError):
        w.Context([3, 4, 4, 4])
    L = [4, 4, 4, 4]
    with w.Context(L) as ctx:
        U = ctx.gauge(w.gauge_unit(L))
        f = ctx.field()
        with pytest.raises(w.WilsonError):
            ctx.apply_M(f, f, U, 0.1, False)
        g32 = ctx.field(prec=w.F32)
        with pytest.raises(w.WilsonError):
            ctx.apply_M(g32, f, U, 0.1, False)
