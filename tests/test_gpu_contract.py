"""Pion correlator of point-source propagators (wd_pion_correlator).

C(dt) = sum_r sum_{x : t(x) = t0 + dt} |S_r(x)|^2 over the 12 spin x colour
solutions S_r = M^-1 delta of one source (refs/python/contraction.py:25-45
for the pion).  Checks: the kernel against a NumPy contraction of the same
data (exact to rounding), the whole chain (12 GPU solves + contraction)
against propagators solved by the oracle, the zero-link closed form
S = delta / (4+m) (C(0) = 12 / (4+m)^2, else 0), and time-reflection symmetry
C(dt) = C(Lt - dt) of the free field with periodic boundaries.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass

pytestmark = pytest.mark.gpu


def _t_of_eo(L):
    """Time coordinate of every eo-ordered site (include/operator.h:56-67)."""
    V = int(np.prod(L))
    i = np.arange(V) % (V // 2)
    return i // (L[0] // 2 * L[1] * L[2])


def _numpy_correlator(props, L, t0):
    t = _t_of_eo(L)
    w = np.zeros(L[3])
    for S in props:
        n2 = np.sum(np.abs(S.reshape(-1, 12)) ** 2, axis=1)
        w += np.bincount(t, weights=n2, minlength=L[3])
    return np.roll(w, -t0)


def _point_sources(w, L, site):
    return [w.spinor_point(L, site, s, c) for s in range(4) for c in range(3)]


def test_kernel_matches_numpy(w):
    L = [4, 6, 4, 10]
    props = [w.spinor_gaussian(L, 91, r) for r in range(14)]
    with w.Context(L) as ctx:
        f = ctx.field(nrhs=14, data=props)
        for first, count, t0 in ((0, 12, 3), (2, 12, 0), (5, 1, 9)):
            got = ctx.pion_correlator(f, t0, first, count)
            ref = _numpy_correlator(props[first:first + count], L, t0)
            assert np.allclose(got, ref, rtol=1e-13, atol=0), (first, count, t0)


def test_point_source_chain_vs_oracle(w, oracle):
    L = [4, 4, 4, 8]
    A = w.gauge_random_su3(L, 93)
    mass = kappa_to_mass(0.12)
    site = [1, 2, 3, 5]
    srcs = _point_sources(w, L, site)
    props = []
    for b in srcs:
        Mdb = oracle.dslash(b, A, L, mass, 1)
        it, x, _ = oracle.cg(b, np.zeros_like(b), A, L, mass, 5000, 1e-13 * np.linalg.norm(Mdb))
        props.append(x)
    ref = _numpy_correlator(props, L, site[3])
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        Uh = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=12, data=srcs)
        x = ctx.field(nrhs=12)
        res = ctx.cg_mixed(x, fb, U, Uh, [mass] * 12, tol=1e-11, even_odd=True)
        got = ctx.pion_correlator(x, site[3])
    print("C(dt) gpu", got, flush=True)
    assert res["status"] == 0 and np.all(got > 0)
    assert np.max(np.abs(got - ref) / ref) < 1e-8
    assert got[0] == np.max(got)  # the correlator peaks at the source


def test_zero_links_closed_form(w):
    L = [4, 4, 4, 6]
    A = [np.zeros(w.volume(L) * 9, dtype=np.complex128) for _ in range(4)]
    mass = 0.3
    site = [0, 1, 2, 4]
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        fb = ctx.field(nrhs=12, data=_point_sources(w, L, site))
        x = ctx.field(nrhs=12)
        for r in range(12):
            xr, br = ctx.field(), ctx.field(data=fb.download(r))
            ctx.cg(xr, br, U, mass, max_iter=10, tol=1e-14, tol_relative=True, use_x0=False)
            x.upload(xr.download(), r)
        C = ctx.pion_correlator(x, site[3])
    assert abs(C[0] - 12 / (4 + mass) ** 2) < 1e-14 * C[0]
    assert np.all(C[1:] == 0)


def test_free_field_time_reflection(w):
    L = [4, 4, 4, 8]
    A = w.gauge_unit(L)
    mass = kappa_to_mass(0.11)
    site = [0, 0, 0, 0]
    with w.Context(L, bc=[1, 1, 1, 1]) as ctx:
        U = ctx.gauge(A)
        x = ctx.field(nrhs=12)
        for r, b in enumerate(_point_sources(w, L, site)):
            xr = ctx.field()
            ctx.cg(xr, ctx.field(data=b), U, mass, max_iter=2000, tol=1e-13, tol_relative=True,
                   use_x0=False, even_odd=True)
            x.upload(xr.download(), r)
        C = ctx.pion_correlator(x, 0)
    print("free C(dt)", C, flush=True)
    for dt in range(1, 8):
        assert abs(C[dt] - C[8 - dt]) < 1e-9 * C[dt]
    assert np.all(np.diff(C[:5]) < 0)  # falls towards the middle of the lattice


def test_argument_errors(w):
    L = [4, 4, 4, 4]
    with w.Context(L) as ctx:
        f = ctx.field(nrhs=2)
        with pytest.raises(w.WilsonError):
            ctx.pion_correlator(f, 0, 1, 2)    # range beyond the field
        with pytest.raises(w.WilsonError):
            ctx.pion_correlator(f, 4)          # t0 outside the lattice
        with pytest.raises(w.WilsonError):
            ctx.pion_correlator(ctx.field(prec=w.F16, nrhs=12), 0)
