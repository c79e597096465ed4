"""wd_pion_correlator_noalloc (no allocation; partial sums in the library's
fixed device array) against wd_pion_correlator, through the C-ABI.

The two entry points sum the same per-slice partials in the same order, so
their results agree bit for bit.  Two calls in a row share the device array:
the second one, on fewer systems and another source time slice, must not see
the first one's partials.  Sizes beyond the array come back as WD_EARG.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass

pytestmark = pytest.mark.gpu


def _propagators(w, ctx, L, n):
    A = w.gauge_near_identity(L, 20211025, 0.3)
    U = ctx.gauge(A)
    Us = ctx.gauge(A, prec=w.F16, recon=12)
    site = [1, 2, 3, 5]
    srcs = [w.spinor_point(L, site, s, c) for s in range(4) for c in range(3)][:n]
    fb = ctx.field(nrhs=n, data=srcs)
    X = ctx.field(nrhs=n)
    ctx.cg_mixed(X, fb, U, Us, [kappa_to_mass(0.13)] * n, tol=1e-10, even_odd=True)
    return X


def test_noalloc_correlator_matches(w):
    L = [8, 8, 8, 16]
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        X = _propagators(w, ctx, L, 12)
        for rhs_first, rhs_count, t0 in ((0, 12, 5), (3, 4, 11)):
            want = ctx.pion_correlator(X, t0, rhs_first, rhs_count)
            got = ctx.pion_correlator_noalloc(X, t0, rhs_first, rhs_count)
            print(f"systems {rhs_first}..{rhs_first + rhs_count - 1} t0={t0}: C(0)={got[0]:.6e}",
                  flush=True)
            assert np.all(want > 0)
            assert np.array_equal(got, want)
    # a second context on a smaller lattice: the array keeps the last call's
    # contents, which must not leak into this one
    L2 = [4, 4, 4, 8]
    with w.Context(L2, bc=[1, 1, 1, -1]) as ctx:
        X = _propagators(w, ctx, L2, 3)
        want = ctx.pion_correlator(X, 2)
        got = ctx.pion_correlator_noalloc(X, 2)
        assert got.shape == (8,) and np.array_equal(got, want)


def test_noalloc_correlator_limits(w):
    L = [4, 4, 4, 8]
    with w.Context(L) as ctx:
        X = ctx.field(nrhs=2)
        with pytest.raises(w.WilsonError) as e:
            ctx.pion_correlator_noalloc(X, 8)           # t0 outside the lattice
        assert e.value.code == w.WD_EARG
    L = [4, 4, 4, 1024]                                  # Lt above 512
    with w.Context(L) as ctx:
        X = ctx.field(nrhs=1)
        with pytest.raises(w.WilsonError) as e:
            ctx.pion_correlator_noalloc(X, 0)
        assert e.value.code == w.WD_EARG
    L = [4, 4, 4, 256]                                   # 40 * 2 * 256 > 16384 partials
    with w.Context(L) as ctx:
        X = ctx.field(nrhs=40)
        with pytest.raises(w.WilsonError) as e:
            ctx.pion_correlator_noalloc(X, 0)
        assert e.value.code == w.WD_EARG
        # the largest allowed count works and matches
        X.upload(w.spinor_gaussian(L, 7), 0)
        got = ctx.pion_correlator_noalloc(X, 3, 0, 32)
        assert np.array_equal(got, ctx.pion_correlator(X, 3, 0, 32))
