"""Solver work memory kept by the context (wd_release_workspace,
wd_workspace_bytes).

The work vectors of a finished solve serve the next solve of the same shape;
every solver zeroes what it takes, so a solve on reused memory must give
bitwise the same solution and iteration count as one on fresh memory, for
the FP64 CG (full and Schur), the mixed CG and the mixed BiCGStab.  Releasing
the cache empties it, and fields allocated after solves start at zero.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass

pytestmark = pytest.mark.gpu


def _solves(w, ctx, U, Us, fb, fb3, masses):
    """One solve of each kind; returns (name, iterations, solution) triples."""
    out = []
    for eo in (False, True):
        x = ctx.field()
        r = ctx.cg(x, fb, U, masses[0], max_iter=2000, tol=1e-10, tol_relative=True,
                   use_x0=False, check_every=8, even_odd=eo)
        out.append((f"cg eo={eo}", r["iterations"], x.download()))
    for name, fn, kw in (("cg_mixed", ctx.cg_mixed, {"even_odd": True}),
                         ("bicgstab_mixed", ctx.bicgstab_mixed, {})):
        X = ctx.field(nrhs=3)
        r = fn(X, fb3, U, Us, masses, tol=1e-10, spinor_prec=w.F16, **kw)
        out.append((name, [s["iterations"] for s in r["systems"]],
                    np.concatenate([X.download(k) for k in range(3)])))
    return out


def test_reused_work_memory_gives_identical_solves(w):
    L = [8, 8, 8, 16]
    A = w.gauge_near_identity(L, 71, 0.3)
    bs = [w.spinor_gaussian(L, 72, r) for r in range(3)]
    masses = [kappa_to_mass(k) for k in (0.12, 0.125, 0.13)]
    print(f"workspace L={L}", flush=True)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(data=bs[0])
        fb3 = ctx.field(nrhs=3, data=bs)
        assert ctx.workspace_bytes() == 0
        first = _solves(w, ctx, U, Us, fb, fb3, masses)
        cached = ctx.workspace_bytes()
        # at least the three F64 work fields of the mixed solvers (24 reals per site)
        assert cached >= 3 * 3 * 24 * 8 * w.volume(L)
        second = _solves(w, ctx, U, Us, fb, fb3, masses)
        assert ctx.workspace_bytes() == cached, "equal shapes must reuse, not grow, the cache"
        ctx.release_workspace()
        assert ctx.workspace_bytes() == 0
        third = _solves(w, ctx, U, Us, fb, fb3, masses)
        for a, b, c in zip(first, second, third):
            assert a[0] == b[0] == c[0]
            assert a[1] == b[1] == c[1], a[0]
            assert np.array_equal(a[2], b[2]) and np.array_equal(a[2], c[2]), a[0]
        # fields created after the solves start at zero
        f = ctx.field(nrhs=3)
        assert all(not np.any(f.download(k)) for k in range(3))


def test_failed_allocation_releases_the_cache(w):
    """A field larger than the device (about 630 GB: hipMalloc refuses it, no
    kernel runs) fails with WD_ENOMEM after the cached work blocks were
    released for the retry; solves afterwards still work."""
    L = [8, 8, 8, 16]
    A = w.gauge_near_identity(L, 73, 0.3)
    b = w.spinor_gaussian(L, 74)
    mass = kappa_to_mass(0.125)
    print(f"workspace release on failed allocation L={L}", flush=True)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        fb = ctx.field(data=b)
        x = ctx.field()
        r1 = ctx.cg(x, fb, U, mass, max_iter=2000, tol=1e-10, tol_relative=True, use_x0=False)
        x1 = x.download()
        assert ctx.workspace_bytes() > 0
        with pytest.raises(w.WilsonError) as e:
            ctx.field(nrhs=400_000)
        assert e.value.code == w.WD_ENOMEM
        assert ctx.workspace_bytes() == 0
        r2 = ctx.cg(x, fb, U, mass, max_iter=2000, tol=1e-10, tol_relative=True, use_x0=False)
        assert r2["iterations"] == r1["iterations"]
        assert np.array_equal(x.download(), x1)
