"""Batched FP64 hops with two systems per thread (WD_OPT_F64_PAIRS).

The mixed solvers' reliable updates and full-system checks apply the FP64
operator to several systems in one launch; with pairs on, a thread loads each
link once for two systems.  Per system the arithmetic is the one-system
kernel's, so every solver must give bit-identical solutions and iteration
counts with the option on and off, also for an odd number of systems (the
last pair holds one) and with masked systems (updates fall due at different
iterations for different masses).
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L", [[8, 8, 8, 8], [4, 6, 4, 8]])
def test_pairs_bit_identical(w, L):
    n = 3
    A = w.gauge_near_identity(L, 101, 0.3)
    bs = [w.spinor_gaussian(L, 102, r) for r in range(n)]
    masses = [kappa_to_mass(k) for k in (0.12, 0.13, 0.125)]
    print(f"f64 pairs L={L}", flush=True)
    out = {}
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        assert ctx.get_option(w.OPT_F64_PAIRS) == 1
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        Us32 = ctx.gauge(A, prec=w.F32, recon=12)
        fb = ctx.field(nrhs=n, data=bs)
        for pairs in (1, 0):
            ctx.set_option(w.OPT_F64_PAIRS, pairs)
            assert ctx.get_option(w.OPT_F64_PAIRS) == pairs
            for name, run in (
                    ("cg_mixed eo", lambda X: ctx.cg_mixed(X, fb, U, Us, masses, even_odd=True,
                                                           check_every=4)),
                    ("cg_mixed full", lambda X: ctx.cg_mixed(X, fb, U, Us32, masses,
                                                             check_every=4)),
                    ("bicgstab", lambda X: ctx.bicgstab_mixed(X, fb, U, Us, masses,
                                                              check_every=4))):
                X = ctx.field(nrhs=n)
                res = run(X)
                assert res["status"] == 0, name
                out[(pairs, name)] = ([s["iterations"] for s in res["systems"]],
                                      [X.download(r) for r in range(n)])
                X.close()
    for name in ("cg_mixed eo", "cg_mixed full", "bicgstab"):
        it1, x1 = out[(1, name)]
        it0, x0 = out[(0, name)]
        assert it1 == it0, name
        for r in range(n):
            assert np.array_equal(x1[r], x0[r]), (name, r)
