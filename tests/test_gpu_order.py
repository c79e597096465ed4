"""Work order of the Dslash kernels (natural vs z-t tiled site blocks).

The tiled order is a permutation of the workgroups: every site is computed by
the same code from the same inputs and reductions are summed in the natural
block order, so every tile width must give bit-identical operators, CG
iterates and mixed-precision iterates.  16x16x8x4: a time slice of one parity
holds 8*16*8 = 1024 sites = 4 blocks, so widths 1 and 3 (a narrower last
column) are both real tilings.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, rel_l2

pytestmark = pytest.mark.gpu

L = [16, 16, 8, 4]
WIDTHS = [0, 1, 3]


def _run_fp64(w, width, A, b, parity_order=0):
    out = {}
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        ctx.set_option(w.OPT_BLOCK_TILE, width)
        ctx.set_option(w.OPT_PARITY_ORDER, parity_order)
        U = ctx.gauge(A)
        fin, fout = ctx.field(data=b), ctx.field()
        for dag in (False, True):
            ctx.apply_M(fout, fin, U, 0.07, dag)
            out[f"M{int(dag)}"] = fout.download()
            for cb in (0, 1):
                ctx.hop(fout, fin, U, dag, cb)
                out[f"hop{int(dag)}{cb}"] = fout.download()
        x = ctx.field()
        res = ctx.cg(x, fin, U, kappa_to_mass(0.12), max_iter=1000, tol=1e-10,
                     tol_relative=True, use_x0=False, check_every=4)
        out["cg_x"] = x.download()
        out["cg_it"] = res["iterations"]
        out["cg_hist"] = np.array(res["history"])
    return out


def test_tiled_order_fp64_bitwise(w, oracle):
    A = w.gauge_random_su3(L, 77)
    b = w.spinor_gaussian(L, 78)
    runs = {W: _run_fp64(w, W, A, b) for W in WIDTHS}
    for W in WIDTHS[1:]:
        print(f"width {W}", flush=True)
        for k, v in runs[0].items():
            assert np.array_equal(np.asarray(runs[W][k]), np.asarray(v)), (W, k)
    # and the tiled result is the oracle's (links flipped for antiperiodic t)
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for dag in (0, 1):
        assert rel_l2(runs[3][f"M{dag}"], oracle.dslash(b, Af, L, 0.07, dag)) < 1e-13


def test_parity_order_fp64_bitwise(w):
    """Both parities side by side (shared links through L2) or one after the
    other: the same values bit for bit, with and without the tiled order."""
    A = w.gauge_random_su3(L, 87)
    b = w.spinor_gaussian(L, 88)
    ref = _run_fp64(w, 0, A, b, 0)
    for width, order in ((0, 1), (3, 1), (1, 0)):
        got = _run_fp64(w, width, A, b, order)
        for k, v in ref.items():
            assert np.array_equal(np.asarray(got[k]), np.asarray(v)), (width, order, k)


def test_tiled_order_sloppy_bitwise(w):
    A = w.gauge_near_identity(L, 79, 0.3)
    n = 5
    bs = [w.spinor_gaussian(L, 80, r) for r in range(n)]
    masses = [0.05 * r for r in range(n)]
    res = {}
    for W, order in [(W, 0) for W in WIDTHS] + [(0, 1), (3, 1)]:
        with w.Context(L) as ctx:
            ctx.set_option(w.OPT_BLOCK_TILE, W)
            ctx.set_option(w.OPT_PARITY_ORDER, order)
            Us = ctx.gauge(A, prec=w.F32, recon=12)
            U = ctx.gauge(A)
            fin = ctx.field(prec=w.F16, nrhs=n, data=bs)
            fout = ctx.field(prec=w.F16, nrhs=n)
            ctx.apply_M_sloppy(fout, fin, Us, masses, False, 2, 1)
            o = [fout.download(r) for r in range(n)]
            fb = ctx.field(nrhs=n, data=bs)
            x = ctx.field(nrhs=n)
            mx = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=2, site_chunk=0,
                              check_every=4)
            res[(W, order)] = (o, [x.download(r) for r in range(n)],
                               [s["iterations"] for s in mx["systems"]])
    base = res[(0, 0)]
    for key, got in res.items():
        for r in range(n):
            assert np.array_equal(got[0][r], base[0][r]), (key, r)
            assert np.array_equal(got[1][r], base[1][r]), (key, r)
        assert got[2] == base[2]


def test_block_tile_option_errors(w):
    with w.Context([4, 4, 4, 4]) as ctx:
        with pytest.raises(w.WilsonError):
            ctx.set_option(w.OPT_BLOCK_TILE, -1)
        ctx.set_option(w.OPT_BLOCK_TILE, 16)  # slices not block aligned: natural order
