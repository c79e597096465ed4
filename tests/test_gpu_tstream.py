"""t-streaming FP64 kernel for the both-parity M / M^dagger launches.

Each thread walks a run of T sites along t of one (x, y, z) column, so the
code path differs from the site-block kernel in every index computation, but
each site is accumulated from the same inputs in the same direction order:
the operator must agree bit for bit with the site-block kernel for every run
length, including shapes whose column count is not a multiple of the
workgroup (idle lanes) and boundary signs in every direction.  The CG's
M^dagger with its fused dot product groups the partial sums differently, so
CG iterates agree to rounding and iteration counts are equal.  The default
(automatic) choice is checked at 32^3x64, where it is on.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, rel_l2

pytestmark = pytest.mark.gpu

SHAPES = [([8, 8, 8, 16], [16, 8, 4, 2]),   # 256 columns per type: one full workgroup
          ([6, 4, 10, 8], [8, 4, 2]),       # 120 columns: idle lanes
          ([4, 6, 6, 12], [12, 6, 4, 3, 2]),  # 72 columns, odd run length
          ([10, 6, 4, 8], [8, 2])]


def _apply_both(w, L, A, b, tstream):
    with w.Context(L, bc=[-1, -1, 1, -1]) as ctx:
        ctx.set_option(w.OPT_TSTREAM, tstream)
        U = ctx.gauge(A)
        fin, fout = ctx.field(data=b), ctx.field()
        out = []
        for dag in (False, True):
            ctx.apply_M(fout, fin, U, 0.07, dag)
            out.append(fout.download())
    return out


@pytest.mark.parametrize("L,runs", SHAPES)
def test_tstream_operator_bitwise(w, oracle, L, runs):
    A = w.gauge_random_su3(L, 91)
    b = w.spinor_gaussian(L, 92)
    ref = _apply_both(w, L, A, b, 0)
    for T in runs:
        print(f"L={L} T={T}", flush=True)
        got = _apply_both(w, L, A, b, T)
        for dag in (0, 1):
            assert np.array_equal(got[dag], ref[dag]), (L, T, dag)
    # and the operator is the oracle's (links flipped for the antiperiodic x, y, t)
    Af = [a.copy() for a in A]
    for mu in (0, 1, 3):
        w.gauge_flip_boundary(L, mu, Af)
    for dag in (0, 1):
        assert rel_l2(ref[dag], oracle.dslash(b, Af, L, 0.07, dag)) < 1e-13


def test_tstream_cg_agrees(w):
    L = [8, 8, 8, 16]
    A = w.gauge_near_identity(L, 93, 0.3)
    b = w.spinor_gaussian(L, 94)
    res, xs = [], []
    for T in (0, 4, 16):
        with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
            ctx.set_option(w.OPT_TSTREAM, T)
            U = ctx.gauge(A)
            x = ctx.field()
            r = ctx.cg(x, ctx.field(data=b), U, kappa_to_mass(0.13), max_iter=2000, tol=1e-10,
                       tol_relative=True, use_x0=False, check_every=8)
            res.append(r)
            xs.append(x.download())
    print([r["iterations"] for r in res], flush=True)
    for k in (1, 2):
        assert res[k]["iterations"] == res[0]["iterations"]
        assert rel_l2(xs[k], xs[0]) < 1e-12
        assert res[k]["true_rsd"] <= 1e-10 * res[k]["rhs_norm"]


def test_tstream_default_at_32x32x32x64(w):
    """Production shape: the automatic choice (runs of 16, 512 workgroups,
    XCD remap) against the site-block kernel, bit for bit."""
    L = [32, 32, 32, 64]
    A = w.gauge_near_identity(L, 95, 0.3)
    b = w.spinor_gaussian(L, 96)
    outs = []
    for T in (1, 0):
        with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
            ctx.set_option(w.OPT_TSTREAM, T)
            U = ctx.gauge(A)
            fin, fout = ctx.field(data=b), ctx.field()
            ctx.apply_M(fout, fin, U, 0.1, False)
            outs.append(fout.download())
            ms = ctx.time_op(w.OP_M, fout, fin, U, 0.1, False, 0, 10)
            print(f"T option {T}: M {ms:.4f} ms", flush=True)
    assert np.array_equal(outs[0], outs[1])


def test_tstream_option_errors(w):
    with w.Context([4, 4, 4, 12]) as ctx:
        with pytest.raises(w.WilsonError):
            ctx.set_option(w.OPT_TSTREAM, -1)
        with pytest.raises(w.WilsonError):
            ctx.set_option(w.OPT_TSTREAM, 8)  # 12 is not a multiple of 8
        ctx.set_option(w.OPT_TSTREAM, 6)
        assert ctx.get_option(w.OPT_TSTREAM) == 6
        ctx.set_option(w.OPT_TSTREAM, 1)
        assert ctx.get_option(w.OPT_TSTREAM) == 0  # automatic: Lt = 12 is off
        ctx.set_option(w.OPT_TSTREAM, 0)
        assert ctx.get_option(w.OPT_TSTREAM) == 0
    with w.Context([32, 32, 32, 64]) as ctx:
        assert ctx.get_option(w.OPT_TSTREAM) == 16  # 512 workgroups: on by default
    with w.Context([24, 24, 24, 48]) as ctx:
        assert ctx.get_option(w.OPT_TSTREAM) == 0   # 162 workgroups: site-block kernel
