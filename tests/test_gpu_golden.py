"""The HIP path against golden outputs of the compiled reference
(tests/golden/, written by tests/golden/make_golden.py from the reference's
DslashoffdNew, DslashNew, DslashEEOONew and CGinvert).

Bars: operator 1e-13 relative L2; CG the same iteration count as the
reference's CGinvert from the same rand() start vector and absolute stop
threshold, residual history equal to the 6 digits the reference prints.
Both FP64 M kernels are held to these bars: the automatic choice (the
site-block kernel at these sizes) and the t-streaming kernel forced on
(the default at 32^3x64 and larger).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle_py import libc_rand_x0, rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
OPS = [k for k in MANIFEST["fixtures"] if k.startswith("op_")]
CGS = [k for k in MANIFEST["fixtures"] if k.startswith("cg_")]


def _load(name):
    return np.load(os.path.join(GOLD, MANIFEST["fixtures"][name]["file"]))


def _tstream_option(kind, Lt):
    """1 = automatic; else t runs forced on (the longest of 16, 8, 4, 2 dividing Lt)."""
    if kind == "auto":
        return 1
    return next(T for T in (16, 8, 4, 2) if Lt % T == 0)


@pytest.mark.parametrize("kind", ["auto", "tstream"])
@pytest.mark.parametrize("name", OPS)
def test_operator_matches_reference_outputs(w, name, kind):
    f = _load(name)
    L = [int(v) for v in f["L"]]
    A = [np.ascontiguousarray(a) for a in f["gauge"]]
    b = f["src"]
    mass = float(f["mass"])
    with w.Context(L) as ctx:
        ctx.set_option(w.OPT_TSTREAM, _tstream_option(kind, L[3]))
        U = ctx.gauge(A)
        fin, fout = ctx.field(data=b), ctx.field()
        for dag in (0, 1):
            for cb in (0, 1):
                ctx.hop(fout, fin, U, bool(dag), cb)
                assert rel_l2(fout.download(), f[f"hopnew_d{dag}_cb{cb}"]) < 1e-13
            ctx.apply_M(fout, fin, U, mass, bool(dag))
            assert rel_l2(fout.download(), f[f"Mnew_d{dag}"]) < 1e-13
        ctx.diag(fout, fin, mass)
        assert rel_l2(fout.download(), f["diag"]) < 1e-15
        f2 = ctx.field(data=f["src2"])
        n2 = ctx.norm2(fin)
        assert abs(n2 - float(f["norm2"])) < 1e-13 * n2
        d = ctx.dot(fin, f2)
        assert abs(d - complex(*f["dot"])) < 1e-13 * abs(d)


@pytest.mark.parametrize("kind", ["auto", "tstream"])
@pytest.mark.parametrize("name", CGS)
def test_cg_matches_reference_iterations(w, name, kind):
    meta = MANIFEST["fixtures"][name]
    f = _load(name)
    L = meta["lattice"]
    print(f"golden cg {name}", flush=True)
    A = (w.gauge_random_su3(L, meta["gauge_seed"]) if meta["gauge"] == "su3"
         else w.gauge_near_identity(L, meta["gauge_seed"], 0.3))
    if meta["antiperiodic_t_by_link_flip"]:
        w.gauge_flip_boundary(L, 3, A)
    b = w.spinor_gaussian(L, meta["src_seed"])
    h = hashlib.sha256()
    for a in (*A, b):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == meta["inputs_sha256"], "generators drifted from the fixture"
    x0_seed = int(meta["x0"].split("(")[1].split(")")[0])
    x0 = libc_rand_x0(x0_seed, w.volume(L) * 12)
    with w.Context(L) as ctx:
        ctx.set_option(w.OPT_TSTREAM, _tstream_option(kind, L[3]))
        U = ctx.gauge(A)
        fb, x = ctx.field(data=b), ctx.field(data=x0)
        res = ctx.cg(x, fb, U, meta["mass"], max_iter=meta["max_iter"], tol=meta["acc_abs"],
                     tol_relative=False, use_x0=True)
        xs = x.download()
        ctx_t = ctx.get_option(w.OPT_TSTREAM)
    print(f"gpu iterations {res['iterations']} reference {int(f['iterations'])} "
          f"(t runs {ctx_t})", flush=True)
    assert res["converged"] and res["iterations"] == int(f["iterations"])
    assert np.allclose(res["history"], f["history"], rtol=1e-5, atol=0)
    assert abs(np.linalg.norm(xs) - float(f["x_norm"])) < 1e-8 * float(f["x_norm"])
    V = w.volume(L)
    assert rel_l2(xs.reshape(V, 12)[f["sample_sites"]], f["x_sites"]) < 1e-7
