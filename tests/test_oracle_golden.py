"""Pins the CPU oracle to golden vectors written by the compiled reference
(tests/golden/make_golden.py: the reference's DslashoffdNew, Dslashoffd,
DslashNew, Dslash, DslashEEOONew, norm_2, vector_p and CGinvert).  CPU only.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle_py import libc_rand_x0, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
OPS = [k for k in MANIFEST["fixtures"] if k.startswith("op_")]
CGS = [k for k in MANIFEST["fixtures"] if k.startswith("cg_")]
FAST_CGS = [k for k in CGS if "24x24" not in k]

TOL = 1e-13


def _load(name):
    return np.load(os.path.join(GOLD, MANIFEST["fixtures"][name]["file"]))


@pytest.mark.parametrize("name", OPS)
def test_oracle_operator_matches_reference(oracle, name):
    f = _load(name)
    L = [int(v) for v in f["L"]]
    A = [np.ascontiguousarray(a) for a in f["gauge"]]
    b = f["src"]
    mass = float(f["mass"])
    for dag in (0, 1):
        for cb in (0, 1):
            mine = oracle.hop(b, A, L, dag, cb)
            assert rel_l2(mine, f[f"hopnew_d{dag}_cb{cb}"]) < TOL
            assert rel_l2(mine, f[f"hop_d{dag}_cb{cb}"]) < TOL
        mine = oracle.dslash(b, A, L, mass, dag)
        assert rel_l2(mine, f[f"Mnew_d{dag}"]) < TOL
        assert rel_l2(mine, f[f"M_d{dag}"]) < TOL
    assert rel_l2(oracle.diag(b, L, mass), f["diag"]) < 1e-15
    n2 = oracle.norm2(b)
    assert abs(n2 - float(f["norm2"])) < 1e-13 * n2
    d = oracle.dot(b, f["src2"])
    assert abs(d - complex(*f["dot"])) < 1e-13 * abs(d)


@pytest.mark.parametrize("name", OPS + CGS)
def test_generators_reproduce_fixture_inputs(w, name):
    meta = MANIFEST["fixtures"][name]
    L = meta["lattice"]
    if w.volume(L) > 70000:
        pytest.skip("large inputs are checked on the GPU side")
    A = (w.gauge_random_su3(L, meta["gauge_seed"]) if meta["gauge"] == "su3"
         else w.gauge_near_identity(L, meta["gauge_seed"], 0.3))
    if meta["antiperiodic_t_by_link_flip"]:
        w.gauge_flip_boundary(L, 3, A)
    b = w.spinor_gaussian(L, meta["src_seed"])
    h = hashlib.sha256()
    for a in (*A, b):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == meta["inputs_sha256"]


def _cg_inputs(w, meta):
    L = meta["lattice"]
    A = (w.gauge_random_su3(L, meta["gauge_seed"]) if meta["gauge"] == "su3"
         else w.gauge_near_identity(L, meta["gauge_seed"], 0.3))
    if meta["antiperiodic_t_by_link_flip"]:
        w.gauge_flip_boundary(L, 3, A)
    b = w.spinor_gaussian(L, meta["src_seed"])
    x0_seed = int(meta["x0"].split("(")[1].split(")")[0])
    x0 = libc_rand_x0(x0_seed, w.volume(L) * 12)
    return L, A, b, x0


@pytest.mark.parametrize("name", FAST_CGS)
def test_oracle_cg_matches_reference(w, oracle, name):
    meta = MANIFEST["fixtures"][name]
    f = _load(name)
    L, A, b, x0 = _cg_inputs(w, meta)
    it, x, hist = oracle.cg(b, x0, A, L, meta["mass"], meta["max_iter"], meta["acc_abs"])
    assert it == int(f["iterations"])
    # the reference prints |r| with 6 significant digits
    assert np.allclose(hist, f["history"], rtol=1e-5, atol=0)
    assert abs(np.linalg.norm(x) - float(f["x_norm"])) < 1e-9 * float(f["x_norm"])
    V = w.volume(L)
    assert rel_l2(x.reshape(V, 12)[f["sample_sites"]], f["x_sites"]) < 1e-8
