"""The reference-style C++ API (libipcc_api.so) end to end: api_selftest calls
DslashoffdNew, Dslashoffd, DslashEONew/DslashOENew, DslashNew, Dslash,
DslashEEOONew, DslashEE, DslashOO and CGinvert on the inputs of a golden
fixture; the outputs must equal the reference's own outputs (1e-13) and the
solution must satisfy the normal equations."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle_py import Oracle, libc_rand_x0, rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
EXE = os.path.join(ROOT, "ipcc2021-final_amd", "api_selftest")


@pytest.mark.parametrize("name", ["op_4x4x4x4_su3", "op_2x4x6x4_near_ap"])
def test_cpp_api_matches_reference(name, tmp_path):
    f = np.load(os.path.join(GOLD, MANIFEST["fixtures"][name]["file"]))
    L = [int(v) for v in f["L"]]
    mass = float(f["mass"])
    f["src"].astype(np.complex128).tofile(tmp_path / "src.bin")
    np.ascontiguousarray(f["gauge"], dtype=np.complex128).tofile(tmp_path / "gauge.bin")
    p = subprocess.run([EXE, str(tmp_path), *map(str, L), repr(mass)], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    out = lambda n: np.fromfile(tmp_path / f"{n}.bin", dtype=np.complex128)
    V = int(np.prod(L))
    for dag in (0, 1):
        for cb in (0, 1):
            assert rel_l2(out(f"hopnew_d{dag}_cb{cb}"), f[f"hopnew_d{dag}_cb{cb}"]) < 1e-13
            assert rel_l2(out(f"hop_d{dag}_cb{cb}"), f[f"hop_d{dag}_cb{cb}"]) < 1e-13
        # DslashEONew = cb 0, DslashOENew = cb 1 (declared, never defined in the reference)
        cb = 0 if dag == 0 else 1
        assert rel_l2(out(f"eo_new_d{dag}"), f[f"hopnew_d{dag}_cb{cb}"]) < 1e-13
        assert rel_l2(out(f"Mnew_d{dag}"), f[f"Mnew_d{dag}"]) < 1e-13
        assert rel_l2(out(f"M_d{dag}"), f[f"M_d{dag}"]) < 1e-13
    diag = f["diag"].reshape(2, -1)
    assert rel_l2(out("diag"), f["diag"]) < 1e-15
    ee, oo = out("diag_ee").reshape(2, -1), out("diag_oo").reshape(2, -1)
    assert rel_l2(ee[0], diag[0]) < 1e-15 and np.all(ee[1] == 0)
    assert rel_l2(oo[1], diag[1]) < 1e-15 and np.all(oo[0] == 0)
    # links kept on the device by gauge_resident: the same outputs bit for bit
    for dag in (0, 1):
        assert np.array_equal(out(f"Mnew_res_d{dag}"), out(f"Mnew_d{dag}"))
        assert np.array_equal(out(f"hopnew_res_d{dag}"), out(f"hopnew_d{dag}_cb1"))
    # after the guard, changed links are read again (zero links: M = 4 + m)
    assert rel_l2(out("Mnew_zero"), (4.0 + mass) * f["src"]) < 1e-15
    # CGinvert: x0 from srand(11), |r| < 1e-9 absolute; check with the FP64 oracle
    o = Oracle()
    A = [np.ascontiguousarray(a) for a in f["gauge"]]
    x = out("cg_x")
    b = f["src"]
    r = o.dslash(b, A, L, mass, 1) - o.dslash(o.dslash(x, A, L, mass, 0), A, L, mass, 1)
    assert np.linalg.norm(r) < 1.5e-9
    it, xo, _ = o.cg(b, libc_rand_x0(11, V * 12), A, L, mass, 1000, 1e-9)
    assert it > 0 and rel_l2(x, xo) < 1e-8


def test_gauge_resident_removes_the_link_upload():
    """DslashNew on the drop-in API at 32^3x64: with a gauge_resident guard the
    1.2 GB of links are not uploaded on every call."""
    p = subprocess.run([EXE, "--time", "32", "32", "32", "64", "5"], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    print(p.stdout.strip(), flush=True)
    m = [float(v) for v in
         __import__("re").findall(r"([0-9.]+) ms", p.stdout)]
    assert len(m) == 2 and m[1] < m[0]
