"""CPU-side checks of the product libraries: the C-ABI loads and exports every
symbol include/*.h declares, host-side generators and argument checks behave,
the reference-style C++ API exports the reference's names.  No GPU compute."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ipcc2021-final_amd")


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wdg?_[A-Za-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", ["wilson.h", "wilson_gen.h"])
def test_cabi_exports_every_declared_symbol(w, header):
    lib = w.load_library()
    names = _declared(header)
    assert names
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/{header} but not exported"
    assert set(names) <= set(w.EXPORTED)


def test_cabi_exports_nothing_else():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG, "libwilson.so")],
                         capture_output=True, text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if " T " in l}
    declared = set(_declared("wilson.h")) | set(_declared("wilson_gen.h"))
    assert syms == declared


def test_create_rejects_bad_shapes_without_touching_a_device(w):
    lib = w.load_library()
    h = ctypes.c_void_p()
    for L in ([3, 4, 4, 4], [4, 4, 4, 0], [4, 4, 4, 5]):
        assert lib.wd_create(ctypes.byref(h), (ctypes.c_int * 4)(*L), None, 0) == w.WD_EARG
    bc = (ctypes.c_int * 4)(1, 1, 1, 2)
    assert lib.wd_create(ctypes.byref(h), (ctypes.c_int * 4)(4, 4, 4, 4), bc, 0) == w.WD_EARG
    assert lib.wd_strerror(w.WD_EARG) == b"invalid argument"
    assert lib.wd_strerror(w.WD_ENOTCONV) == b"not converged"


def test_cg_params_default_are_the_reference_settings(w):
    lib = w.load_library()
    p = w.CGParams()
    lib.wd_cg_params_default(ctypes.byref(p))
    # src/main.cpp:47-48: MaxCG = 500, Accuracy = 1e-9 (absolute)
    assert p.max_iter == 500 and p.tol == 1e-9 and p.tol_relative == 0 and p.use_x0 == 1


def test_philox_known_answers(w):
    """Known-answer vectors of Philox4x32-10 (Random123 kat_vectors)."""
    assert w.philox4x32([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert w.philox4x32([0xffffffff] * 4, [0xffffffff] * 2) == [
        0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert w.philox4x32([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                        [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420,
                                                      0x24126ea1]


@pytest.mark.parametrize("kind", ["su3", "near"])
def test_generated_links_are_su3(w, kind):
    L = [4, 4, 4, 6]
    A = w.gauge_random_su3(L, 5) if kind == "su3" else w.gauge_near_identity(L, 5, 0.3)
    for a in A:
        U = a.reshape(-1, 3, 3)
        eye = np.einsum("nab,ncb->nac", U, U.conj())
        assert np.abs(eye - np.eye(3)).max() < 1e-14
        assert np.abs(np.linalg.det(U) - 1).max() < 1e-14
    if kind == "near":
        # exp(0.3 i H): close to the identity, Re tr U / 3 ~ 0.89
        tr = np.mean([np.trace(a.reshape(-1, 3, 3), axis1=1, axis2=2).real.mean() / 3 for a in A])
        assert 0.8 < tr < 0.95


def test_generators_are_deterministic_and_volume_independent(w):
    L = [4, 4, 4, 4]
    a1 = w.spinor_gaussian(L, 11, 0)
    assert np.array_equal(a1, w.spinor_gaussian(L, 11, 0))
    assert not np.array_equal(a1, w.spinor_gaussian(L, 11, 1))
    assert not np.array_equal(a1, w.spinor_gaussian(L, 12, 0))
    assert abs(np.std(np.concatenate([a1.real, a1.imag])) - 1) < 0.1
    # the value at a site depends only on (seed, stream, global lex site): a lattice
    # with more time slices holds the same values on the common sites
    L2 = [4, 4, 4, 8]
    a2 = w.spinor_gaussian(L2, 11, 0)
    import dense_wilson as dw
    s1 = dw.spinor_to_lex(a1, L)
    s2 = dw.spinor_to_lex(a2, L2)
    assert np.array_equal(s1, s2[:4])


def test_point_source(w):
    L = [4, 4, 4, 4]
    a = w.spinor_point(L, [1, 2, 3, 0], 2, 1)
    assert np.count_nonzero(a) == 1
    import dense_wilson as dw
    s = dw.spinor_to_lex(a, L)
    assert s[0, 3, 2, 1, 2, 1] == 1.0


def test_reference_api_exports():
    out = subprocess.run(["nm", "-DC", "--defined-only", os.path.join(PKG, "libipcc_api.so")],
                         capture_output=True, text=True, check=True).stdout
    for sig in ["DslashNew(lattice_fermion&, lattice_fermion&, lattice_gauge&, double, bool)",
                "Dslash(lattice_fermion&, lattice_fermion&, lattice_gauge&, double, bool)",
                "DslashoffdNew(lattice_fermion&, lattice_fermion&, lattice_gauge&, bool, int)",
                "Dslashoffd(lattice_fermion&, lattice_fermion&, lattice_gauge&, bool, int)",
                "DslashEEOONew(lattice_fermion&, lattice_fermion&, double)",
                "DslashEE(lattice_fermion&, lattice_fermion&, double)",
                "DslashOO(lattice_fermion&, lattice_fermion&, double)",
                "DslashEONew(", "DslashOENew(", "DslashEO(", "DslashOE(",
                "CGinvert(lattice_fermion&, lattice_fermion&, lattice_gauge&, double, int, double)",
                "CGinvert(std::complex<double>*, std::complex<double>*, std::complex<double>**, "
                "double, int, double, int*, int*)",
                "check(lattice_fermion&, lattice_fermion&, lattice_gauge&, double)",
                "load_gauge(std::complex<double>**, std::__cxx11::basic_string<char",
                "lattice_fermion::lattice_fermion(int*, int*)",
                "lattice_gauge::lattice_gauge(int*, int*)",
                "Minus(lattice_fermion&, lattice_fermion&, lattice_fermion&)"]:
        assert sig in out, sig


def test_main_usage_and_partition_errors():
    exe = os.path.join(PKG, "wilson_main")
    p = subprocess.run([exe, "1", "2"], capture_output=True, text=True, timeout=30)
    assert p.returncode == 1 and "wilson_main" in p.stderr
    p = subprocess.run([exe, "0.1", "UNIT", "8", "8", "8", "8", "8", "8", "4", "8"],
                       capture_output=True, text=True, timeout=30)
    assert p.returncode == 1 and "partition" in p.stderr


def test_bench_refuses_multi_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr
