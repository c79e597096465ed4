#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Runs in the development container only, where the reference sources exist and
oracle/Makefile builds them into oracle/_ref/libref.so (`make -C oracle ref`).
The reference's own DslashoffdNew / Dslashoffd / DslashNew / Dslash /
DslashEEOONew / CGinvert produce the expected outputs on inputs from the
seeded generators of libwilson.so.  Small inputs are stored in the fixture;
large ones are regenerated from their seeds and checked against the stored
SHA-256 digest.  Antiperiodic time is fed to the periodic-only reference by
negating U_t on the last time slice.

Usage: python tests/golden/make_golden.py [--cg24]
"""
import argparse
import ctypes
import hashlib
import json
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402
sys.path.insert(0, HERE)

w = ge.load_package().wilson_hip
REF = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref.so"))


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _L(L):
    return (ctypes.c_int * 4)(*L)


def _U(A):
    return (ctypes.c_void_p * 4)(*[a.ctypes.data for a in A])


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def ref_hop(which, src, A, L, dag, cb):
    out = np.empty_like(src)
    REF.ref_hop(which, _p(src), _p(out), _U(A), _L(L), dag, cb)
    return out


def ref_dslash(which, src, A, L, mass, dag):
    out = np.empty_like(src)
    REF.ref_dslash(which, _p(src), _p(out), _U(A), _L(L), ctypes.c_double(mass), dag)
    return out


def ref_diag(src, L, mass):
    out = np.empty_like(src)
    REF.ref_diag(_p(src), _p(out), _L(L), ctypes.c_double(mass))
    return out


def ref_cg(src, A, L, mass, max_iter, acc, seed):
    out = np.empty_like(src)
    log = ctypes.create_string_buffer(1 << 22)
    REF.ref_cg(_p(src), _p(out), _U(A), _L(L), ctypes.c_double(mass), max_iter,
               ctypes.c_double(acc), ctypes.c_uint(seed), log, len(log))
    text = log.value.decode()
    hist = [float(m.group(2)) for m in re.finditer(r"CG: (\d+) iter, rsd \|r\| = (\S+)", text)]
    conv = re.search(r"CG: (\d+) iterations, convergence residual \|r\| = (\S+)", text)
    iters = int(conv.group(1)) if conv else -1
    if conv:
        hist.append(float(conv.group(2)))
    return out, iters, np.array(hist)


def operator_fixture(name, L, gauge, mass, seed_g, seed_s, flip_t):
    if gauge == "su3":
        A = w.gauge_random_su3(L, seed_g)
    else:
        A = w.gauge_near_identity(L, seed_g, 0.3)
    if flip_t:
        w.gauge_flip_boundary(L, 3, A)
    b = w.spinor_gaussian(L, seed_s)
    out = {"L": np.array(L), "mass": mass, "gauge": np.stack(A), "src": b}
    for dag in (0, 1):
        for cb in (0, 1):
            out[f"hopnew_d{dag}_cb{cb}"] = ref_hop(0, b, A, L, dag, cb)
            out[f"hop_d{dag}_cb{cb}"] = ref_hop(1, b, A, L, dag, cb)
        out[f"Mnew_d{dag}"] = ref_dslash(0, b, A, L, mass, dag)
        out[f"M_d{dag}"] = ref_dslash(1, b, A, L, mass, dag)
    out["diag"] = ref_diag(b, L, mass)
    # reductions of the reference (norm_2, vector_p)
    b2 = w.spinor_gaussian(L, seed_s, 1)
    out["src2"] = b2
    out["norm2"] = REF.ref_norm2(_p(b), _L(L))
    d = np.zeros(2)
    REF.ref_dot(_p(b), _p(b2), _L(L), _p(d))
    out["dot"] = d
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    return {"file": name + ".npz", "lattice": L, "gauge": gauge, "gauge_seed": seed_g,
            "src_seed": seed_s, "mass": mass, "antiperiodic_t_by_link_flip": flip_t,
            "inputs_sha256": sha(*A, b)}


def cg_fixture(name, L, gauge, kappa, seed_g, seed_s, x0_seed, flip_t, max_iter, tol_rel):
    A = (w.gauge_random_su3(L, seed_g) if gauge == "su3"
         else w.gauge_near_identity(L, seed_g, 0.3))
    if flip_t:
        w.gauge_flip_boundary(L, 3, A)
    b = w.spinor_gaussian(L, seed_s)
    mass = 1.0 / (2 * kappa) - 4.0
    Mdb = ref_dslash(0, b, A, L, mass, 1)
    acc = tol_rel * float(np.linalg.norm(Mdb))
    t0 = time.time()
    x, iters, hist = ref_cg(b, A, L, mass, max_iter, acc, x0_seed)
    el = time.time() - t0
    rng = np.random.default_rng(7)
    probes = rng.standard_normal((4, x.size)) + 1j * rng.standard_normal((4, x.size))
    V = x.size // 12
    sample = np.linspace(0, V - 1, 64).astype(np.int64)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), L=np.array(L), kappa=kappa, mass=mass,
                        acc=acc, iterations=iters, history=hist, x_norm=np.linalg.norm(x),
                        x_probe=probes.conj() @ x, x_sites=x.reshape(V, 12)[sample],
                        sample_sites=sample, probe_seed=7)
    return {"file": name + ".npz", "lattice": L, "gauge": gauge, "gauge_seed": seed_g,
            "src_seed": seed_s, "kappa": kappa, "mass": mass, "x0": f"srand({x0_seed}); rand()",
            "acc_abs": acc, "tol_rel": tol_rel, "max_iter": max_iter, "iterations": iters,
            "antiperiodic_t_by_link_flip": flip_t, "inputs_sha256": sha(*A, b),
            "reference_seconds_1core": round(el, 2)}


def main_flow_fixture(name, L, mass, gauge_seed, seed):
    """The reference's driver steps (src/main.cpp:77-117): rand() source, load_gauge
    ("UNIT" or a file written from the seeded near-identity generator), CGinvert with
    MaxCG 500 / Accuracy 1e-9, check().  The printed text is the fixture."""
    import tempfile
    from gaugefile import write_gauge_file
    meta = {"file": name + ".txt", "lattice": L, "mass": mass, "seed": seed}
    with tempfile.TemporaryDirectory() as td:
        if gauge_seed is None:
            gauge = "UNIT"
        else:
            A = w.gauge_near_identity(L, gauge_seed, 0.3)
            gauge = write_gauge_file(os.path.join(td, "gauge.bin"), A, L)
            meta["gauge_seed"] = gauge_seed
            meta["gauge_file_sha256"] = hashlib.sha256(open(gauge, "rb").read()).hexdigest()
        log = ctypes.create_string_buffer(1 << 22)
        REF.ref_main_flow(_L(L), ctypes.c_double(mass), gauge.encode(), ctypes.c_uint(seed),
                          log, len(log))
    open(os.path.join(HERE, name + ".txt"), "w").write(log.value.decode())
    meta["gauge"] = "UNIT" if gauge_seed is None else "file"
    return meta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cg24", action="store_true", help="also the 24^3x48 CG fixture (slow)")
    args = ap.parse_args()
    REF.ref_mpi_size()
    REF.ref_norm2.restype = ctypes.c_double
    manifest_path = os.path.join(HERE, "manifest.json")
    manifest = json.load(open(manifest_path)) if os.path.exists(manifest_path) else {}
    manifest["generator"] = "tests/golden/make_golden.py"
    manifest["reference_build"] = ("oracle/Makefile `ref` target: g++ -std=c++11 -Ofast "
                                   "-ffast-math -march=core-avx2 -fomit-frame-pointer "
                                   "-DNO_MPI_IO, MPICH singleton, 1 rank")
    fx = manifest.setdefault("fixtures", {})
    fx["op_4x4x4x4_su3"] = operator_fixture("op_4x4x4x4_su3", [4, 4, 4, 4], "su3", 0.1,
                                            20211025, 20211026, False)
    fx["op_2x4x6x4_near_ap"] = operator_fixture("op_2x4x6x4_near_ap", [2, 4, 6, 4], "near",
                                                -0.15384615384615419, 11, 12, True)
    fx["cg_4x4x4x8_near"] = cg_fixture("cg_4x4x4x8_near", [4, 4, 4, 8], "near", 0.12, 31, 32,
                                       1234, False, 500, 1e-10)
    fx["cg_8x8x8x8_su3_ap"] = cg_fixture("cg_8x8x8x8_su3_ap", [8, 8, 8, 8], "su3", 0.12,
                                         20211025, 20211026, 20211027, True, 1000, 1e-10)
    fx["main_8x8x8x8_unit"] = main_flow_fixture("main_8x8x8x8_unit", [8, 8, 8, 8], 0.05, None, 0)
    fx["main_8x8x8x8_file"] = main_flow_fixture("main_8x8x8x8_file", [8, 8, 8, 8], 0.1, 41, 3)
    if args.cg24:
        fx["cg_24x24x24x48_near_ap"] = cg_fixture(
            "cg_24x24x24x48_near_ap", [24, 24, 24, 48], "near", 0.13, 20211025, 20211026,
            20211027, True, 20000, 1e-10)
    json.dump(manifest, open(manifest_path, "w"), indent=1, sort_keys=True)
    print(json.dumps(fx, indent=1))


if __name__ == "__main__":
    main()
