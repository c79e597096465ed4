#!/usr/bin/env python3
"""Benchmark of the Wilson-Dslash hot path on one MI355X.

A step is one application of the full-lattice Wilson operator M (the
reference's DslashNew, src/invert.cpp:138-156) to a resident FP64 spinor,
through the fused HIP kernel (both parities, (4+m) diagonal included).
Workload: 32^3x64, near-identity SU(3) links exp(0.3 i H), Gaussian source,
kappa 0.135.  Inputs are synthetic (seeded generators) and resident in HBM
before the timed region.

Printed (one JSON line): GFLOP/s of M (1320 hop + 48 diagonal flop per site),
the roofline of the dominant kernel (algorithmic bytes 8*144 + 3*192 = 1728 B
per site for M, 8*144 + 2*192 = 1536 B per output site for the one-parity hop)
timed with HIP events on the library stream, and a CPU baseline: the
reference's own DslashNew (oracle/_ref, compiled from the reference sources)
on one host core over a bounded sample, or the oracle restatement when the
reference build is absent.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FLOP_HOP = 1320            # per output site, one hop
FLOP_M = 1320 + 48         # per site, hop + (4+m) diagonal (fused FMA)
BYTES_HOP_F64 = 8 * 144 + 2 * 192
BYTES_M_F64 = 8 * 144 + 3 * 192


def load_pkg():
    import __graft_entry__ as ge
    return ge.load_package()


def cpu_baseline_worker(L, kind, seconds):
    """Runs in a child process: times DslashNew on the host (1 core for the reference)."""
    import numpy as np
    V = int(np.prod(L))
    rng = np.random.default_rng(20211026)
    src = rng.standard_normal(V * 24)
    U = [rng.standard_normal(V * 18) for _ in range(4)]
    Up = (ctypes.c_void_p * 4)(*[u.ctypes.data for u in U])
    Lc = (ctypes.c_int * 4)(*L)
    out = np.zeros(V * 24)
    mass = ctypes.c_double(1.0 / (2 * 0.135) - 4.0)
    if kind == "reference":
        lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
        lib.ref_mpi_size()
        call = lambda: lib.ref_dslash(0, ctypes.c_void_p(src.ctypes.data),
                                      ctypes.c_void_p(out.ctypes.data), Up, Lc, mass, 0)
        cores = 1
    else:
        lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        lib.or_dslash.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                  ctypes.POINTER(ctypes.c_int), ctypes.c_double, ctypes.c_int]
        call = lambda: lib.or_dslash(ctypes.c_void_p(src.ctypes.data),
                                     ctypes.c_void_p(out.ctypes.data), Up, Lc, mass, 0)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    call()  # first touch
    n, t0 = 0, time.perf_counter()
    while True:
        call()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    print(json.dumps({"value": FLOP_M * V * n / el / 1e9, "unit": "GFLOP/s", "cores": cores,
                      "kind": kind, "sample": f"{n} DslashNew applications at "
                      f"{'x'.join(map(str, L))}, {el:.1f} s"}), flush=True)


def cpu_baseline(L, seconds):
    have_ref = os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
    kinds = (["reference"] if have_ref else []) + ["port"]
    for kind in kinds:
        env = dict(os.environ)
        if kind == "port":
            env["OMP_NUM_THREADS"] = env.get("OMP_NUM_THREADS", "16")
        try:
            p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", kind,
                                "--cpu-seconds", str(seconds), "--lattice", *map(str, L)],
                               capture_output=True, text=True, timeout=seconds * 4 + 120, env=env)
            if p.returncode == 0 and p.stdout.strip():
                return json.loads(p.stdout.strip().splitlines()[-1])
            print(f"cpu baseline ({kind}) failed: {p.stderr[-500:]}", file=sys.stderr)
        except subprocess.TimeoutExpired:
            print(f"cpu baseline ({kind}) timed out", file=sys.stderr)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--lattice", type=int, nargs=4, default=[32, 32, 32, 64])
    ap.add_argument("--kappa", type=float, default=0.135)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-worker", default=None)
    args = ap.parse_args()

    if args.cpu_worker:
        cpu_baseline_worker(args.lattice, args.cpu_worker, args.cpu_seconds)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs (--gpus > 1) are not "
              "part of it", file=sys.stderr)
        sys.exit(2)

    pkg = load_pkg()
    w = pkg.wilson_hip
    L = args.lattice
    V = w.volume(L)
    mass = 1.0 / (2 * args.kappa) - 4.0
    print(f"bench: generating {'x'.join(map(str, L))} inputs", file=sys.stderr, flush=True)
    A = w.gauge_near_identity(L, 20211025, 0.3)
    b = w.spinor_gaussian(L, 20211026)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        del A
        fin = ctx.field(data=b)
        fout = ctx.field()
        for _ in range(args.warmup):
            ctx.apply_M(fout, fin, U, mass, False)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.apply_M(fout, fin, U, mass, False)
        ctx.sync()
        el = time.perf_counter() - t0
        ms_step = el * 1e3 / args.steps
        # dominant kernel, HIP events on the library stream
        ms_m = ctx.time_op(w.OP_M, fout, fin, U, mass, False, 0, max(args.steps, 10))
        ms_hop = ctx.time_op(w.OP_HOP, fout, fin, U, mass, False, 0, max(args.steps, 10))
    gflops = FLOP_M * V / (ms_step * 1e-3) / 1e9
    gbs_m = BYTES_M_F64 * V / (ms_m * 1e-3) / 1e9
    gbs_hop = BYTES_HOP_F64 * (V // 2) / (ms_hop * 1e-3) / 1e9
    line = {
        "metric": "Dslash GB/s & GFLOP/s vs HBM roofline; CG solves/sec at 1e-10, mixed vs FP64",
        "value": round(gflops, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded near-identity SU(3) links, Gaussian spinor)",
        "config": {"workload": f"wilson_M_fp64_{'x'.join(map(str, L))}", "lattice": L,
                   "kappa": args.kappa, "bc": "periodic xyz, antiperiodic t",
                   "step": "one full-lattice M application (fused diag + hop, both parities)"},
        "roofline": {"bound": "hbm", "achieved": round(gbs_m, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs_m / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_hop_f64<HOP_M>", "bytes_per_site": BYTES_M_F64,
                     "ms_per_launch": round(ms_m, 4)},
        "hop_roofline": {"achieved": round(gbs_hop, 1), "frac": round(gbs_hop / HBM_PEAK_GBS, 4),
                         "unit": "GB/s", "bytes_per_site": BYTES_HOP_F64,
                         "ms_per_launch": round(ms_hop, 4),
                         "gflops": round(FLOP_HOP * (V // 2) / (ms_hop * 1e-3) / 1e9, 1)},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(L, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
