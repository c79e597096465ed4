#!/usr/bin/env python3
"""Benchmark of the Wilson-Dslash hot path on one MI355X.

A step is one application of the full-lattice Wilson operator M (the
reference's DslashNew, src/invert.cpp:138-156) to a resident FP64 spinor
through the fused HIP kernel (both parities, (4+m) diagonal included).
Workload: 32^3x64, near-identity SU(3) links exp(0.3 i H), Gaussian source,
kappa 0.135, periodic x/y/z and antiperiodic t.  Inputs are synthetic (seeded
generators) and resident in HBM before the timed region.

The JSON line carries
  value          GFLOP/s of M, 1320 (hop) + 48 (diagonal) flop per site
  roofline       the step's kernel (k_hop_f64, HOP_M): algorithmic bytes
                 8*144 + 3*192 = 1728 per site / average launch time from HIP
                 events on the library stream; traffic from the committed
                 rocprofv3 FETCH_SIZE/WRITE_SIZE summary when present
  hop_roofline   one-parity FP64 hop (the reference's DslashoffdNew):
                 1536 B and 1320 flop per output site
  fp64_recon12   the FP64 M and hop on 12-real FP64 links (SU(3) links,
                 row 2 rebuilt in registers), beside the 18-real step
  sloppy_roofline  FP32-arithmetic M on F16 spinors and F32 12-real links:
                 8*48 + 3*52 = 540 B per site
  sloppy_temporal_gauge  the same on the links in temporal gauge (the mixed
                 solvers' links here): t links only on the boundary slices,
                 6*48 + 2*48/Lt + 3*52 B per site
  cg             time to a 1e-10 relative residual of M^dag M x = M^dag b:
                 FP64 CG (the reference's algorithm) vs the mixed solver
  cpu_baseline   the reference's own DslashNew (compiled from its sources,
                 oracle/_ref/ref_bench) on as many MPI ranks as the host's
                 cores allow (its one-rank-per-core decomposition), over a
                 bounded sample; the OpenMP oracle restatement on all cores
                 when MPICH or that build is absent
  batched_32x32x32x64  12 Gaussian systems at kappa 0.135 in one batched
                 mixed solve against 12 FP64 CG solves one after another
  batched        24^3x48, 12 point sources x 4 kappas = 48 systems
  large          48^3x96, 12 Gaussian systems, F16 spinors and F16 links
The mixed solvers run on sloppy links in temporal gauge (WD_RECON_12T).
--workload single runs only the single-system part.
"""
import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FLOP_HOP = 1320            # per output site, one hop
FLOP_M = 1320 + 48         # per site, hop + (4+m) diagonal (fused FMA)
BYTES_HOP_F64 = 8 * 144 + 2 * 192
BYTES_M_F64 = 8 * 144 + 3 * 192
BYTES_M_SLOPPY = 8 * 48 + 3 * 52   # F32 12-real links, F16 spinors + float scale
# the same links in temporal gauge: the t links are loaded on the boundary
# slices only (2 of 8 links on 1 of Lt slices each)
def bytes_m_sloppy_tg(Lt):
    return 6 * 48 + 2 * 48 / Lt + 3 * 52


TG_LABEL = " in temporal gauge"
PMC_SUMMARY = {0: os.path.join(ROOT, "profiles", "k_hop_f64_M_32x32x32x64_pmc.txt"),
               16: os.path.join(ROOT, "profiles", "k_hop_f64_ts_M_32x32x32x64_pmc.txt")}


def load_pkg():
    import __graft_entry__ as ge
    return ge.load_package()


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def kappa_mass(k):
    return 1.0 / (2.0 * k) - 4.0


# ---- CPU baseline (child process) -------------------------------------------------
def cpu_baseline_worker(L, kind, seconds):
    V = int(np.prod(L))
    rng = np.random.default_rng(20211026)
    src = rng.standard_normal(V * 24)
    U = [rng.standard_normal(V * 18) for _ in range(4)]
    Up = (ctypes.c_void_p * 4)(*[u.ctypes.data for u in U])
    Lc = (ctypes.c_int * 4)(*L)
    out = np.zeros(V * 24)
    mass = ctypes.c_double(kappa_mass(0.135))
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
    print(json.dumps({"value": round(FLOP_M * V * n / el / 1e9, 3), "unit": "GFLOP/s",
                      "cores": cores, "kind": kind,
                      "sample": f"{n} DslashNew applications (full-lattice M, FP64) at "
                                f"{'x'.join(map(str, L))} in {el:.1f} s"}), flush=True)


MPIEXEC = "/opt/conda/bin/mpiexec"
REF_BENCH = os.path.join(ROOT, "oracle", "_ref", "ref_bench")


def host_cores():
    """Cores this process may use (the box's CPU share), at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def rank_grid(L, nranks):
    """The reference's decomposition for up to nranks ranks: x whole (its
    parity bookkeeping assumes an unsplit x, src/dslash_new.cpp:1407-1409),
    y, z, t split so every sub-lattice extent stays even; the largest rank
    count that fits, and among those the smallest halo (faces per site)."""
    best = None
    for gt in range(1, nranks + 1):
        for gz in range(1, nranks // gt + 1):
            for gy in range(1, nranks // (gt * gz) + 1):
                g = (1, gy, gz, gt)
                if any(L[i] % g[i] or (L[i] // g[i]) % 2 for i in range(4)):
                    continue
                halo = sum(g[i] / L[i] for i in range(1, 4) if g[i] > 1)
                key = (gy * gz * gt, -halo, gt)
                if best is None or key > best[0]:
                    best = (key, g)
    return list(best[1])


def cpu_baseline_mpi(L, seconds):
    n = host_cores()
    g = rank_grid(L, n)
    cmd = [MPIEXEC, "-n", str(int(np.prod(g))), REF_BENCH, *map(str, L), *map(str, g),
           str(seconds)]
    log(f"cpu baseline: {' '.join(cmd[1:])}")
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=seconds * 4 + 120)
    if p.returncode == 0 and p.stdout.strip():
        out = json.loads(p.stdout.strip().splitlines()[-1])
        out["host_cores"] = n
        out["rank_grid"] = g
        return out
    log(f"cpu baseline (mpi) failed: {p.stderr[-400:]}")
    return None


def cpu_baseline(L, seconds):
    if os.path.exists(MPIEXEC) and os.path.exists(REF_BENCH):
        try:
            out = cpu_baseline_mpi(L, seconds)
            if out:
                return out
        except (subprocess.TimeoutExpired, OSError, ValueError) as e:
            log(f"cpu baseline (mpi) failed: {e}")
    have_ref = os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
    for kind in ["port"] + (["reference"] if have_ref else []):
        env = dict(os.environ)
        if kind == "port":
            env["OMP_NUM_THREADS"] = str(host_cores())
        try:
            p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", kind,
                                "--cpu-seconds", str(seconds), "--lattice", *map(str, L)],
                               capture_output=True, text=True, timeout=seconds * 4 + 120, env=env)
            if p.returncode == 0 and p.stdout.strip():
                return json.loads(p.stdout.strip().splitlines()[-1])
            log(f"cpu baseline ({kind}) failed: {p.stderr[-400:]}")
        except subprocess.TimeoutExpired:
            log(f"cpu baseline ({kind}) timed out")
    return None


def pmc_traffic(L, tstream):
    """HBM bytes per launch of the step's kernel from the committed PMC summary
    (taken at 32^3x64 for the kernel the context runs)."""
    path = PMC_SUMMARY.get(tstream)
    if list(L) != [32, 32, 32, 64] or path is None or not os.path.exists(path):
        return None, None
    text = open(path).read()
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", text, re.M)
    return (float(m.group(1)) if m else None), os.path.relpath(path, ROOT)


# ---- GPU workloads -------------------------------------------------------------------
CG_VARIANTS = ("fp64", "fp64_even_odd", "mixed", "mixed_even_odd", "mixed_bicgstab")
DELTA = 1e-2   # reliable update at every residual drop of 1e-2 (the mixed configuration)


def require_converged(res, what):
    """Speed-ups are only reported for solves that met the stop test."""
    if "systems" in res:
        ok = res["status"] == 0 and all(s["converged"] for s in res["systems"])
    else:
        ok = res["status"] == 0 and res["converged"]
    if not ok:
        raise RuntimeError(f"bench: {what} did not converge: status {res['status']}")
    return res


def cg_compare(w, ctx, U, A, b, mass, max_iter=20000, variants=CG_VARIANTS):
    """Time to solution of one system at 1e-10 relative residual (x0 = 0): the
    plain FP64 CG (the reference's algorithm, the baseline of the speedups),
    FP64 CG on the Schur even-odd system, the mixed-precision CG on the full
    and on the Schur system, and mixed-precision BiCGStab on the Schur system
    (a Krylov method beside CG, reported apart from the CG speedup).  Every
    solution is checked by the library
    against the FP64 |M^dag b - M^dag M x| of the full system; the relative
    difference to the plain FP64 solution is reported beside it."""
    fb, x = ctx.field(data=b), ctx.field()
    warm_fp64(ctx, x, fb, U, mass)
    fp = require_converged(ctx.cg(x, fb, U, mass, max_iter=max_iter, tol=1e-10,
                                  tol_relative=True, use_x0=False, check_every=16), "FP64 CG")
    x64 = x.download()

    def err(v):
        return float(np.linalg.norm(v - x64) / np.linalg.norm(x64))

    out = {
        "fp64": {"algorithm": "CGNR on M^dag M (reference)", "iterations": fp["iterations"],
                 "seconds": round(fp["seconds"], 4), "solves_per_s": round(1.0 / fp["seconds"], 3),
                 "true_rel_residual": fp["true_rsd"] / fp["rhs_norm"]},
    }
    if "fp64_even_odd" in variants:
        xe = ctx.field()
        warm_fp64(ctx, xe, fb, U, mass, even_odd=True)
        fpe = require_converged(ctx.cg(xe, fb, U, mass, max_iter=max_iter, tol=1e-10,
                                       tol_relative=True, use_x0=False, check_every=16,
                                       even_odd=True), "FP64 Schur CG")
        out["fp64_even_odd"] = {"algorithm": "CGNR on the Schur complement",
                                "iterations": fpe["iterations"],
                                "seconds": round(fpe["seconds"], 4),
                                "solves_per_s": round(1.0 / fpe["seconds"], 3),
                                "true_rel_residual": fpe["true_rsd"] / fpe["rhs_norm"],
                                "rel_error_vs_fp64_solution": err(xe.download())}
    for key, eo, gprec, label in (("mixed", False, w.F32, "F32 12-real links" + TG_LABEL),
                                  ("mixed_even_odd", True, w.F16, "F16 12-real links" + TG_LABEL)):
        if key not in variants:
            continue
        Us = ctx.gauge(A, prec=gprec, recon=w.RECON_12T)
        for rep in range(2):  # the first call also autotunes the operator
            xm = ctx.field()
            mx = ctx.cg_mixed(xm, fb, U, Us, [mass], tol=1e-10, delta=DELTA, spinor_prec=w.F16,
                              max_iter=max_iter, even_odd=eo)
        require_converged(mx, key)
        s = mx["systems"][0]
        out[key] = {"algorithm": "CGNR on the Schur complement" if eo else "CGNR on M^dag M",
                    "iterations": s["iterations"], "reliable_updates": s["reliable_updates"],
                    "seconds": round(mx["seconds"], 4), "solves_per_s": round(1.0 / mx["seconds"], 3),
                    "true_rel_residual": s["true_rel_rsd"],
                    "rel_error_vs_fp64_solution": err(xm.download()),
                    "precision": f"FP32 arithmetic, F16 spinors, {label}, FP64 reliable updates"}
        del Us
    if "mixed_bicgstab" in variants:
        Us = ctx.gauge(A, prec=w.F16, recon=w.RECON_12T)
        for rep in range(2):  # the first call also autotunes the operator
            xm = ctx.field()
            mx = ctx.bicgstab_mixed(xm, fb, U, Us, [mass], tol=1e-10, delta=DELTA,
                                    spinor_prec=w.F16, max_iter=max_iter)
        require_converged(mx, "mixed BiCGStab")
        s = mx["systems"][0]
        out["mixed_bicgstab"] = {
            "algorithm": "BiCGStab on the Schur complement (A itself, not A^dag A)",
            "iterations": s["iterations"], "reliable_updates": s["reliable_updates"],
            "seconds": round(mx["seconds"], 4), "solves_per_s": round(1.0 / mx["seconds"], 3),
            "true_rel_residual": s["true_rel_rsd"],
            "rel_error_vs_fp64_solution": err(xm.download()),
            "precision": "FP32 arithmetic, F16 spinors, F16 12-real links" + TG_LABEL
                         + ", FP64 reliable updates"}
        out["speedup_mixed_bicgstab_vs_fp64"] = round(fp["seconds"] / mx["seconds"], 3)
        del Us
    mixed = [k for k in ("mixed", "mixed_even_odd") if k in out]
    if mixed:
        best = min(mixed, key=lambda k: out[k]["seconds"])
        out["speedup_mixed_vs_fp64"] = round(fp["seconds"] / out[best]["seconds"], 3)
        out["speedup_basis"] = f"{best} vs fp64 (plain FP64 GPU CG, the reference's algorithm)"
    for k, name in (("mixed", "speedup_mixed_full_system_vs_fp64"),
                    ("fp64_even_odd", "speedup_fp64_even_odd_vs_fp64")):
        if k in out:
            out[name] = round(fp["seconds"] / out[k]["seconds"], 3)
    # precision alone, same algorithm on both sides
    if "mixed" in out:
        out["speedup_precision_full_system"] = round(fp["seconds"] / out["mixed"]["seconds"], 3)
    if "mixed_even_odd" in out and "fp64_even_odd" in out:
        out["speedup_precision_schur"] = round(out["fp64_even_odd"]["seconds"]
                                               / out["mixed_even_odd"]["seconds"], 3)
    out["reliable_update_delta"] = DELTA
    return out


def warm_fp64(ctx, x, fb, U, mass, even_odd=False):
    """Untimed two-iteration FP64 CG: its work vectors then come from the
    context's cache in the timed solve, as the mixed solvers' do after their
    warm-up (a first allocation is part of a solve's time)."""
    ctx.cg(x, fb, U, mass, max_iter=2, tol=1e-10, tol_relative=True, use_x0=False,
           even_odd=even_odd)


def tune_solver(w, ctx, x, fb, U, Us, masses, even_odd=False):
    """Untimed warm-up: the first mixed solve of a context autotunes the sloppy
    operator for this number of systems, link format and operator (full
    lattice, or one parity for the Schur system), about 0.1 s; a solve stopped
    after its first iteration block does that, so the timed solves exclude it."""
    ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, spinor_prec=w.F16, max_iter=1,
                 even_odd=even_odd)


def loop_kernel_case(w, ctx, fb, U, A, fin, masses):
    """The sloppy operator with every system of a site in one thread (links
    loaded once per site, packed-half products) against R = 2 systems per
    thread (R = 1 for one system), both on F16 spinors and F16 12-real links
    (the autotuner picks between them from 8 systems on): the full M and the
    one-parity launch of the Schur solvers, and the batched Schur CG and
    BiCGStab solves on each (every system to 1e-10 on the full system)."""
    n = fin.nrhs
    V = w.volume(fin.ctx.L)
    Us = ctx.gauge(A, prec=w.F16, recon=w.RECON_12T)
    fout = ctx.field(prec=w.F16, nrhs=n)
    Rt = 1 if n == 1 else 2
    out = {"links": "F16 12-real" + TG_LABEL, "rhs_per_thread_compared": Rt}
    for key, R in (("rhs2" if Rt == 2 else "rhs1", Rt), ("loop", w.RHS_LOOP)):
        full = ctx.time_sloppy(fout, fin, Us, masses, R, 0, -1, 10)
        half = ctx.time_sloppy(fout, fin, Us, masses, R, 0, 0, 10)
        out[f"M_ms_{key}"] = round(full, 4)
        out[f"one_parity_ms_{key}"] = round(half, 4)
        # links once per site (t links on the boundary slices only), input
        # spinor twice (neighbours + diagonal), output once
        Lt = fin.ctx.L[3]
        out[f"M_GBps_{key}"] = round(((6 + 2 / Lt) * 24 / n + 3 * 52) * V * n / (full * 1e-3)
                                     / 1e9, 1)
        x = ctx.field(nrhs=n)
        cg = require_converged(ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=R,
                                            site_chunk=0, even_odd=True), f"Schur CG R={R}")
        bi = require_converged(ctx.bicgstab_mixed(x, fb, U, Us, masses, tol=1e-10,
                                                  rhs_per_thread=R, site_chunk=0),
                               f"BiCGStab R={R}")
        out[f"mixed_even_odd_seconds_{key}"] = round(cg["seconds"], 4)
        out[f"mixed_bicgstab_seconds_{key}"] = round(bi["seconds"], 4)
        out[f"max_iterations_{key}"] = [max(s["iterations"] for s in cg["systems"]),
                                        max(s["iterations"] for s in bi["systems"])]
    out["M_frac_of_8TBps_loop"] = round(out["M_GBps_loop"] / HBM_PEAK_GBS, 4)
    return out


def batched32_case(w, L=(32, 32, 32, 64), n=12, kappa=0.135):
    """n Gaussian systems at one kappa on 32^3x64 in one batched mixed solve
    (FP32 arithmetic, F16 spinors, FP64 reliable updates at every residual drop
    of 1e-2) against the plain FP64 CG (the reference's algorithm) of the same
    systems one after another.  The speed-ups: the best batched mixed CG over
    the FP64 CG (time to solution of all n), and the same algorithm on both
    sides (full system: batched mixed CGNR vs FP64 CGNR; Schur system: batched
    mixed vs FP64 Schur CGNR)."""
    L = list(L)
    mass = kappa_mass(kappa)
    log(f"batched {'x'.join(map(str, L))}: generating {n} systems")
    A = w.gauge_near_identity(L, 20211025, 0.3)
    srcs = [w.spinor_gaussian(L, 20211026, r) for r in range(n)]
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        fb = ctx.field(nrhs=n, data=srcs)
        log(f"batched {'x'.join(map(str, L))}: FP64 CG and FP64 Schur CG, one system after another")
        x1, xe1 = ctx.field(), ctx.field()
        b1 = ctx.field(data=srcs[0])
        warm_fp64(ctx, x1, b1, U, mass)
        warm_fp64(ctx, xe1, b1, U, mass, even_odd=True)
        t_fp = t_fpe = 0.0
        its_fp, its_fpe, x64 = [], [], []
        for r in range(n):
            b1.upload(srcs[r])
            res = require_converged(ctx.cg(x1, b1, U, mass, max_iter=20000, tol=1e-10,
                                           tol_relative=True, use_x0=False, check_every=16),
                                    "FP64 CG")
            t_fp += res["seconds"]
            its_fp.append(res["iterations"])
            x64.append(x1.download())
            res = require_converged(ctx.cg(xe1, b1, U, mass, max_iter=20000, tol=1e-10,
                                           tol_relative=True, use_x0=False, check_every=16,
                                           even_odd=True), "FP64 Schur CG")
            t_fpe += res["seconds"]
            its_fpe.append(res["iterations"])
        del x1, xe1, b1
        out = {"workload": f"{'x'.join(map(str, L))}, {n} Gaussian systems, kappa {kappa}, "
                           "stop at 1e-10 relative residual, x0 = 0",
               "fp64_sequential_seconds": round(t_fp, 3),
               "fp64_solves_per_s": round(n / t_fp, 3),
               "fp64_iterations": max(its_fp),
               "fp64_even_odd_sequential_seconds": round(t_fpe, 3),
               "fp64_even_odd_solves_per_s": round(n / t_fpe, 3),
               "fp64_even_odd_iterations": max(its_fpe),
               "reliable_update_delta": DELTA}
        X = ctx.field(nrhs=n)
        for key, gprec, eo in (("mixed", w.F32, False), ("mixed_even_odd", w.F16, True),
                               ("mixed_bicgstab", w.F16, True)):
            log(f"batched {'x'.join(map(str, L))}: {key} of {n} systems")
            Us = ctx.gauge(A, prec=gprec, recon=w.RECON_12T)
            if key == "mixed_bicgstab":
                ctx.bicgstab_mixed(X, fb, U, Us, [mass] * n, tol=1e-10, delta=DELTA,
                                   spinor_prec=w.F16, max_iter=1)  # untimed: autotuning
                mx = ctx.bicgstab_mixed(X, fb, U, Us, [mass] * n, tol=1e-10, delta=DELTA,
                                        spinor_prec=w.F16)
            else:
                tune_solver(w, ctx, X, fb, U, Us, [mass] * n, even_odd=eo)
                mx = ctx.cg_mixed(X, fb, U, Us, [mass] * n, tol=1e-10, delta=DELTA,
                                  spinor_prec=w.F16, even_odd=eo)
            require_converged(mx, f"batched {key}")
            errs = [float(np.linalg.norm(X.download(r) - x64[r]) / np.linalg.norm(x64[r]))
                    for r in range(n)]
            out[key] = {"seconds": round(mx["seconds"], 4),
                        "solves_per_s": round(n / mx["seconds"], 3),
                        "iterations": max(s["iterations"] for s in mx["systems"]),
                        "reliable_updates": max(s["reliable_updates"] for s in mx["systems"]),
                        "max_true_rel_residual": max(s["true_rel_rsd"] for s in mx["systems"]),
                        "max_rel_error_vs_fp64_solution": max(errs),
                        "rhs_per_thread": mx["systems"][0]["rhs_per_thread"],
                        "site_chunk": mx["systems"][0]["site_chunk"],
                        "links": ("F32 12-real" if gprec == w.F32 else "F16 12-real") + TG_LABEL}
            del Us
    best = min(("mixed", "mixed_even_odd"), key=lambda k: out[k]["seconds"])
    out["speedup_mixed_vs_fp64"] = round(t_fp / out[best]["seconds"], 3)
    out["speedup_basis"] = f"batched {best} vs {n} FP64 CG solves (the reference's algorithm)"
    out["speedup_same_algorithm_full_system"] = round(t_fp / out["mixed"]["seconds"], 3)
    out["speedup_same_algorithm_schur"] = round(t_fpe / out["mixed_even_odd"]["seconds"], 3)
    out["speedup_bicgstab_vs_fp64"] = round(t_fp / out["mixed_bicgstab"]["seconds"], 3)
    return out


def batched_case(w):
    """24^3x48: 12 point sources (spin x colour at a seeded site) x 4 kappas = 48 systems."""
    L = [24, 24, 24, 48]
    kappas = [0.120, 0.125, 0.130, 0.135]
    rng = np.random.RandomState(20211027)
    site = [int(rng.randint(0, n)) for n in L]
    A = w.gauge_near_identity(L, 20211025, 0.3)
    srcs, masses = [], []
    for k in kappas:
        for s in range(4):
            for c in range(3):
                srcs.append(w.spinor_point(L, site, s, c))
                masses.append(kappa_mass(k))
    n = len(srcs)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F32, recon=w.RECON_12T)
        fb = ctx.field(nrhs=n, data=srcs)
        x = ctx.field(nrhs=n)
        log("batched: mixed solve of 48 systems")
        tune_solver(w, ctx, x, fb, U, Us, masses)
        mx = require_converged(ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, spinor_prec=w.F16),
                               "batched mixed CG")
        log("batched: mixed solve of 48 systems on the Schur system")
        Us16 = ctx.gauge(A, prec=w.F16, recon=w.RECON_12T)
        xe = ctx.field(nrhs=n)
        tune_solver(w, ctx, xe, fb, U, Us16, masses, even_odd=True)
        mxe = require_converged(ctx.cg_mixed(xe, fb, U, Us16, masses, tol=1e-10,
                                             spinor_prec=w.F16, even_odd=True),
                                "batched mixed Schur CG")
        log("batched: mixed BiCGStab of 48 systems on the Schur system")
        xb = ctx.field(nrhs=n)
        mxb = require_converged(ctx.bicgstab_mixed(xb, fb, U, Us16, masses, tol=1e-10,
                                                   spinor_prec=w.F16), "batched BiCGStab")
        del Us16
        sp = ctx.field(prec=w.F16, nrhs=n, data=srcs)
        bestR, bestC, ms = ctx.tune_report(sp, Us)
        loop = loop_kernel_case(w, ctx, fb, U, A, sp, masses)
        # FP64 baseline: the reference algorithm, one system after the other
        log("batched: FP64 CG of the 48 systems one by one")
        t_fp = 0.0
        its_fp = []
        b1, x1 = ctx.field(data=srcs[0]), ctx.field()
        warm_fp64(ctx, x1, b1, U, masses[0])
        c_fp = np.zeros((len(kappas), L[3]))
        err_e = err_b = 0.0
        for r in range(n):
            b1.upload(srcs[r])
            res = require_converged(ctx.cg(x1, b1, U, masses[r], max_iter=20000, tol=1e-10,
                                           tol_relative=True, use_x0=False, check_every=16),
                                    "FP64 CG")
            t_fp += res["seconds"]
            its_fp.append(res["iterations"])
            c_fp[r // 12] += ctx.pion_correlator(x1, site[3])
            x64 = x1.download()
            nx = np.linalg.norm(x64)
            err_e = max(err_e, float(np.linalg.norm(xe.download(r) - x64) / nx))
            err_b = max(err_b, float(np.linalg.norm(xb.download(r) - x64) / nx))
        # the consumer of the 12 point sources per kappa: pion correlator and
        # effective mass, from the batched mixed Schur solve
        c_mx = np.array([ctx.pion_correlator(xe, site[3], 12 * k, 12)
                         for k in range(len(kappas))])
        c_bi = np.array([ctx.pion_correlator(xb, site[3], 12 * k, 12)
                         for k in range(len(kappas))])
    its = [s["iterations"] for s in mx["systems"]]
    its_e = [s["iterations"] for s in mxe["systems"]]
    its_b = [s["iterations"] for s in mxb["systems"]]
    V = w.volume(L)
    return {
        "workload": "24x24x24x48, 12 point sources x kappa {0.120,0.125,0.130,0.135} = 48 systems",
        "mixed_seconds": round(mx["seconds"], 3), "mixed_solves_per_s": round(n / mx["seconds"], 2),
        "mixed_even_odd_seconds": round(mxe["seconds"], 3),
        "mixed_even_odd_solves_per_s": round(n / mxe["seconds"], 2),
        "fp64_sequential_seconds": round(t_fp, 3), "fp64_solves_per_s": round(n / t_fp, 2),
        "speedup": round(t_fp / min(mx["seconds"], mxe["seconds"]), 3),
        "speedup_full_system": round(t_fp / mx["seconds"], 3),
        "all_converged": all(s["converged"] for s in mx["systems"] + mxe["systems"]),
        "max_true_rel_residual": max(s["true_rel_rsd"] for s in mx["systems"] + mxe["systems"]),
        "mixed_bicgstab_seconds": round(mxb["seconds"], 3),
        "mixed_bicgstab_solves_per_s": round(n / mxb["seconds"], 2),
        "speedup_bicgstab": round(t_fp / mxb["seconds"], 3),
        "bicgstab_all_converged": all(s["converged"] for s in mxb["systems"]),
        "bicgstab_max_true_rel_residual": max(s["true_rel_rsd"] for s in mxb["systems"]),
        "max_rel_error_vs_fp64_solution_mixed_even_odd": err_e,
        "max_rel_error_vs_fp64_solution_bicgstab": err_b,
        "iterations_per_kappa_mixed_bicgstab": [int(np.max(its_b[12 * k:12 * k + 12]))
                                                for k in range(4)],
        "pion_correlator_max_diff_over_C0_bicgstab_vs_fp64":
            float(np.max(np.abs(c_bi - c_fp) / c_fp[:, :1])),
        "iterations_per_kappa_mixed_even_odd": [int(np.max(its_e[12 * k:12 * k + 12]))
                                                for k in range(4)],
        "iterations_per_kappa_mixed": [int(np.max(its[12 * k:12 * k + 12])) for k in range(4)],
        "pion_effective_mass_dt6": [round(float(np.log(c[6] / c[7])), 5) for c in c_mx],
        # the solves are accurate to 1e-10 relative to the whole solution, so
        # far from the source (C decays like e^-m dt) only C(0)-relative
        # differences are meaningful
        "pion_correlator_max_diff_over_C0_mixed_vs_fp64":
            float(np.max(np.abs(c_mx - c_fp) / c_fp[:, :1])),
        "pion_correlator_max_rel_diff_dt_le_8_mixed_vs_fp64":
            float(np.max(np.abs(c_mx[:, :9] - c_fp[:, :9]) / c_fp[:, :9])),
        "iterations_per_kappa_fp64": [int(np.max(its_fp[12 * k:12 * k + 12])) for k in range(4)],
        "rhs_per_thread": mx["systems"][0]["rhs_per_thread"],
        "site_chunk": mx["systems"][0]["site_chunk"],
        "sloppy_M_ms_by_R_chunk": {f"{R}/{C}": round(v, 4) for (R, C), v in ms.items()},
        "sloppy_M_GBps_best": round((8 * 48 / n + 3 * 52) * V * n / (ms[(bestR, bestC)] * 1e-3)
                                    / 1e9, 1),
        "loop_kernel": loop,
    }


def link_product_case(w, ctx, nsys, products_per_M, ms_M):
    """MFMA against VALU for the operator's link product, U h for one link and
    the half spinors of nsys systems (wd_link_product): 2^18 sites, 64
    repetitions on registers per launch (compute-bound), the same F16 inputs
    for both kernels.  144 flop per product (2 spins x 3 x 3 complex
    multiply-adds).  `product_ms_per_M` scales each rate to the products of
    one loop-kernel M launch (8 per site and system) and `share_of_M` divides
    by the measured launch time: what the product would cost if it ran
    alone at that rate."""
    nsites, inner = 1 << 18, 64
    rng = np.random.default_rng(20211028)
    U = rng.standard_normal((nsites, 3, 3)) + 1j * rng.standard_normal((nsites, 3, 3))
    h = rng.standard_normal((nsites, nsys, 2, 3)) + 1j * rng.standard_normal((nsites, nsys, 2, 3))
    out, outs = {}, {}
    for name, impl in (("valu", w.LINKPROD_VALU), ("mfma", w.LINKPROD_MFMA)):
        o, ms = ctx.link_product(impl, U, h, inner=inner, iters=5)
        prods = nsites * nsys * inner
        ps = ms * 1e9 / prods
        out[name] = {"ms": round(ms, 4),
                     "tflops": round(144 * prods / (ms * 1e-3) / 1e12, 2),
                     "ps_per_product": round(ps, 3),
                     "product_ms_per_M": round(ps * products_per_M * 1e-9, 4),
                     "share_of_M": round(ps * products_per_M * 1e-9 / ms_M, 4)}
        outs[name] = o
    out["mfma_speedup_on_the_product"] = round(out["valu"]["ms"] / out["mfma"]["ms"], 3)
    out["max_rel_diff_mfma_vs_valu"] = float(np.abs(outs["mfma"] - outs["valu"]).max()
                                            / np.abs(outs["valu"]).max())
    out["workload"] = (f"{nsites} sites x {nsys} systems, {inner} repetitions per launch; "
                       "VALU: one site per lane, v_dot2_f32_f16; MFMA: v_mfma_f32_16x16x16f16, "
                       "two sites per tile")
    return out


def large_case(w, full_system=False):
    """48^3x96, 12 Gaussian systems at kappa 0.135, F16 spinors and F16 12-real
    links in temporal gauge; the mixed CG on the full M^dag M (about 13 s) only
    with full_system."""
    L = [48, 48, 48, 96]
    n = 12
    mass = kappa_mass(0.135)
    log("large: generating 48^3x96 inputs")
    A = w.gauge_near_identity(L, 20211025, 0.3)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=w.RECON_12T)
        fb = ctx.field(nrhs=n)
        for r in range(n):
            fb.upload(w.spinor_gaussian(L, 20211026, r), r)
        del A
        x = ctx.field(nrhs=n)
        mx = None
        if full_system:
            log("large: mixed solve of 12 systems")
            tune_solver(w, ctx, x, fb, U, Us, [mass] * n)
            mx = require_converged(ctx.cg_mixed(x, fb, U, Us, [mass] * n, tol=1e-10,
                                                spinor_prec=w.F16), "large mixed CG")
        log("large: mixed solve of 12 systems on the Schur system")
        tune_solver(w, ctx, x, fb, U, Us, [mass] * n, even_odd=True)
        mxe = require_converged(ctx.cg_mixed(x, fb, U, Us, [mass] * n, tol=1e-10,
                                             spinor_prec=w.F16, even_odd=True),
                                "large mixed Schur CG")
        xe0 = x.download(0)
        log("large: mixed BiCGStab of 12 systems on the Schur system")
        mxb = require_converged(ctx.bicgstab_mixed(x, fb, U, Us, [mass] * n, tol=1e-10,
                                                   spinor_prec=w.F16), "large BiCGStab")
        xb0 = x.download(0)
        # FP64 time to solution of one of the systems (all 12 share kappa and source statistics);
        # the mixed solvers' cached work memory (about 150 GB) is returned first
        ctx.release_workspace()
        b1 = ctx.field()
        b1.upload(w.spinor_gaussian(L, 20211026, 0))
        x1 = ctx.field()
        warm_fp64(ctx, x1, b1, U, mass)
        log("large: FP64 CG of one system")
        fp = require_converged(ctx.cg(x1, b1, U, mass, max_iter=20000, tol=1e-10,
                                      tol_relative=True, use_x0=False, check_every=16), "FP64 CG")
        x64 = x1.download()
        nx = np.linalg.norm(x64)
        err_e = float(np.linalg.norm(xe0 - x64) / nx)
        err_b = float(np.linalg.norm(xb0 - x64) / nx)
        del xe0, xb0, x64
        fin = ctx.field(prec=w.F16, nrhs=n)
        fin.copy_from(fb)
        bestR, bestC, ms = ctx.tune_report(fin, Us)
        fout = ctx.field(prec=w.F16, nrhs=n)
        ms_loop = ctx.time_sloppy(fout, fin, Us, [mass] * n, w.RHS_LOOP, 0, -1, 5)
        del fout
        log("large: MFMA vs VALU link product")
        linkprod = link_product_case(w, ctx, n, 8 * w.volume(L) * n, ms_loop)
    V = w.volume(L)
    # F16 12-real links in temporal gauge shared by n systems, F16 spinors
    bytes_site = (6 + 2 / L[3]) * 24 / n + 3 * 52
    cg_systems = mxe["systems"] + (mx["systems"] if mx else [])
    out = {
        "workload": ("48x48x48x96, 12 Gaussian systems, kappa 0.135, F16 spinors, F16 12-real "
                     "links" + TG_LABEL),
        "mixed_even_odd_seconds": round(mxe["seconds"], 3),
        "mixed_even_odd_solves_per_s": round(n / mxe["seconds"], 3),
        "fp64_seconds_per_solve": round(fp["seconds"], 3),
        "fp64_solves_per_s": round(1.0 / fp["seconds"], 3),
        "speedup": round(fp["seconds"] * n / mxe["seconds"], 3),
        "speedup_basis": "batched mixed Schur CG vs 12 x the FP64 CG time of system 0",
        "all_converged": all(s["converged"] for s in cg_systems),
        "max_true_rel_residual": max(s["true_rel_rsd"] for s in cg_systems),
        "mixed_bicgstab_seconds": round(mxb["seconds"], 3),
        "mixed_bicgstab_solves_per_s": round(n / mxb["seconds"], 3),
        "speedup_bicgstab": round(fp["seconds"] * n / mxb["seconds"], 3),
        "bicgstab_all_converged": all(s["converged"] for s in mxb["systems"]),
        "bicgstab_max_true_rel_residual": max(s["true_rel_rsd"] for s in mxb["systems"]),
        "rel_error_vs_fp64_solution_system0_mixed_even_odd": err_e,
        "rel_error_vs_fp64_solution_system0_bicgstab": err_b,
        "iterations_mixed_bicgstab": max(s["iterations"] for s in mxb["systems"]),
        "iterations_mixed_even_odd": max(s["iterations"] for s in mxe["systems"]),
        "iterations_fp64": fp["iterations"],
        "rhs_per_thread": mxe["systems"][0]["rhs_per_thread"],
        "site_chunk": mxe["systems"][0]["site_chunk"],
        "sloppy_M_ms_by_R_chunk": {f"{R}/{C}": round(v, 3) for (R, C), v in ms.items()},
        "sloppy_M_GBps_best": round(bytes_site * V * n / (ms[(bestR, bestC)] * 1e-3) / 1e9, 1),
        "sloppy_M_ms_loop_kernel": round(ms_loop, 3),
        "sloppy_M_GBps_loop_kernel": round(bytes_site * V * n / (ms_loop * 1e-3) / 1e9, 1),
        "mfma_vs_valu": linkprod,
    }
    if mx:
        out.update({"mixed_seconds": round(mx["seconds"], 3),
                    "mixed_solves_per_s": round(n / mx["seconds"], 3),
                    "speedup_full_system": round(fp["seconds"] * n / mx["seconds"], 3),
                    "iterations_mixed": max(s["iterations"] for s in mx["systems"])})
    return out


PROFILE_CASES = ("mixed32", "eo32", "bicgstab32", "batched32", "batched48", "batched48_bicgstab",
                 "large12", "sloppy48")


def profile_case(w, name, R=None):
    """One solve (or a few operator launches) per process with a fixed work
    order (no autotuning), for rocprofv3 kernel statistics and counters:
      mixed32              32^3x64, one Gaussian system, mixed CG on M^dag M
                           (F16 spinors, F32 12-real links), delta 1e-2
                           (every sloppy link field here in temporal gauge)
      eo32 / bicgstab32    the same system, mixed CG on the Schur system / mixed
                           BiCGStab (F16 spinors and F16 12-real links)
      batched32            32^3x64, 12 Gaussian systems, mixed CG on M^dag M
                           (F16 spinors, F32 12-real links), delta 1e-2
      batched48            24^3x48, 48 point-source systems at 4 kappas, mixed
                           CG on the Schur system (F16 spinors and links)
      batched48_bicgstab   the same systems, mixed BiCGStab
      large12              48^3x96, 12 Gaussian systems, mixed Schur CG
      sloppy48             24^3x48, 48 systems: 10 launches of the sloppy M
                           (F16 spinors, F16 12-real links, 2 systems per thread)
    R: systems per thread of the sloppy operator (1, 2, 4 or w.RHS_LOOP);
    default 1 for one system, 2 for many.  Prints one JSON line."""
    t0 = time.perf_counter()
    if name in ("mixed32", "eo32", "bicgstab32"):
        L, masses = [32, 32, 32, 64], [kappa_mass(0.135)]
        srcs = [w.spinor_gaussian(L, 20211026)]
    elif name == "batched32":
        L, masses = [32, 32, 32, 64], [kappa_mass(0.135)] * 12
        srcs = [w.spinor_gaussian(L, 20211026, r) for r in range(12)]
    elif name in ("batched48", "batched48_bicgstab", "sloppy48"):
        L = [24, 24, 24, 48]
        rng = np.random.RandomState(20211027)
        site = [int(rng.randint(0, n)) for n in L]
        srcs = [w.spinor_point(L, site, s, c) for k in range(4) for s in range(4)
                for c in range(3)]
        masses = [kappa_mass(k) for k in (0.120, 0.125, 0.130, 0.135) for _ in range(12)]
    else:
        L, masses = [48, 48, 48, 96], [kappa_mass(0.135)] * 12
        srcs = None
    n = len(masses)
    A = w.gauge_near_identity(L, 20211025, 0.3)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F32 if name in ("mixed32", "batched32") else w.F16,
                       recon=w.RECON_12T)
        if srcs is None:
            fb = ctx.field(nrhs=n)
            for r in range(n):
                fb.upload(w.spinor_gaussian(L, 20211026, r), r)
        else:
            fb = ctx.field(nrhs=n, data=srcs)
        del A
        if R is None:
            R = 1 if n == 1 else 2
        if name == "sloppy48":
            fin = ctx.field(prec=w.F16, nrhs=n)
            fin.copy_from(fb)
            fout = ctx.field(prec=w.F16, nrhs=n)
            ctx.sync()
            t1 = time.perf_counter()
            for _ in range(10):
                ctx.apply_M_sloppy(fout, fin, Us, masses, False, R, 0)
            ctx.sync()
            out = {"case": name, "launches": 10, "rhs_per_thread": R,
                   "ms_per_launch": round((time.perf_counter() - t1) * 100, 4)}
        else:
            x = ctx.field(nrhs=n)
            if name in ("batched48_bicgstab", "bicgstab32"):
                res = ctx.bicgstab_mixed(x, fb, U, Us, masses, tol=1e-10, delta=DELTA,
                                         rhs_per_thread=R, site_chunk=0)
            else:
                res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, delta=DELTA,
                                   rhs_per_thread=R, site_chunk=0,
                                   even_odd=name not in ("mixed32", "batched32"))
            require_converged(res, name)
            out = {"case": name, "systems": n, "rhs_per_thread": R,
                   "seconds": round(res["seconds"], 4),
                   "iterations": max(s["iterations"] for s in res["systems"]),
                   "max_true_rel_residual": max(s["true_rel_rsd"] for s in res["systems"])}
    out["lattice"] = L
    out["wall_s"] = round(time.perf_counter() - t0, 2)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--lattice", type=int, nargs=4, default=[32, 32, 32, 64])
    ap.add_argument("--kappa", type=float, default=0.135)
    ap.add_argument("--workload", choices=["default", "full", "single", "batched32", "batched",
                                           "large"], default="default",
                    help="default: the single-system line plus the batched 32^3x64 (12 "
                         "systems), batched 24^3x48 (48 systems) and large 48^3x96 (12 "
                         "systems) solves; full: also the large case's full-system mixed CG; "
                         "single: only the single-system part; batched32 / batched / large: "
                         "the single-system line plus that case")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--cg", nargs="+", choices=CG_VARIANTS, default=list(CG_VARIANTS),
                    help="CG solvers timed (the plain FP64 CG always runs: it is the baseline)")
    ap.add_argument("--cpu-worker", default=None)
    ap.add_argument("--profile-case", choices=PROFILE_CASES, default=None,
                    help="run one solve or operator case for rocprofv3 and exit")
    ap.add_argument("--profile-r", type=int, default=None,
                    help="systems per thread of the profiled case (1, 2, 4 or 64 = every "
                         "system of a site in one thread)")
    args = ap.parse_args()

    if args.cpu_worker:
        cpu_baseline_worker(args.lattice, args.cpu_worker, args.cpu_seconds)
        return
    if args.profile_case:
        profile_case(load_pkg().wilson_hip, args.profile_case, args.profile_r)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs (--gpus > 1) are not "
              "part of it", file=sys.stderr)
        sys.exit(2)

    w = load_pkg().wilson_hip
    L = args.lattice
    V = w.volume(L)
    mass = kappa_mass(args.kappa)
    log(f"generating {'x'.join(map(str, L))} inputs")
    A = w.gauge_near_identity(L, 20211025, 0.3)
    b = w.spinor_gaussian(L, 20211026)
    extra = {}
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        fin = ctx.field(data=b)
        fout = ctx.field()
        for _ in range(args.warmup):
            ctx.apply_M(fout, fin, U, mass, False)
        ctx.sync()
        log(f"timing {args.steps} steps")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.apply_M(fout, fin, U, mass, False)
        ctx.sync()
        el = time.perf_counter() - t0
        ms_step = el * 1e3 / args.steps
        # the step's kernel and the one-parity hop, HIP events on the library stream
        iters = max(args.steps, 10)
        ms_m = ctx.time_op(w.OP_M, fout, fin, U, mass, False, 0, iters)
        tstream = ctx.get_option(w.OPT_TSTREAM)
        ms_hop = ctx.time_op(w.OP_HOP, fout, fin, U, mass, False, 0, iters)
        # the same kernels on 12-real FP64 links (SU(3) links only; not the
        # reference's format, so not the step)
        U12 = ctx.gauge(A, prec=w.F64, recon=12)
        ms_m12 = ctx.time_op(w.OP_M, fout, fin, U12, mass, False, 0, iters)
        ms_hop12 = ctx.time_op(w.OP_HOP, fout, fin, U12, mass, False, 0, iters)
        del U12
        # sloppy operator (the mixed solver's inner kernel), one system
        Us = ctx.gauge(A, prec=w.F32, recon=12)
        s_in = ctx.field(prec=w.F16, data=b)
        s_out = ctx.field(prec=w.F16)
        best_R, best_C, ms_t = ctx.tune_report(s_in, Us)
        # one system: chunk has no effect, and the loop kernel is not a candidate
        ms_s = {R: ms_t[(R, 0)] for R in w.TUNE_R if (R, 0) in ms_t}
        # the same links in temporal gauge (the mixed solvers' links here)
        Ust = ctx.gauge(A, prec=w.F32, recon=w.RECON_12T)
        ms_tg = ctx.time_sloppy(s_out, s_in, Ust, [mass], best_R, 0, -1, max(args.steps, 10))
        del Ust
        if not args.no_cg:
            log("CG: FP64 vs mixed")
            extra["cg"] = cg_compare(w, ctx, U, A, b, mass, variants=args.cg)
            extra["cg"]["workload"] = (f"{'x'.join(map(str, L))}, kappa {args.kappa}, Gaussian "
                                       "source, x0 = 0, stop at 1e-10 relative residual")
        del s_out
    if args.workload in ("default", "full", "batched32"):
        extra["batched_32x32x32x64"] = batched32_case(w)
    if args.workload in ("default", "full", "batched"):
        extra["batched"] = batched_case(w)
    if args.workload in ("default", "full", "large"):
        extra["large"] = large_case(w, full_system=args.workload == "full")

    gflops = FLOP_M * V / (ms_step * 1e-3) / 1e9
    gbs_m = BYTES_M_F64 * V / (ms_m * 1e-3) / 1e9
    gbs_hop = BYTES_HOP_F64 * (V // 2) / (ms_hop * 1e-3) / 1e9
    traffic, traffic_src = pmc_traffic(L, tstream)
    m_kernel = (f"k_hop_f64_ts<dagger=0, HOP_M>, t runs of {tstream}" if tstream
                else "k_hop_f64<dagger=0, HOP_M>")
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
        "data": "synthetic (seeded near-identity SU(3) links exp(0.3iH), Gaussian spinor)",
        "config": {"workload": f"wilson_M_fp64_{'x'.join(map(str, L))}", "lattice": L,
                   "kappa": args.kappa, "bc": "periodic xyz, antiperiodic t",
                   "step": "one full-lattice M application (fused diagonal + hop, both parities)"},
        "roofline": {"bound": "hbm", "achieved": round(gbs_m, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs_m / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": m_kernel,
                     "algorithmic_bytes_per_site": BYTES_M_F64, "sites_per_launch": V,
                     "ms_per_launch": round(ms_m, 4)},
        "hop_roofline": {"kernel": "k_hop_f64<dagger=0, HOP_ONLY>, one parity",
                         "achieved": round(gbs_hop, 1), "frac": round(gbs_hop / HBM_PEAK_GBS, 4),
                         "unit": "GB/s", "algorithmic_bytes_per_site": BYTES_HOP_F64,
                         "sites_per_launch": V // 2, "ms_per_launch": round(ms_hop, 4),
                         "gflops": round(FLOP_HOP * (V // 2) / (ms_hop * 1e-3) / 1e9, 1)},
        "fp64_recon12": {"links": "F64 12-real (row 2 rebuilt; SU(3) links only)",
                         "M_ms_per_launch": round(ms_m12, 4),
                         "M_gflops": round(FLOP_M * V / (ms_m12 * 1e-3) / 1e9, 1),
                         "M_algorithmic_bytes_per_site": 8 * 96 + 3 * 192,
                         "hop_ms_per_launch": round(ms_hop12, 4),
                         "hop_gflops": round(FLOP_HOP * (V // 2) / (ms_hop12 * 1e-3) / 1e9, 1)},
        "sloppy_roofline": {"kernel": "k_hop_sloppy<F16 spinors, F32 12-real links, HOP_M>",
                            "rhs_per_thread": best_R,
                            "achieved": round(BYTES_M_SLOPPY * V / (ms_s[best_R] * 1e-3) / 1e9, 1),
                            "frac": round(BYTES_M_SLOPPY * V / (ms_s[best_R] * 1e-3) / 1e9
                                          / HBM_PEAK_GBS, 4),
                            "unit": "GB/s", "algorithmic_bytes_per_site": BYTES_M_SLOPPY,
                            "ms_per_launch": round(ms_s[best_R], 4),
                            "gflops": round(FLOP_M * V / (ms_s[best_R] * 1e-3) / 1e9, 1)},
        "sloppy_temporal_gauge": {
            "kernel": "k_hop_sloppy<F16 spinors, F32 12-real links in temporal gauge, HOP_M>",
            "rhs_per_thread": best_R,
            "achieved": round(bytes_m_sloppy_tg(L[3]) * V / (ms_tg * 1e-3) / 1e9, 1),
            "frac": round(bytes_m_sloppy_tg(L[3]) * V / (ms_tg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "unit": "GB/s", "algorithmic_bytes_per_site": round(bytes_m_sloppy_tg(L[3]), 2),
            "ms_per_launch": round(ms_tg, 4)},
    }
    if traffic_src:
        line["roofline"]["traffic_source"] = traffic_src
    if traffic:
        # HBM-side bytes the launch actually moved (FETCH_SIZE x 2.0 + WRITE_SIZE,
        # PMC) over the same launch time: a fraction of the physical peak, <= 1
        line["roofline"]["achieved_counter"] = round(traffic / (ms_m * 1e-3) / 1e9, 1)
        line["roofline"]["frac_counter"] = round(traffic / (ms_m * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        line["roofline"]["traffic_over_algorithmic"] = round(traffic / (BYTES_M_F64 * V), 4)
    line.update(extra)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(L, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
