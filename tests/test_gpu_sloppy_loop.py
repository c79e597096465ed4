"""The sloppy operator with every system of a site in one thread
(rhs_per_thread = WD_RHS_LOOP: links loaded once per site; F16 links kept as
packed halves, SU(3) products by packed F16 dot products with FP32
accumulation; F32 links kept as FP32, the arithmetic of the R kernels).

Tolerances: against the FP64 oracle the same 2e-3 relative L2 as the other
F16-spinor formats; against the R = 1 kernel (FP32 products of the same F16
inputs) 1e-3, the size of the extra F16 roundings of the half spinor and the
rebuilt third link row.  Solvers running on it must still meet the FP64 stop
test 1e-10 (re-checked with the oracle), every system on its own iteration
count (per-system masking), on lattices whose half volume is not a multiple
of the workgroup (idle lanes).
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, rel_l2

pytestmark = pytest.mark.gpu


def _setup(w, L, nrhs, seed=20211025):
    A = w.gauge_near_identity(L, seed, 0.3)
    bs = [w.spinor_gaussian(L, seed + 1, r) for r in range(nrhs)]
    return A, bs


@pytest.mark.parametrize("L", [[4, 4, 6, 8], [6, 4, 6, 10]], ids=lambda L: "x".join(map(str, L)))
@pytest.mark.parametrize("dag", [0, 1])
def test_loop_operator_matches_oracle(w, oracle, L, dag):
    n = 5
    A, bs = _setup(w, L, n)
    masses = [kappa_to_mass(k) for k in (0.12, 0.125, 0.13, 0.135, 0.11)]
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fin = ctx.field(prec=w.F16, nrhs=n, data=bs)
        fl, f1 = ctx.field(prec=w.F16, nrhs=n), ctx.field(prec=w.F16, nrhs=n)
        ctx.apply_M_sloppy(fl, fin, Us, masses, bool(dag), w.RHS_LOOP)
        ctx.apply_M_sloppy(f1, fin, Us, masses, bool(dag), 1)
        got = [fl.download(r) for r in range(n)]
        ref1 = [f1.download(r) for r in range(n)]
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in range(n):
        ref = oracle.dslash(bs[r], Af, L, masses[r], dag)
        e_or, e_r1 = rel_l2(got[r], ref), rel_l2(got[r], ref1[r])
        print(f"L={L} dag={dag} system {r}: vs oracle {e_or:.2e} (R=1: "
              f"{rel_l2(ref1[r], ref):.2e}), vs R=1 {e_r1:.2e}", flush=True)
        assert e_or < 2e-3 and e_r1 < 1e-3


@pytest.mark.parametrize("solver", ["cg_mixed", "bicgstab_mixed"])
def test_loop_solvers_reach_fp64_residual(w, oracle, solver):
    L = [6, 4, 6, 10]
    n = 6
    A, bs = _setup(w, L, n)
    masses = [kappa_to_mass(k) for k in (0.11, 0.12, 0.125, 0.13, 0.135, 0.12)]
    bs[5] = np.zeros_like(bs[5])   # solved by x = 0 without iterating
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=n, data=bs)
        x = ctx.field(nrhs=n)
        kw = {"even_odd": True} if solver == "cg_mixed" else {}
        res = getattr(ctx, solver)(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=w.RHS_LOOP,
                                   check_every=4, **kw)
        xs = [x.download(r) for r in range(n)]
    its = [s["iterations"] for s in res["systems"]]
    print(solver, its, [s["rhs_per_thread"] for s in res["systems"]], flush=True)
    assert res["status"] == 0
    assert all(s["rhs_per_thread"] == w.RHS_LOOP for s in res["systems"])
    assert its[0] < its[4] and its[5] == 0 and np.all(xs[5] == 0)
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in range(5):
        Mdb = oracle.dslash(bs[r], Af, L, masses[r], 1)
        rr = Mdb - oracle.dslash(oracle.dslash(xs[r], Af, L, masses[r], 0), Af, L, masses[r], 1)
        assert res["systems"][r]["converged"]
        assert np.linalg.norm(rr) / np.linalg.norm(Mdb) < 1.05e-10


def test_loop_graph_and_direct_launch_agree(w):
    L = [4, 4, 4, 8]
    A, bs = _setup(w, L, 3)
    masses = [0.1, 0.2, -0.1]
    out = []
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=3, data=bs)
        for g in (True, False):
            x = ctx.field(nrhs=3)
            res = ctx.bicgstab_mixed(x, fb, U, Us, masses, use_graph=g, check_every=3,
                                     rhs_per_thread=w.RHS_LOOP)
            out.append(([x.download(r) for r in range(3)], res))
    for r in range(3):
        assert np.array_equal(out[0][0][r], out[1][0][r])


def test_loop_argument_errors(w):
    L = [4, 4, 4, 4]
    with w.Context(L) as ctx:
        U32 = ctx.gauge(w.gauge_unit(L), prec=w.F32, recon=18)
        f16 = ctx.field(prec=w.F16, nrhs=2)
        o16 = ctx.field(prec=w.F16, nrhs=2)
        with pytest.raises(w.WilsonError) as e:   # 18-real links: not a loop-kernel format
            ctx.apply_M_sloppy(o16, f16, U32, [0.1, 0.1], rhs_per_thread=w.RHS_LOOP)
        assert e.value.code == w.WD_EARG
        U = ctx.gauge(w.gauge_unit(L))
        x, b = ctx.field(nrhs=2), ctx.field(nrhs=2)
        with pytest.raises(w.WilsonError) as e:
            ctx.cg_mixed(x, b, U, U32, [0.1, 0.1], rhs_per_thread=w.RHS_LOOP)
        assert e.value.code == w.WD_EARG


def test_time_sloppy(w):
    """wd_time_sloppy: positive launch times for both kernels and both launch
    shapes, argument errors before any launch."""
    L = [8, 8, 8, 8]
    A, bs = _setup(w, L, 4)
    masses = [0.1, 0.2, 0.3, 0.0]
    with w.Context(L) as ctx:
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fin = ctx.field(prec=w.F16, nrhs=4, data=bs)
        fout = ctx.field(prec=w.F16, nrhs=4)
        for R in (1, 2, w.RHS_LOOP):
            for cb in (-1, 0, 1):
                ms = ctx.time_sloppy(fout, fin, Us, masses, R, 0, cb, 3)
                assert ms > 0, (R, cb)
        for kw in ({"cb": 2}, {"iters": 0}, {"rhs_per_thread": 3}):
            with pytest.raises(w.WilsonError) as e:
                ctx.time_sloppy(fout, fin, Us, masses, **kw)
            assert e.value.code == w.WD_EARG


def test_loop_groups_and_prefetch_bit_identical(w):
    """WD_OPT_LOOP_SYSTEMS (systems per workgroup group) and
    WD_OPT_LOOP_PREFETCH only reorder work: M, M^dagger and a mixed Schur
    solve give the same bits for every group size and both prefetch depths,
    including groups that do not divide the system count."""
    L = [6, 4, 6, 10]
    n = 11
    A, bs = _setup(w, L, n)
    masses = [kappa_to_mass(0.11 + 0.002 * r) for r in range(n)]
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        assert ctx.get_option(w.OPT_LOOP_SYSTEMS) == 0
        assert ctx.get_option(w.OPT_LOOP_PREFETCH) == 0
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=n, data=bs)
        fin = ctx.field(prec=w.F16, nrhs=n, data=bs)
        fout = ctx.field(prec=w.F16, nrhs=n)
        runs = {}
        for K in (0, 1, 3):
            for per in (0, 1, 4, 5, 11, 64):
                ctx.set_option(w.OPT_LOOP_PREFETCH, K)
                ctx.set_option(w.OPT_LOOP_SYSTEMS, per)
                assert ctx.get_option(w.OPT_LOOP_SYSTEMS) == per
                got = []
                for dag in (False, True):
                    ctx.apply_M_sloppy(fout, fin, Us, masses, dag, w.RHS_LOOP)
                    got += [fout.download(r) for r in range(n)]
                x = ctx.field(nrhs=n)
                res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=w.RHS_LOOP,
                                   even_odd=True)
                assert res["status"] == 0
                got += [x.download(r) for r in range(n)]
                runs[(K, per)] = (got, [s["iterations"] for s in res["systems"]])
        with pytest.raises(w.WilsonError) as e:
            ctx.set_option(w.OPT_LOOP_SYSTEMS, -1)
        assert e.value.code == w.WD_EARG
        # parity-major work order (both parities of M)
        ctx.set_option(w.OPT_LOOP_PREFETCH, 0)
        ctx.set_option(w.OPT_LOOP_SYSTEMS, 0)
        ctx.set_option(w.OPT_PARITY_ORDER, 0)
        got = []
        for dag in (False, True):
            ctx.apply_M_sloppy(fout, fin, Us, masses, dag, w.RHS_LOOP)
            got += [fout.download(r) for r in range(n)]
        ctx.set_option(w.OPT_PARITY_ORDER, 1)
        assert all(np.array_equal(a, b) for a, b in zip(got, runs[(0, 0)][0][:2 * n]))
        # the site-chunk argument of the loop kernel counts systems per pass
        for chunk in (3, 16):
            got = []
            for dag in (False, True):
                ctx.apply_M_sloppy(fout, fin, Us, masses, dag, w.RHS_LOOP, chunk)
                got += [fout.download(r) for r in range(n)]
            x = ctx.field(nrhs=n)
            res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=w.RHS_LOOP,
                               site_chunk=chunk, even_odd=True)
            assert res["status"] == 0
            assert all(s["site_chunk"] == chunk for s in res["systems"])
            got += [x.download(r) for r in range(n)]
            runs[("chunk", chunk)] = (got, [s["iterations"] for s in res["systems"]])
    base_got, base_its = runs[(0, 0)]
    for key, (got, its) in runs.items():
        assert its == base_its, key
        assert all(np.array_equal(a, b) for a, b in zip(got, base_got)), key


@pytest.mark.parametrize("L", [[4, 4, 6, 8], [6, 4, 6, 10]], ids=lambda L: "x".join(map(str, L)))
def test_loop_f32_links_bit_identical_to_r_kernels(w, oracle, L):
    """With F32 12-real links the loop kernel keeps the links in FP32 and does
    the FP32 arithmetic of the R-systems-per-thread kernel, so M and M^dagger
    are bit-identical to R = 1, 2, 4 for every group size; a Schur CG on it
    reaches the oracle's FP64 residual (its dot products sum per wavefront,
    so its iterates may differ from R = 1 by roundings)."""
    n = 9
    A, bs = _setup(w, L, n)
    masses = [kappa_to_mass(0.11 + 0.003 * r) for r in range(n)]
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F32, recon=12)
        fin = ctx.field(prec=w.F16, nrhs=n, data=bs)
        fout = ctx.field(prec=w.F16, nrhs=n)
        for dag in (False, True):
            ref = None
            for R, chunk in ((1, 0), (2, 0), (4, 16), (w.RHS_LOOP, 0), (w.RHS_LOOP, 4),
                             (w.RHS_LOOP, 64)):
                ctx.apply_M_sloppy(fout, fin, Us, masses, dag, R, chunk)
                got = [fout.download(r) for r in range(n)]
                if ref is None:
                    ref = got
                assert all(np.array_equal(a, b) for a, b in zip(got, ref)), (dag, R, chunk)
        fb = ctx.field(nrhs=n, data=bs)
        x = ctx.field(nrhs=n)
        res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=w.RHS_LOOP,
                           even_odd=True, check_every=4)
        xs = [x.download(r) for r in range(n)]
    assert res["status"] == 0
    assert all(s["rhs_per_thread"] == w.RHS_LOOP for s in res["systems"])
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in (0, n - 1):
        Mdb = oracle.dslash(bs[r], Af, L, masses[r], 1)
        rr = Mdb - oracle.dslash(oracle.dslash(xs[r], Af, L, masses[r], 0), Af, L, masses[r], 1)
        assert np.linalg.norm(rr) / np.linalg.norm(Mdb) < 1.05e-10


def test_loop_more_than_64_systems(w, oracle):
    """70 systems: the loop kernel scans the applied-system flags in chunks of
    64 lanes; the operator (F32 links) matches R = 1 bit for bit for every
    group size, and a masked Schur CG over all 70 (systems converging at
    different iterations, one zero source) meets the oracle's FP64 residual."""
    L = [4, 4, 4, 8]
    n = 70
    A, bs = _setup(w, L, n)
    masses = [kappa_to_mass(0.10 + 0.0005 * r) for r in range(n)]
    bs[66] = np.zeros_like(bs[66])
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F32, recon=12)
        fin = ctx.field(prec=w.F16, nrhs=n, data=bs)
        fout = ctx.field(prec=w.F16, nrhs=n)
        ctx.apply_M_sloppy(fout, fin, Us, masses, False, 1, 0)
        ref = [fout.download(r) for r in range(n)]
        for chunk in (0, 5, 64):
            ctx.apply_M_sloppy(fout, fin, Us, masses, False, w.RHS_LOOP, chunk)
            assert all(np.array_equal(fout.download(r), ref[r]) for r in range(n)), chunk
        fb = ctx.field(nrhs=n, data=bs)
        x = ctx.field(nrhs=n)
        res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=w.RHS_LOOP,
                           site_chunk=16, even_odd=True, check_every=4)
        xs = {r: x.download(r) for r in (0, 63, 64, 66, 69)}
    its = [s["iterations"] for s in res["systems"]]
    assert res["status"] == 0 and its[66] == 0 and np.all(xs[66] == 0)
    assert len(set(its)) > 3   # systems stop at different iterations
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in (0, 63, 64, 69):
        Mdb = oracle.dslash(bs[r], Af, L, masses[r], 1)
        rr = Mdb - oracle.dslash(oracle.dslash(xs[r], Af, L, masses[r], 0), Af, L, masses[r], 1)
        assert np.linalg.norm(rr) / np.linalg.norm(Mdb) < 1.05e-10, r
