"""Mixed-precision and multi-rhs paths vs the FP64 CPU oracle, through the C-ABI.

Tolerances (stated per precision): the sloppy operator with F32 spinors and
F32 links agrees with the FP64 oracle to 1e-6 relative L2, with F16 spinors
(per-site scale) and 12-real F16 or F32 links to 2e-3.  The mixed solver must
reach the FP64 true residual |M^dag b - M^dag M x| < 1e-10 |M^dag b|, which is
re-checked here with the oracle's FP64 operator.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, rel_l2

pytestmark = pytest.mark.gpu


def _setup(w, L, nrhs, seed=20211025):
    A = w.gauge_near_identity(L, seed, 0.3)
    bs = [w.spinor_gaussian(L, seed + 1, r) for r in range(nrhs)]
    return A, bs


@pytest.mark.parametrize("sprec,gprec,recon,tol", [
    ("F32", "F32", 18, 1e-6), ("F32", "F32", 12, 1e-6),
    ("F16", "F32", 12, 2e-3), ("F16", "F16", 12, 2e-3)])
@pytest.mark.parametrize("dag", [0, 1])
def test_sloppy_operator_matches_oracle(w, oracle, sprec, gprec, recon, tol, dag):
    L = [4, 4, 6, 8]
    A, bs = _setup(w, L, 3)
    masses = [kappa_to_mass(k) for k in (0.12, 0.13, 0.135)]
    sp, gp = getattr(w, sprec), getattr(w, gprec)
    print(f"sloppy {sprec}/{gprec}/{recon} dag={dag}", flush=True)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        Us = ctx.gauge(A, prec=gp, recon=recon)
        fin = ctx.field(prec=sp, nrhs=3, data=bs)
        fout = ctx.field(prec=sp, nrhs=3)
        ctx.apply_M_sloppy(fout, fin, Us, masses, bool(dag), 1)
        got = [fout.download(r) for r in range(3)]
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in range(3):
        ref = oracle.dslash(bs[r], Af, L, masses[r], dag)
        assert rel_l2(got[r], ref) < tol


def test_rhs_per_thread_variants_agree_bitwise(w):
    L = [4, 4, 4, 8]
    A, bs = _setup(w, L, 5)
    masses = [0.1, -0.1, 0.2, 0.0, 0.05]
    outs = {}
    with w.Context(L) as ctx:
        Us = ctx.gauge(A, prec=w.F32, recon=12)
        fin = ctx.field(prec=w.F16, nrhs=5, data=bs)
        for R in (1, 2, 4):
            fout = ctx.field(prec=w.F16, nrhs=5)
            ctx.apply_M_sloppy(fout, fin, Us, masses, False, R)
            outs[R] = [fout.download(r) for r in range(5)]
    for R in (2, 4):
        for r in range(5):
            assert np.array_equal(outs[R][r], outs[1][r])


def test_site_chunk_orders_agree_bitwise(w):
    # 8^4: 2048 sites per parity = 8 site blocks, 16 over both parities; chunks
    # of 3 blocks leave a short last chunk, 1 is the finest order, 64 > 16 a
    # single chunk.  The work order must not change a single bit.
    L = [8, 8, 8, 8]
    A, bs = _setup(w, L, 7)
    masses = [0.1, -0.1, 0.2, 0.0, 0.05, 0.3, -0.2]
    with w.Context(L) as ctx:
        Us = ctx.gauge(A, prec=w.F32, recon=12)
        fin = ctx.field(prec=w.F16, nrhs=7, data=bs)
        for dag in (False, True):
            ref = None
            for R in (1, 2, 4):
                for C in (0, 1, 3, 64):
                    fout = ctx.field(prec=w.F16, nrhs=7)
                    ctx.apply_M_sloppy(fout, fin, Us, masses, dag, R, C)
                    got = [fout.download(r) for r in range(7)]
                    if ref is None:
                        ref = got
                        continue
                    for r in range(7):
                        assert np.array_equal(got[r], ref[r]), (dag, R, C, r)


def _true_rel_residual(oracle, b, x, A, L, mass):
    Mdb = oracle.dslash(b, A, L, mass, 1)
    r = Mdb - oracle.dslash(oracle.dslash(x, A, L, mass, 0), A, L, mass, 1)
    return np.linalg.norm(r) / np.linalg.norm(Mdb)


@pytest.mark.parametrize("sprec,gprec,recon", [("F16", "F32", 12), ("F32", "F32", 18),
                                                ("F16", "F16", 12)])
def test_mixed_cg_reaches_fp64_true_residual(w, oracle, sprec, gprec, recon):
    L = [8, 8, 8, 8]
    nrhs = 3
    A, bs = _setup(w, L, nrhs)
    kappas = [0.12, 0.125, 0.13]
    masses = [kappa_to_mass(k) for k in kappas]
    print(f"mixed cg {sprec}/{gprec}/{recon}", flush=True)
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=getattr(w, gprec), recon=recon)
        fb = ctx.field(nrhs=nrhs, data=bs)
        x = ctx.field(nrhs=nrhs)
        res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, spinor_prec=getattr(w, sprec))
        xs = [x.download(r) for r in range(nrhs)]
        # FP64 reference solve on the GPU for the same systems
        xref = []
        for r in range(nrhs):
            b1 = ctx.field(data=bs[r])
            x1 = ctx.field()
            ctx.cg(x1, b1, U, masses[r], max_iter=20000, tol=1e-12, tol_relative=True,
                   use_x0=False)
            xref.append(x1.download())
    assert res["status"] == 0
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    for r in range(nrhs):
        s = res["systems"][r]
        print(r, s, flush=True)
        assert s["converged"] and s["true_rel_rsd"] < 1e-10
        assert s["reliable_updates"] >= 2
        assert _true_rel_residual(oracle, bs[r], xs[r], Af, L, masses[r]) < 1.2e-10
        assert rel_l2(xs[r], xref[r]) < 1e-7


def test_mixed_cg_masks_converged_systems(w, oracle):
    """Systems of very different difficulty in one solve: each stops on its own."""
    L = [4, 4, 4, 8]
    A, bs = _setup(w, L, 4)
    masses = [2.0, 0.5, kappa_to_mass(0.13), 1.0]
    bs[3] = np.zeros_like(bs[3])  # a zero right-hand side is solved by x = 0 at once
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F32, recon=12)
        fb = ctx.field(nrhs=4, data=bs)
        x = ctx.field(nrhs=4)
        res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=2, check_every=4)
        xs = [x.download(r) for r in range(4)]
    its = [s["iterations"] for s in res["systems"]]
    print(its, flush=True)
    assert its[0] < its[1] < its[2] and its[3] == 0
    assert np.all(xs[3] == 0)
    for r in range(3):
        assert _true_rel_residual(oracle, bs[r], xs[r], A, L, masses[r]) < 1.2e-10


def test_mixed_graph_and_direct_launch_agree(w):
    L = [4, 4, 4, 8]
    A, bs = _setup(w, L, 2)
    masses = [0.1, 0.2]
    out = []
    with w.Context(L) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=2, data=bs)
        for g in (True, False):
            x = ctx.field(nrhs=2)
            res = ctx.cg_mixed(x, fb, U, Us, masses, use_graph=g, check_every=3,
                               rhs_per_thread=1)
            out.append(([x.download(r) for r in range(2)], res))
    for r in range(2):
        assert np.array_equal(out[0][0][r], out[1][0][r])
        assert out[0][1]["systems"][r]["iterations"] == out[1][1]["systems"][r]["iterations"]


@pytest.mark.parametrize("even_odd", [False, True])
def test_reliable_update_slots_do_not_change_iterates(w, even_odd):
    """A system whose reliable update falls due waits for the next slot without
    iterating, so the slot spacing changes neither iterates nor counts: four
    systems of different masses (updates in different iterations, two per
    thread so a waiting system shares its workgroups with an iterating one),
    slots after every iteration, every 3rd, only at the end of each block of 4
    (spacing 5 > block), and every 10th in blocks of 20, where groups of 8
    iterations (the most search directions kept for one x_s pass) end
    between the slots."""
    L = [4, 4, 6, 8]
    A, bs = _setup(w, L, 4)
    masses = [kappa_to_mass(k) for k in (0.11, 0.12, 0.13, 0.135)]
    out = []
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=4, data=bs)
        for K, chunk in ((1, 4), (3, 4), (5, 4), (10, 20)):
            x = ctx.field(nrhs=4)
            res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, rhs_per_thread=2,
                               check_every=chunk, even_odd=even_odd, reliable_every=K)
            out.append(([x.download(r) for r in range(4)], res))
            x.close()
    its = [[s["iterations"] for s in o[1]["systems"]] for o in out]
    ups = [[s["reliable_updates"] for s in o[1]["systems"]] for o in out]
    print("iterations", its, "updates", ups, flush=True)
    assert len(set(its[0])) > 1  # the systems do stop at different iterations
    for k in (1, 2, 3):
        assert its[k] == its[0] and ups[k] == ups[0]
        assert [s["true_rel_rsd"] for s in out[k][1]["systems"]] == \
            [s["true_rel_rsd"] for s in out[0][1]["systems"]]
        for r in range(4):
            assert np.array_equal(out[k][0][r], out[0][0][r])
    assert all(s["converged"] for s in out[0][1]["systems"])


def test_f16_storage_of_underflowing_sites(w):
    """Sites whose components are far below FP32 normal range (as far from a
    point source) are stored as exact zeros in F16 storage, never as NaN."""
    L = [4, 4, 4, 4]
    b = w.spinor_gaussian(L, 3)
    b[:12] *= 1e-40          # one site entirely denormal in FP32
    b[12:24] = 0.0           # one exactly zero site
    A = w.gauge_near_identity(L, 4, 0.3)
    with w.Context(L) as ctx:
        f = ctx.field(prec=w.F16, data=b)
        got = f.download()
        Us = ctx.gauge(A, prec=w.F32, recon=12)
        out = ctx.field(prec=w.F16)
        ctx.apply_M_sloppy(out, f, Us, [0.1], False, 1)
        Mb = out.download()
    assert np.all(np.isfinite(got)) and np.all(got[:24] == 0)
    assert np.all(np.isfinite(Mb))
    assert rel_l2(got[24:], b[24:]) < 1e-3


def test_tune_report(w):
    L = [8, 8, 8, 8]
    _, bs = _setup(w, L, 6)
    A = w.gauge_unit(L)
    with w.Context(L) as ctx:
        Us = ctx.gauge(A, prec=w.F32, recon=12)
        f = ctx.field(prec=w.F16, nrhs=6, data=bs)
        best, chunk, ms = ctx.tune_report(f, Us)
        # 6 systems: R = 1, 2, 4 all leave more than one group, so every
        # candidate is timed; F32 links: no loop-kernel candidate
        assert set(ms) == {(R, C) for R in (1, 2, 4) for C in w.TUNE_CHUNK}
        assert best in w.TUNE_R and chunk in w.TUNE_CHUNK
        assert all(v > 0 for v in ms.values())
        assert ms[(best, chunk)] == min(ms.values())
        f1 = ctx.field(prec=w.F16, nrhs=1, data=bs[:1])
        _, chunk1, ms1 = ctx.tune_report(f1, Us)
        # one system: a single group for every R, chunk candidates not timed
        assert set(ms1) == {(R, 0) for R in (1, 2, 4)} and chunk1 == 0
        # F16 links: the loop kernel is a candidate from 8 systems on (chunk 0)
        Uh = ctx.gauge(A, prec=w.F16, recon=12)
        _, _, ms6 = ctx.tune_report(f, Uh)
        assert (w.RHS_LOOP, 0) not in ms6
        f8 = ctx.field(prec=w.F16, nrhs=8, data=bs + bs[:2])
        best8, chunk8, ms8 = ctx.tune_report(f8, Uh)
        assert (w.RHS_LOOP, 0) in ms8 and (w.RHS_LOOP, 16) not in ms8
        assert ms8[(best8, chunk8)] == min(ms8.values())
        # 40 systems: the loop kernel's chunk candidates are systems per pass,
        # 12 (chunk 0, 4 passes), 16 (3), 32 (2), 64 (1 pass): all timed
        f40 = ctx.field(prec=w.F16, nrhs=40, data=(bs * 7)[:40])
        best40, chunk40, ms40 = ctx.tune_report(f40, Uh)
        assert {(w.RHS_LOOP, C) for C in w.TUNE_CHUNK} <= set(ms40)
        assert ms40[(best40, chunk40)] == min(ms40.values())


def test_sloppy_argument_errors(w):
    L = [4, 4, 4, 4]
    with w.Context(L) as ctx:
        with pytest.raises(w.WilsonError):
            ctx.gauge(w.gauge_unit(L), prec=w.F16, recon=18)
        Us = ctx.gauge(w.gauge_unit(L), prec=w.F32, recon=12)
        f64 = ctx.field(nrhs=2)
        out = ctx.field(nrhs=2)
        with pytest.raises(w.WilsonError):
            ctx.apply_M_sloppy(out, f64, Us, [0.1, 0.1])  # F64 fields are not sloppy
        f16 = ctx.field(prec=w.F16, nrhs=2)
        o16 = ctx.field(prec=w.F16, nrhs=2)
        with pytest.raises(w.WilsonError):
            ctx.apply_M_sloppy(o16, f16, Us, [0.1, 0.1], rhs_per_thread=3)
        with pytest.raises(w.WilsonError):
            ctx.apply_M_sloppy(o16, f16, Us, [0.1, 0.1], site_chunk=-1)
        with pytest.raises(w.WilsonError):
            ctx.cg_mixed(out, f64, ctx.gauge(w.gauge_unit(L)), Us, [0.1, 0.1],
                         reliable_every=-1)


def test_sloppy_work_orders_bit_identical(w):
    """The autotuner's work-order axis: the sloppy operator gives the same bits
    in the z-t tiled order and in the natural site order, for every kernel
    shape; the solvers report the order they ran with."""
    L = [8, 8, 24, 8]   # a time slice of one parity = 3 site blocks: tiled columns of 2
    n = 9
    _, bs = _setup(w, L, n)
    A = w.gauge_near_identity(L, 5, 0.3)
    masses = [0.1 + 0.01 * r for r in range(n)]
    with w.Context(L, bc=[1, 1, 1, -1]) as ctx:
        tile0 = ctx.get_option(w.OPT_BLOCK_TILE)
        assert tile0 > 0
        fin = ctx.field(prec=w.F16, nrhs=n, data=bs)
        fout = ctx.field(prec=w.F16, nrhs=n)
        for gp in (w.F32, w.F16):
            Us = ctx.gauge(A, prec=gp, recon=12)
            for R in (1, 2, w.RHS_LOOP):
                outs = []
                for tile in (tile0, 0):
                    ctx.set_option(w.OPT_BLOCK_TILE, tile)
                    ctx.apply_M_sloppy(fout, fin, Us, masses, False, R, 0)
                    outs.append([fout.download(r) for r in range(n)])
                ctx.set_option(w.OPT_BLOCK_TILE, tile0)
                assert all(np.array_equal(a, b) for a, b in zip(*outs)), (gp, R)
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=12)
        fb = ctx.field(nrhs=n, data=bs)
        x = ctx.field(nrhs=n)
        res = ctx.cg_mixed(x, fb, U, Us, masses, tol=1e-10, even_odd=True)
        assert res["status"] == 0
        assert all(s["block_tile"] in (0, -1) for s in res["systems"])
