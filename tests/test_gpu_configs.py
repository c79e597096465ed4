"""The benchmark configurations at their full sizes, through the C-ABI.

Operator parity at production shape: wd_hop (both output parities, M and
M^dag hops) and wd_apply_M with the context's default kernels -- at 32^3x64
the t-streaming M (runs of 16, 512 workgroups, XCD remap), at 24^3x48 the
site-block kernel, z-t tiled block columns and the parities side by side --
against the OpenMP oracle (pinned to the compiled reference's golden vectors)
at 1e-13 relative L2, the drop-in Dslash bar.  Antiperiodic t on the GPU is
checked against the periodic oracle on links negated on the last time slice.

Solver parity at the configuration sizes (the reference has no mixed or
batched solver, so the target is the solution): every system's FP64 true
residual |M^dag b - M^dag M x| <= 1e-10 |M^dag b|, recomputed with the
oracle's FP64 operator for the systems named in each test, and agreement
with the FP64 CG's (wd_cg, the reference's algorithm) solution to 1e-7
relative L2 (the two differ by the conditioning of the normal equations,
observed 1e-8).
  32^3x64, kappa 0.135, one Gaussian source: mixed CG on M^dag M, mixed CG on
      the Schur system and mixed BiCGStab, reliable updates at every residual
      drop of 1e-2.
  24^3x48, 12 point sources (spin x colour at one seeded site) x kappa
      {0.120, 0.125, 0.130, 0.135} = 48 systems in one batched solve: per-kappa
      iteration counts differ (per-system masking), one kappa's 12 solutions
      against 12 FP64 solves.
  48^3x96, 12 Gaussian sources at kappa 0.135, F16 spinors and F16 12-real
      links: one batched Schur solve.
The sloppy links of the first two are tested as stored and in temporal gauge
(WD_RECON_12T, the bench's choice); the third runs in temporal gauge.
"""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, rel_l2

pytestmark = pytest.mark.gpu

TOL_OP = 1e-13
BC = [1, 1, 1, -1]
KAPPAS = [0.120, 0.125, 0.130, 0.135]


def _flipped(w, L, A):
    Af = [a.copy() for a in A]
    w.gauge_flip_boundary(L, 3, Af)
    return Af


def _normal_residual(oracle, x, b, A, L, mass):
    Mdb = oracle.dslash(b, A, L, mass, 1)
    r = Mdb - oracle.dslash(oracle.dslash(x, A, L, mass, 0), A, L, mass, 1)
    return float(np.linalg.norm(r) / np.linalg.norm(Mdb))


@pytest.mark.parametrize("L", [[24, 24, 24, 48], [32, 32, 32, 64]],
                         ids=lambda L: "x".join(map(str, L)))
def test_operator_production_shape(w, oracle, L):
    A = w.gauge_near_identity(L, 20211025, 0.3)
    b = w.spinor_gaussian(L, 20211026)
    Af = _flipped(w, L, A)
    mass = kappa_to_mass(0.13)
    got = {}
    with w.Context(L, bc=BC) as ctx:
        opts = {name: ctx.get_option(getattr(w, name))
                for name in ("OPT_TSTREAM", "OPT_BLOCK_TILE", "OPT_PARITY_ORDER",
                             "OPT_XCD_REMAP")}
        print(f"L={L} default options {opts}", flush=True)
        U = ctx.gauge(A)
        fin, fout = ctx.field(data=b), ctx.field()
        for dag in (0, 1):
            ctx.apply_M(fout, fin, U, mass, bool(dag))
            got[("M", dag)] = fout.download()
            for cb in (0, 1):
                ctx.hop(fout, fin, U, bool(dag), cb)
                got[("hop", dag, cb)] = fout.download()
    # the defaults the bench and the solvers run with at these shapes
    assert opts["OPT_TSTREAM"] == (16 if L[3] == 64 else 0)
    assert opts["OPT_BLOCK_TILE"] > 0 and opts["OPT_PARITY_ORDER"] == 1
    assert opts["OPT_XCD_REMAP"] == 1
    Vh = w.volume(L) // 2
    for key, val in got.items():
        if key[0] == "M":
            ref = oracle.dslash(b, Af, L, mass, key[1])
        else:
            ref = oracle.hop(b, Af, L, key[1], key[2])
            assert np.all(val.reshape(2, Vh * 12)[1 - key[2]] == 0)
        err = rel_l2(val, ref)
        print(f"{key}: rel L2 vs oracle {err:.3e}", flush=True)
        assert err < TOL_OP, key


RECONS = {"12real": 12, "temporal_gauge": 13}   # 13 = WD_RECON_12T, the bench's sloppy links


@pytest.mark.parametrize("recon", list(RECONS), ids=list(RECONS))
def test_config_32x32x32x64_mixed(w, oracle, recon):
    """Mixed precision at 32^3x64, kappa 0.135, reliable updates every residual
    drop of 1e-2: FP32 arithmetic, F16 spinors; 12-real links F32 on the full
    system and F16 on the Schur system, as stored or in temporal gauge."""
    L = [32, 32, 32, 64]
    mass = kappa_to_mass(0.135)
    A = w.gauge_near_identity(L, 20211025, 0.3)
    b = w.spinor_gaussian(L, 20211026)
    sols, reps = {}, {}
    with w.Context(L, bc=BC) as ctx:
        U = ctx.gauge(A)
        fb = ctx.field(data=b)
        x = ctx.field()
        fp = ctx.cg(x, fb, U, mass, max_iter=20000, tol=1e-10, tol_relative=True,
                    use_x0=False, check_every=16)
        print(f"fp64 CG: {fp['iterations']} iterations", flush=True)
        assert fp["converged"]
        sols["fp64"] = x.download()
        for name, gprec in (("mixed", w.F32), ("mixed_even_odd", w.F16),
                            ("mixed_bicgstab", w.F16)):
            Us = ctx.gauge(A, prec=gprec, recon=RECONS[recon])
            xm = ctx.field()
            if name == "mixed_bicgstab":
                res = ctx.bicgstab_mixed(xm, fb, U, Us, [mass], tol=1e-10, delta=1e-2)
            else:
                res = ctx.cg_mixed(xm, fb, U, Us, [mass], tol=1e-10, delta=1e-2,
                                   even_odd=name == "mixed_even_odd")
            s = res["systems"][0]
            print(f"{name}: {s['iterations']} iterations, {s['reliable_updates']} reliable "
                  f"updates, true rel residual {s['true_rel_rsd']:.3e}", flush=True)
            reps[name] = dict(s, status=res["status"])
            sols[name] = xm.download()
            del Us
    Af = _flipped(w, L, A)
    for name, s in reps.items():
        assert s["status"] == 0 and s["converged"], name
        assert s["true_rel_rsd"] <= 1e-10, name
        meas = _normal_residual(oracle, sols[name], b, Af, L, mass)
        err = rel_l2(sols[name], sols["fp64"])
        print(f"{name}: oracle residual {meas:.3e}, rel error vs FP64 CG {err:.3e}", flush=True)
        assert meas <= 1.05e-10, name
        assert err < 1e-7, name
    # the Schur solvers take fewer iterations than the full system
    assert reps["mixed_even_odd"]["iterations"] < reps["mixed"]["iterations"]
    assert reps["mixed_bicgstab"]["iterations"] < reps["mixed_even_odd"]["iterations"]


def _point_sources(w, L):
    rng = np.random.RandomState(20211027)
    site = [int(rng.randint(0, n)) for n in L]
    srcs, masses = [], []
    for k in KAPPAS:
        for s in range(4):
            for c in range(3):
                srcs.append(w.spinor_point(L, site, s, c))
                masses.append(kappa_to_mass(k))
    return site, srcs, masses


@pytest.mark.timeout(600)
@pytest.mark.parametrize("recon", list(RECONS), ids=list(RECONS))
def test_config_24x24x24x48_48_systems(w, oracle, recon):
    L = [24, 24, 24, 48]
    A = w.gauge_near_identity(L, 20211025, 0.3)
    site, srcs, masses = _point_sources(w, L)
    n = len(srcs)
    out = {}
    with w.Context(L, bc=BC) as ctx:
        U = ctx.gauge(A)
        Us = ctx.gauge(A, prec=w.F16, recon=RECONS[recon])
        fb = ctx.field(nrhs=n, data=srcs)
        for name in ("mixed_even_odd", "mixed_bicgstab"):
            X = ctx.field(nrhs=n)
            if name == "mixed_bicgstab":
                res = ctx.bicgstab_mixed(X, fb, U, Us, masses, tol=1e-10)
            else:
                res = ctx.cg_mixed(X, fb, U, Us, masses, tol=1e-10, even_odd=True)
            its = [s["iterations"] for s in res["systems"]]
            print(f"{name}: {res['seconds']:.3f} s, max iterations per kappa "
                  f"{[max(its[12 * k:12 * k + 12]) for k in range(4)]}", flush=True)
            out[name] = (res, [X.download(r) for r in range(n)])
            del X
        # one kappa (0.120) against the FP64 CG, system by system
        x64 = []
        for r in range(12):
            b1, x1 = ctx.field(data=srcs[r]), ctx.field()
            fp = ctx.cg(x1, b1, U, masses[r], max_iter=20000, tol=1e-10, tol_relative=True,
                        use_x0=False, check_every=16)
            assert fp["converged"]
            x64.append(x1.download())
    Af = _flipped(w, L, A)
    for name, (res, xs) in out.items():
        assert res["status"] == 0, name
        its = [s["iterations"] for s in res["systems"]]
        for s in res["systems"]:
            assert s["converged"] and s["true_rel_rsd"] <= 1e-10, name
        # per-system masking: the lighter kappas need more iterations and the
        # systems of a lighter kappa keep iterating after the heavier ones stopped
        per_k = [max(its[12 * k:12 * k + 12]) for k in range(4)]
        assert all(per_k[k] < per_k[k + 1] for k in range(3)), (name, per_k)
        for r in range(12):
            err = rel_l2(xs[r], x64[r])
            assert err < 1e-7, (name, r, err)
        # FP64 stop test with the oracle's operator: one system per kappa
        for k in range(4):
            r = 12 * k + 5
            meas = _normal_residual(oracle, xs[r], srcs[r], Af, L, masses[r])
            print(f"{name} system {r}: oracle residual {meas:.3e}", flush=True)
            assert meas <= 1.05e-10, (name, r)


@pytest.mark.timeout(900)
def test_config_48x48x48x96_12_systems(w, oracle):
    L = [48, 48, 48, 96]
    n = 12
    mass = kappa_to_mass(0.135)
    print("generating 48^3x96 links", flush=True)
    A = w.gauge_near_identity(L, 20211025, 0.3)
    with w.Context(L, bc=BC) as ctx:
        U = ctx.gauge(A)
        # the bench's sloppy links: F16 12-real in temporal gauge
        Us = ctx.gauge(A, prec=w.F16, recon=RECONS["temporal_gauge"])
        fb = ctx.field(nrhs=n)
        for r in range(n):
            fb.upload(w.spinor_gaussian(L, 20211026, r), r)
        X = ctx.field(nrhs=n)
        print("batched Schur CG of 12 systems", flush=True)
        res = ctx.cg_mixed(X, fb, U, Us, [mass] * n, tol=1e-10, even_odd=True)
        print(f"{res['seconds']:.2f} s, iterations "
              f"{[s['iterations'] for s in res['systems']]}", flush=True)
        xs = {r: X.download(r) for r in (0, n - 1)}
        ctx.release_workspace()
    assert res["status"] == 0
    for s in res["systems"]:
        assert s["converged"] and s["true_rel_rsd"] <= 1e-10
    Af = _flipped(w, L, A)
    del A
    for r, x in xs.items():
        b = w.spinor_gaussian(L, 20211026, r)
        meas = _normal_residual(oracle, x, b, Af, L, mass)
        print(f"system {r}: oracle residual {meas:.3e}", flush=True)
        assert meas <= 1.05e-10
