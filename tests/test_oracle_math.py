"""Pins the CPU oracle (oracle/wilson_oracle.c) by mathematics that does not
depend on its code: a dense gamma-matrix operator written from the problem
statement, the free-field plane-wave closed form, gamma5-hermiticity,
adjointness, gauge covariance, the zero-link case, a dense linear solve for
CG, and planted mistakes that these checks must catch.  CPU only.
"""
import numpy as np
import pytest

import dense_wilson as dw
from oracle_py import kappa_to_mass, rel_l2

TOL = 1e-13


def test_gamma_algebra():
    I4 = np.eye(4)
    for mu in range(4):
        g = dw.GAMMA[mu]
        assert np.allclose(g, g.conj().T)
        for nu in range(4):
            ac = g @ dw.GAMMA[nu] + dw.GAMMA[nu] @ g
            assert np.allclose(ac, 2 * I4 * (mu == nu))
    # gamma5 = gx gy gz gt = diag(1, 1, -1, -1) in the DeGrand-Rossi basis
    assert np.allclose(dw.GAMMA5, np.diag([1, 1, -1, -1]))


SHAPES = [[2, 2, 2, 2], [4, 4, 4, 4], [2, 4, 6, 4], [6, 2, 4, 2]]


@pytest.mark.parametrize("L", SHAPES, ids=lambda L: "x".join(map(str, L)))
@pytest.mark.parametrize("dag", [0, 1])
def test_oracle_equals_dense_operator(w, oracle, L, dag):
    A = w.gauge_random_su3(L, 99)
    b = w.spinor_gaussian(L, 100)
    mass = kappa_to_mass(0.13)
    ref = dw.apply_M_ref(b, A, L, mass, bool(dag))
    assert rel_l2(oracle.dslash(b, A, L, mass, dag), ref) < TOL
    # the hop alone, per output parity
    V = w.volume(L)
    hop_dense = ref - (4 + mass) * b
    for cb in (0, 1):
        h = oracle.hop(b, A, L, dag, cb).reshape(2, -1)
        assert rel_l2(h[cb], hop_dense.reshape(2, -1)[cb]) < TOL
        assert np.all(h[1 - cb] == 0)
    assert V % 2 == 0


def test_plane_wave_closed_form(w, oracle):
    """Unit links: M e^{ipx} u = [m + sum(1 - cos p) + i sum gamma_mu sin p_mu] e^{ipx} u."""
    L = [4, 6, 4, 8]
    A = w.gauge_unit(L)
    mass = 0.37
    rng = np.random.default_rng(5)
    Lx, Ly, Lz, Lt = L
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx),
                             indexing="ij")
    for _ in range(4):
        n = rng.integers(0, 8, size=4)
        p = 2 * np.pi * n / np.array(L)
        u = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
        phase = np.exp(1j * (p[0] * x + p[1] * y + p[2] * z + p[3] * t))
        psi = phase[..., None, None] * u
        Mp = (mass + np.sum(1 - np.cos(p))) * np.eye(4) + 1j * sum(
            dw.GAMMA[mu] * np.sin(p[mu]) for mu in range(4))
        expect = phase[..., None, None] * (Mp @ u)
        b = dw.spinor_from_lex(psi, L)
        got = oracle.dslash(b, A, L, mass, 0)
        assert rel_l2(got, dw.spinor_from_lex(expect, L)) < TOL


def test_antiperiodic_plane_wave(w, oracle):
    """Antiperiodic t (links negated on the last t slice) shifts p_t to (2n+1) pi / Lt."""
    L = [4, 4, 4, 6]
    A = w.gauge_unit(L)
    w.gauge_flip_boundary(L, 3, A)
    mass = -0.2
    Lx, Ly, Lz, Lt = L
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx),
                             indexing="ij")
    p = np.array([2 * np.pi / Lx, 0.0, 4 * np.pi / Lz, 3 * np.pi / Lt])
    u = np.arange(12).reshape(4, 3) + 1j
    phase = np.exp(1j * (p[0] * x + p[1] * y + p[2] * z + p[3] * t))
    Mp = (mass + np.sum(1 - np.cos(p))) * np.eye(4) + 1j * sum(
        dw.GAMMA[mu] * np.sin(p[mu]) for mu in range(4))
    got = oracle.dslash(dw.spinor_from_lex(phase[..., None, None] * u, L), A, L, mass, 0)
    assert rel_l2(got, dw.spinor_from_lex(phase[..., None, None] * (Mp @ u), L)) < TOL


def test_gamma5_hermiticity_and_adjointness(w, oracle):
    L = [4, 4, 4, 8]
    A = w.gauge_near_identity(L, 3, 0.3)
    a, b = w.spinor_gaussian(L, 1, 0), w.spinor_gaussian(L, 1, 1)
    mass = kappa_to_mass(0.135)

    def g5(v):
        s = v.reshape(-1, 4, 3).copy()
        s[:, 2:] *= -1
        return s.reshape(-1)

    Md = oracle.dslash(b, A, L, mass, 1)
    assert rel_l2(Md, g5(oracle.dslash(g5(b), A, L, mass, 0))) < TOL
    lhs = np.vdot(a, oracle.dslash(b, A, L, mass, 0))
    rhs = np.vdot(oracle.dslash(a, A, L, mass, 1), b)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_gauge_covariance(w, oracle):
    L = [4, 4, 6, 4]
    A = w.gauge_random_su3(L, 8)
    G = w.gauge_random_su3(L, 9)[0]  # one SU(3) matrix per site
    b = w.spinor_gaussian(L, 10)
    mass = 0.1
    Ul = dw.gauge_to_lex(A, L)
    g = dw.gauge_to_lex([G] * 4, L)[0]
    Up = []
    for mu in range(4):
        g_next = np.roll(g, -1, axis=dw.AXIS[mu])
        Up.append(g @ Ul[mu] @ np.conj(np.swapaxes(g_next, -1, -2)))
    psi = dw.spinor_to_lex(b, L)
    gpsi = np.einsum("...ab,...sb->...sa", g, psi)
    Ap = dw.gauge_from_lex(Up, L)
    lhs = oracle.dslash(dw.spinor_from_lex(gpsi, L), Ap, L, mass, 0)
    Mpsi = dw.spinor_to_lex(oracle.dslash(b, A, L, mass, 0), L)
    rhs = dw.spinor_from_lex(np.einsum("...ab,...sb->...sa", g, Mpsi), L)
    assert rel_l2(lhs, rhs) < 1e-12


def test_zero_links_and_diag(w, oracle):
    L = [4, 4, 4, 4]
    A = [np.zeros(w.volume(L) * 9, dtype=complex) for _ in range(4)]
    b = w.spinor_gaussian(L, 4)
    assert np.array_equal(oracle.dslash(b, A, L, 0.25, 0), 4.25 * b)
    assert np.array_equal(oracle.diag(b, L, -0.5), 3.5 * b)
    it, x, hist = oracle.cg(b, np.zeros_like(b), A, L, 0.25, 10, 1e-12)
    assert it == 1 and rel_l2(x, b / 4.25) < 1e-14


def test_norm_and_dot_closed_forms(oracle):
    n = 12 * 64
    v = (np.arange(n) % 7 - 3) + 1j * (np.arange(n) % 5 - 2)
    assert oracle.norm2(v) == float(np.sum(np.abs(v) ** 2))
    a = np.full(n, 1 + 2j)
    b = np.full(n, 3 - 1j)
    assert oracle.dot(a, b) == n * (1 - 2j) * (3 - 1j)


def test_cg_equals_dense_solve(w, oracle):
    L = [2, 2, 2, 4]
    A = w.gauge_random_su3(L, 21)
    b = w.spinor_gaussian(L, 22)
    mass = kappa_to_mass(0.12)
    M = dw.dense_matrix(A, L, mass)
    Md = dw.dense_matrix(A, L, mass, dag=True)
    assert np.allclose(Md, M.conj().T, atol=1e-14)
    x_dense = np.linalg.solve(M, b)
    it, x, hist = oracle.cg(b, np.zeros_like(b), A, L, mass, 500, 1e-13)
    assert it > 0
    assert rel_l2(x, x_dense) < 1e-11
    # CG on M^dag M: the residual history ends below the stop threshold
    assert hist[-1] < 1e-13 <= hist[-2]


@pytest.mark.parametrize("mistake", ["transposed_link", "y_gamma_sign", "dropped_t_backward"])
def test_planted_mistakes_are_caught(w, oracle, mistake):
    """A plausible slip in the operator must fail the dense comparison."""
    L = [4, 4, 4, 4]
    A = w.gauge_random_su3(L, 31)
    b = w.spinor_gaussian(L, 32)
    mass = 0.1
    good = oracle.dslash(b, A, L, mass, 0)
    psi, U = dw.spinor_to_lex(b, L), dw.gauge_to_lex(A, L)
    if mistake == "transposed_link":
        bad = dw.apply_M_lex(psi, [np.swapaxes(u, -1, -2) for u in U], mass)
    elif mistake == "y_gamma_sign":
        gam = list(dw.GAMMA)
        gam[1] = -gam[1]
        bad = dw.apply_M_lex(psi, U, mass, gamma=gam)
    else:
        Uz = list(U)
        full = dw.apply_M_lex(psi, Uz, mass)
        # remove the t-backward term by hand
        Ub = np.roll(U[3], 1, axis=0)
        bwd = np.einsum("...ba,...sb->...sa", Ub.conj(), np.roll(psi, 1, axis=0))
        bwd = bwd + np.einsum("st,...tb->...sb", dw.GAMMA[3], bwd)
        bad = full + 0.5 * bwd
    assert rel_l2(dw.spinor_from_lex(bad, L), good) > 1e-3
