"""The NumPy temporal gauge against what the mathematics fixes (CPU): the
rotated t links are the identity off the last slice and the last slice holds
the Polyakov line; rotated links stay in SU(3); and the oracle's operator on
the rotated links is Omega M Omega^dag (M'(Omega x) = Omega M x to rounding,
also with the antiperiodic t boundary as negated links of the last slice)."""
import numpy as np
import pytest

from oracle_py import kappa_to_mass, rel_l2
import temporal_gauge as tgm


@pytest.fixture(scope="module")
def case(w):
    L = [4, 4, 6, 8]
    return L, w.gauge_random_su3(L, 20211101), w.spinor_gaussian(L, 20211102)


def test_t_links_are_unit_off_the_last_slice(case):
    L, A, _ = case
    R = tgm.links(tgm.rotate_links(A, L))
    idx = tgm.eo_index(L)
    I = np.eye(3)
    for t in range(L[3] - 1):
        assert np.abs(R[3][idx[:, :, :, t].ravel()] - I).max() < 1e-13
    # last slice: U_t(0) ... U_t(Lt - 1), the Polyakov line of the column
    U = tgm.links(A)
    P = np.broadcast_to(I, idx.shape[:3] + (3, 3)).copy()
    for t in range(L[3]):
        P = P @ U[3][idx[:, :, :, t]]
    assert np.abs(R[3][idx[:, :, :, -1]] - P).max() < 1e-13
    assert np.abs(P - I).max() > 0.1   # random links: a non-trivial check


def test_rotated_links_are_su3(case):
    L, A, _ = case
    for R in tgm.links(tgm.rotate_links(A, L)):
        RR = R @ np.conj(np.transpose(R, (0, 2, 1)))
        assert np.abs(RR - np.eye(3)).max() < 1e-13
        assert np.abs(np.linalg.det(R) - 1).max() < 1e-13


@pytest.mark.parametrize("dag", [0, 1])
@pytest.mark.parametrize("antiperiodic", [False, True])
def test_oracle_operator_is_covariant(w, oracle, case, dag, antiperiodic):
    L, A, x = case
    mass = kappa_to_mass(0.13)
    Om = tgm.omega(A, L)
    R = tgm.rotate_links(A, L, Om)
    if antiperiodic:
        A = w.gauge_flip_boundary(L, 3, [a.copy() for a in A])
        R = w.gauge_flip_boundary(L, 3, [r.copy() for r in R])
    lhs = oracle.dslash(tgm.rotate_spinor(x, Om), R, L, mass, dag)
    rhs = tgm.rotate_spinor(oracle.dslash(x, A, L, mass, dag), Om)
    assert rel_l2(lhs, rhs) < 1e-13
    # and the rotation is not the identity
    assert rel_l2(tgm.rotate_spinor(x, Om), x) > 0.1
