"""Independent dense Wilson operator written from the problem statement.

M = (4+m) 1 - 1/2 sum_mu [(1 - f g_mu) U_mu(x) d_{x+mu,y} + (1 + f g_mu) U_mu^dag(x-mu) d_{x-mu,y}]
(doc/IPCC_2021.pdf eq. 2.4-2.6; f = +1 for M, -1 for M^dagger), with the
DeGrand-Rossi gamma matrices of eq. 2.7 and periodic boundaries.  It works on
lexicographic arrays psi[t, z, y, x, spin, colour] with np.roll and full 4x4
gamma matrices, i.e. none of the half-spinor tables of the oracle or the HIP
kernels, so it is a check on both.  Test infrastructure only.
"""
import numpy as np

GAMMA = [
    np.array([[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]]),  # x
    np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], dtype=complex),  # y
    np.array([[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]]),  # z
    np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=complex),  # t
]
GAMMA5 = GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3]

# lattice direction mu -> axis of a [t, z, y, x, ...] array
AXIS = {0: 3, 1: 2, 2: 1, 3: 0}


def eo_permutation(L):
    """perm[lex] = eo index (include/operator.h:56-67 of the reference)."""
    Lx, Ly, Lz, Lt = L
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx),
                             indexing="ij")
    lex = x + Lx * (y + Ly * (z + Lz * t))
    V = Lx * Ly * Lz * Lt
    eo = lex // 2 + ((x + y + z + t) % 2) * (V // 2)
    return eo.reshape(-1)


def spinor_to_lex(a, L):
    perm = eo_permutation(L)
    Lx, Ly, Lz, Lt = L
    return a.reshape(-1, 4, 3)[perm].reshape(Lt, Lz, Ly, Lx, 4, 3)


def spinor_from_lex(psi, L):
    perm = eo_permutation(L)
    out = np.empty((perm.size, 4, 3), dtype=complex)
    out[perm] = psi.reshape(-1, 4, 3)
    return out.reshape(-1)


def gauge_to_lex(A, L):
    perm = eo_permutation(L)
    Lx, Ly, Lz, Lt = L
    return [a.reshape(-1, 3, 3)[perm].reshape(Lt, Lz, Ly, Lx, 3, 3) for a in A]


def gauge_from_lex(U, L):
    perm = eo_permutation(L)
    out = []
    for u in U:
        a = np.empty((perm.size, 3, 3), dtype=complex)
        a[perm] = u.reshape(-1, 3, 3)
        out.append(a.reshape(-1))
    return out


def apply_M_lex(psi, U, mass, dag=False, gamma=GAMMA):
    f = -1.0 if dag else 1.0
    out = (4.0 + mass) * psi
    for mu in range(4):
        ax = AXIS[mu]
        g = gamma[mu]
        fwd = np.einsum("...ab,...sb->...sa", U[mu], np.roll(psi, -1, axis=ax))
        fwd = fwd - f * np.einsum("st,...tb->...sb", g, fwd)
        Ub = np.roll(U[mu], 1, axis=ax)
        bwd = np.einsum("...ba,...sb->...sa", Ub.conj(), np.roll(psi, 1, axis=ax))
        bwd = bwd + f * np.einsum("st,...tb->...sb", g, bwd)
        out = out - 0.5 * (fwd + bwd)
    return out


def apply_M_ref(a, A, L, mass, dag=False, gamma=GAMMA):
    """M on reference-layout arrays (spinor complex[V*12], links 4 x complex[V*9])."""
    return spinor_from_lex(apply_M_lex(spinor_to_lex(a, L), gauge_to_lex(A, L), mass, dag, gamma),
                           L)


def dense_matrix(A, L, mass, dag=False):
    n = int(np.prod(L)) * 12
    M = np.empty((n, n), dtype=complex)
    e = np.zeros(n, dtype=complex)
    for j in range(n):
        e[:] = 0
        e[j] = 1
        M[:, j] = apply_M_ref(e, A, L, mass, dag)
    return M
