"""Temporal gauge in NumPy, for the tests of WD_RECON_12T (test
infrastructure): Omega(x, y, z, t) = U_t(t = 0) U_t(t = 1) ... U_t(t - 1) along
each t column, the rotated links Omega(x) U_mu(x) Omega(x + mu)^dag, and the
rotation of spinors, all in the reference's even-odd layout (site e = parity *
V/2 + x/2 + Lx/2 (y + Ly (z + Lz t)), include/operator.h:56-67)."""
import numpy as np


def eo_index(L):
    """idx[x, y, z, t] = even-odd site index."""
    Lx, Ly, Lz, Lt = L
    x, y, z, t = np.meshgrid(*[np.arange(n) for n in L], indexing="ij")
    par = (x + y + z + t) & 1
    Vh = Lx * Ly * Lz * Lt // 2
    return par * Vh + x // 2 + (Lx // 2) * (y + Ly * (z + Lz * t))


def links(A):
    """Reference link arrays A[mu] (complex [V * 9]) -> [4, V, 3, 3]."""
    return np.stack([np.asarray(a).reshape(-1, 3, 3) for a in A])


def omega(A, L):
    """Omega per site, [V, 3, 3]."""
    idx = eo_index(L)
    U = links(A)
    V = int(np.prod(L))
    Om = np.zeros((V, 3, 3), dtype=np.complex128)
    W = np.broadcast_to(np.eye(3, dtype=np.complex128), idx.shape[:3] + (3, 3)).copy()
    for t in range(L[3]):
        e = idx[:, :, :, t]
        Om[e] = W
        W = W @ U[3][e]
    return Om


def rotate_links(A, L, Om=None):
    """Omega(x) U_mu(x) Omega(x + mu)^dag for every mu, as reference arrays."""
    if Om is None:
        Om = omega(A, L)
    idx = eo_index(L)
    U = links(A)
    out = []
    for mu in range(4):
        nb = np.roll(idx, -1, axis=mu)          # site x + mu
        e, f = idx.ravel(), nb.ravel()
        R = np.empty_like(U[mu])
        R[e] = Om[e] @ U[mu][e] @ np.conj(np.transpose(Om[f], (0, 2, 1)))
        out.append(R.reshape(-1))
    return out


def rotate_spinor(v, Om, dag=False):
    """(Omega v)(x) on every spin (dag: Omega^dag v)."""
    s = np.asarray(v).reshape(-1, 4, 3)
    M = np.conj(np.transpose(Om, (0, 2, 1))) if dag else Om
    return np.einsum("vab,vsb->vsa", M, s).reshape(-1)
