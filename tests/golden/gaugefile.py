"""Writes gauge fields in the reference's file format (test infrastructure).

Format read by the reference's load_gauge (src/load_gauge.cpp:110-161, the
NO_MPI_IO path): raw big-endian FP64, sites in lexicographic order (x fastest),
72 doubles per site = [mu][row][col][re, im].
"""
import numpy as np


def lex_to_eo(L):
    Lx, Ly, Lz, Lt = L
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx),
                             indexing="ij")
    lex = (x + Lx * (y + Ly * (z + Lz * t))).reshape(-1)
    V = Lx * Ly * Lz * Lt
    eo = (lex // 2 + ((x + y + z + t) % 2).reshape(-1) * (V // 2))
    out = np.empty(V, dtype=np.int64)
    out[lex] = eo
    return out


def write_gauge_file(path, A, L):
    """A: four complex128[V*9] arrays in the reference's even-odd layout."""
    eo = lex_to_eo(L)
    V = eo.size
    U4 = np.empty((V, 4, 9), dtype=np.complex128)
    for mu in range(4):
        U4[:, mu, :] = A[mu].reshape(-1, 9)[eo]
    data = np.empty((V, 4, 9, 2), dtype=">f8")
    data[..., 0] = U4.real
    data[..., 1] = U4.imag
    with open(path, "wb") as f:
        f.write(data.tobytes())
    return path
