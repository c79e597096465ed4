"""ctypes wrapper of oracle/liboracle.so (the CPU restatement of the reference path).

Test infrastructure only: the checker for the GPU results.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


class Oracle:
    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle oracle`")
        lib = ctypes.CDLL(path)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        ip = ctypes.POINTER(ctypes.c_int)
        lib.or_hop.argtypes = [vp, vp, ctypes.POINTER(vp), ip, i, i]
        lib.or_hop.restype = None
        lib.or_diag.argtypes = [vp, vp, ip, d]
        lib.or_diag.restype = None
        lib.or_dslash.argtypes = [vp, vp, ctypes.POINTER(vp), ip, d, i]
        lib.or_dslash.restype = None
        lib.or_norm2.argtypes = [vp, ctypes.c_long]
        lib.or_norm2.restype = d
        lib.or_dot.argtypes = [vp, vp, ctypes.c_long, vp]
        lib.or_dot.restype = None
        lib.or_cg.argtypes = [vp, vp, ctypes.POINTER(vp), ip, d, i, d, vp]
        lib.or_cg.restype = i
        lib.or_site2eoindex.argtypes = [ip, ip]
        lib.or_site2eoindex.restype = ctypes.c_long
        self.lib = lib

    @staticmethod
    def _L(L):
        return (ctypes.c_int * 4)(*[int(x) for x in L])

    @staticmethod
    def _U(A):
        return (ctypes.c_void_p * 4)(*[a.ctypes.data for a in A])

    def hop(self, src, A, L, dag, cb):
        out = np.empty_like(src)
        self.lib.or_hop(_p(src), _p(out), self._U(A), self._L(L), int(dag), int(cb))
        return out

    def diag(self, src, L, mass):
        out = np.empty_like(src)
        self.lib.or_diag(_p(src), _p(out), self._L(L), float(mass))
        return out

    def dslash(self, src, A, L, mass, dag):
        out = np.empty_like(src)
        self.lib.or_dslash(_p(src), _p(out), self._U(A), self._L(L), float(mass), int(dag))
        return out

    def norm2(self, v):
        return self.lib.or_norm2(_p(v), v.size)

    def dot(self, a, b):
        o = np.zeros(2)
        self.lib.or_dot(_p(a), _p(b), a.size, _p(o))
        return complex(o[0], o[1])

    def cg(self, b, x0, A, L, mass, max_iter, acc):
        x = np.array(x0, dtype=np.complex128, copy=True)
        hist = np.zeros(max(max_iter, 1))
        it = self.lib.or_cg(_p(b), _p(x), self._U(A), self._L(L), float(mass), int(max_iter),
                            float(acc), _p(hist))
        n = abs(it)
        return it, x, hist[:n]

    def eo(self, x, L):
        return self.lib.or_site2eoindex((ctypes.c_int * 4)(*x), self._L(L))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def kappa_to_mass(kappa):
    """m = 1/(2 kappa) - 4 (mass form of the operator, PDF eq. 2.4)."""
    return 1.0 / (2.0 * kappa) - 4.0


def libc_rand_x0(seed, n):
    """The reference's CG start vector: srand(seed); x0_i = rand()/RAND_MAX + 0i
    (src/invert.cpp:42-44), drawn from this process's C library."""
    lib = ctypes.CDLL(ORACLE_LIB)
    lib.or_rand_x0.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_uint]
    lib.or_rand_x0.restype = None
    x = np.empty(n, dtype=np.complex128)
    lib.or_rand_x0(_p(x), n, seed)
    return x
