"""ctypes binding of libwilson.so (include/wilson.h, include/wilson_gen.h).

Host arrays use the reference layout: a spinor is complex128[V*12] indexed
eo(x)*12 + spin*3 + colour, a gauge field is four complex128[V*9] arrays indexed
eo(x)*9 + row*3 + col (include/operator.h:56-67 of the reference).

The library is the HIP build in this directory.  There is no CPU fallback:
loading fails loudly when libwilson.so is missing, and every compute call
needs a GPU.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwilson.so")

WD_OK = 0
WD_EARG = -1
WD_EHIP = -2
WD_ENOMEM = -3
WD_EBREAKDOWN = -4
WD_ENOTCONV = 1

F64, F32, F16 = 0, 1, 2
OP_HOP, OP_M = 0, 1
RHS_LOOP = 64   # rhs_per_thread: every system of a site in one thread (F16 spinors and links)
OPT_XCD_REMAP = 1
OPT_BLOCK_TILE = 2
OPT_PARITY_ORDER = 3
OPT_TSTREAM = 4
OPT_F64_PAIRS = 5
OPT_LOOP_PREFETCH = 6
OPT_LOOP_SYSTEMS = 7
LINKPROD_VALU, LINKPROD_MFMA = 0, 1
RECON_18, RECON_12 = 18, 12
# sloppy (F32, F16) links stored 12-real in temporal gauge (include/wilson.h WD_RECON_12T)
RECON_12T = 13
# autotuner candidates (include/wilson.h WD_TUNE_NR x WD_TUNE_NC)
TUNE_R = (1, 2, 4, RHS_LOOP)
TUNE_CHUNK = (0, 16, 32, 64)
TUNE_NR, TUNE_NC = len(TUNE_R), len(TUNE_CHUNK)

# Every symbol include/wilson.h and include/wilson_gen.h declare.
EXPORTED = [
    "wd_create", "wd_destroy", "wd_strerror", "wd_last_error", "wd_sync", "wd_volume",
    "wd_release_workspace", "wd_workspace_bytes",
    "wd_gauge_create", "wd_gauge_upload_ref", "wd_gauge_destroy",
    "wd_field_create", "wd_field_destroy", "wd_field_upload_ref", "wd_field_download_ref",
    "wd_field_zero", "wd_field_copy",
    "wd_hop", "wd_diag", "wd_apply_M", "wd_apply_MdagM",
    "wd_norm2", "wd_dot",
    "wd_cg_params_default", "wd_cg", "wd_cg_eo", "wd_time_op", "wd_set_option",
    "wd_get_option",
    "wd_mixed_params_default", "wd_cg_mixed", "wd_bicgstab_mixed", "wd_apply_M_sloppy",
    "wd_tune_report", "wd_time_sloppy", "wd_link_product",
    "wd_pion_correlator", "wd_pion_correlator_noalloc",
    "wdg_gauge_random_su3", "wdg_gauge_near_identity", "wdg_gauge_unit",
    "wdg_gauge_flip_boundary", "wdg_spinor_gaussian", "wdg_spinor_point", "wdg_philox4x32",
]


class WilsonError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


class CGParams(ctypes.Structure):
    _fields_ = [
        ("max_iter", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("tol_relative", ctypes.c_int),
        ("use_x0", ctypes.c_int),
        ("check_every", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("rsd_history", ctypes.POINTER(ctypes.c_double)),
    ]


class CGReport(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("rsd", ctypes.c_double),
        ("true_rsd", ctypes.c_double),
        ("rhs_norm", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("operator_applies", ctypes.c_long),
    ]


class MixedParams(ctypes.Structure):
    _fields_ = [
        ("max_iter", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("spinor_prec", ctypes.c_int),
        ("rhs_per_thread", ctypes.c_int),
        ("site_chunk", ctypes.c_int),
        ("check_every", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("even_odd", ctypes.c_int),
        ("reliable_every", ctypes.c_int),
    ]


class MixedReport(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("reliable_updates", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("true_rel_rsd", ctypes.c_double),
        ("rhs_norm", ctypes.c_double),
        ("rhs_per_thread", ctypes.c_int),
        ("site_chunk", ctypes.c_int),
        ("block_tile", ctypes.c_int),
    ]


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libwilson.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP library first (python -c "
            "'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p
    i, d = ctypes.c_int, ctypes.c_double
    sig = {
        "wd_create": ([ctypes.POINTER(vp), ip, ip, i], i),
        "wd_destroy": ([vp], None),
        "wd_strerror": ([i], ctypes.c_char_p),
        "wd_last_error": ([vp], ctypes.c_char_p),
        "wd_sync": ([vp], i),
        "wd_volume": ([vp], ctypes.c_long),
        "wd_gauge_create": ([vp, ctypes.POINTER(vp), i, i], i),
        "wd_gauge_upload_ref": ([vp, vp, ctypes.POINTER(dp)], i),
        "wd_gauge_destroy": ([vp], None),
        "wd_field_create": ([vp, ctypes.POINTER(vp), i, i], i),
        "wd_field_destroy": ([vp], None),
        "wd_field_upload_ref": ([vp, vp, dp, i], i),
        "wd_field_download_ref": ([vp, vp, dp, i], i),
        "wd_field_zero": ([vp, vp], i),
        "wd_field_copy": ([vp, vp, vp], i),
        "wd_hop": ([vp, vp, vp, vp, i, i], i),
        "wd_diag": ([vp, vp, vp, d], i),
        "wd_apply_M": ([vp, vp, vp, vp, d, i], i),
        "wd_apply_MdagM": ([vp, vp, vp, vp, vp, d], i),
        "wd_norm2": ([vp, vp, ctypes.POINTER(d)], i),
        "wd_dot": ([vp, vp, vp, dp], i),
        "wd_cg_params_default": ([ctypes.POINTER(CGParams)], None),
        "wd_cg": ([vp, vp, vp, vp, d, ctypes.POINTER(CGParams), ctypes.POINTER(CGReport)], i),
        "wd_cg_eo": ([vp, vp, vp, vp, d, ctypes.POINTER(CGParams), ctypes.POINTER(CGReport)], i),
        "wd_mixed_params_default": ([ctypes.POINTER(MixedParams)], None),
        "wd_cg_mixed": ([vp, vp, vp, vp, vp, dp, ctypes.POINTER(MixedParams),
                         ctypes.POINTER(MixedReport), ctypes.POINTER(d)], i),
        "wd_bicgstab_mixed": ([vp, vp, vp, vp, vp, dp, ctypes.POINTER(MixedParams),
                               ctypes.POINTER(MixedReport), ctypes.POINTER(d)], i),
        "wd_apply_M_sloppy": ([vp, vp, vp, vp, dp, i, i, i], i),
        "wd_tune_report": ([vp, vp, vp, ctypes.POINTER(i), ctypes.POINTER(i), dp], i),
        "wd_time_sloppy": ([vp, vp, vp, vp, dp, i, i, i, i, ctypes.POINTER(d)], i),
        "wd_link_product": ([vp, i, i, i, vp, vp, i, i, vp, ctypes.POINTER(d)], i),
        "wd_time_op": ([vp, i, vp, vp, vp, d, i, i, i, ctypes.POINTER(d)], i),
        "wd_pion_correlator": ([vp, vp, i, i, i, dp], i),
        "wd_pion_correlator_noalloc": ([vp, vp, i, i, i, dp], i),
        "wd_set_option": ([vp, i, i], i),
        "wd_get_option": ([vp, i, ctypes.POINTER(ctypes.c_int)], i),
        "wd_release_workspace": ([vp], i),
        "wd_workspace_bytes": ([vp, ctypes.POINTER(ctypes.c_size_t)], i),
        "wdg_gauge_random_su3": ([ip, ctypes.c_uint64, ctypes.POINTER(dp)], None),
        "wdg_gauge_near_identity": ([ip, ctypes.c_uint64, d, ctypes.POINTER(dp)], None),
        "wdg_gauge_unit": ([ip, ctypes.POINTER(dp)], None),
        "wdg_gauge_flip_boundary": ([ip, i, ctypes.POINTER(dp)], None),
        "wdg_spinor_gaussian": ([ip, ctypes.c_uint64, ctypes.c_uint32, dp], None),
        "wdg_spinor_point": ([ip, ip, i, i, dp], None),
        "wdg_philox4x32": ([ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                            ctypes.POINTER(ctypes.c_uint32)], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _ivec(v: Sequence[int]):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def _gauge_ptrs(A: Sequence[np.ndarray]):
    return (ctypes.c_void_p * 4)(*[a.ctypes.data for a in A])


def volume(L: Sequence[int]) -> int:
    return int(np.prod([int(x) for x in L]))


# ---- generators (host code of libwilson.so) ---------------------------------
def gauge_random_su3(L, seed: int):
    lib = load_library()
    A = [np.empty(volume(L) * 9, dtype=np.complex128) for _ in range(4)]
    lib.wdg_gauge_random_su3(_ivec(L), seed, _gauge_ptrs(A))
    return A


def gauge_near_identity(L, seed: int, eps: float = 0.3):
    lib = load_library()
    A = [np.empty(volume(L) * 9, dtype=np.complex128) for _ in range(4)]
    lib.wdg_gauge_near_identity(_ivec(L), seed, eps, _gauge_ptrs(A))
    return A


def gauge_unit(L):
    lib = load_library()
    A = [np.empty(volume(L) * 9, dtype=np.complex128) for _ in range(4)]
    lib.wdg_gauge_unit(_ivec(L), _gauge_ptrs(A))
    return A


def gauge_flip_boundary(L, mu: int, A):
    """In place: U_mu on the last mu-slice times -1 (antiperiodic in mu)."""
    lib = load_library()
    lib.wdg_gauge_flip_boundary(_ivec(L), mu, _gauge_ptrs(A))
    return A


def spinor_gaussian(L, seed: int, stream: int = 0):
    lib = load_library()
    a = np.empty(volume(L) * 12, dtype=np.complex128)
    lib.wdg_spinor_gaussian(_ivec(L), seed, stream, _ptr(a))
    return a


def spinor_point(L, x, spin: int, colour: int):
    lib = load_library()
    a = np.empty(volume(L) * 12, dtype=np.complex128)
    lib.wdg_spinor_point(_ivec(L), _ivec(x), spin, colour, _ptr(a))
    return a


def philox4x32(counter, key):
    lib = load_library()
    c = (ctypes.c_uint32 * 4)(*counter)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib.wdg_philox4x32(c, k, o)
    return list(o)


# ---- device objects ----------------------------------------------------------
class Context:
    """One device + one stream (wd_ctx)."""

    def __init__(self, L, bc=(1, 1, 1, 1), device: int = 0):
        self.lib = load_library()
        self.L = [int(x) for x in L]
        self.bc = [int(x) for x in bc]
        self.V = volume(self.L)
        h = ctypes.c_void_p()
        rc = self.lib.wd_create(ctypes.byref(h), _ivec(self.L), _ivec(self.bc), device)
        if rc != WD_OK:
            raise WilsonError(rc, f"wd_create failed: {self.lib.wd_strerror(rc).decode()}")
        self.h = h

    def check(self, rc: int, what: str):
        if rc < 0:
            raise WilsonError(rc, f"{what}: {self.lib.wd_last_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.wd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        self.check(self.lib.wd_sync(self.h), "wd_sync")

    # objects
    def gauge(self, A=None, prec: int = F64, recon: int = RECON_18) -> "Gauge":
        g = Gauge(self, prec, recon)
        if A is not None:
            g.upload(A)
        return g

    def field(self, prec: int = F64, nrhs: int = 1, data=None) -> "Field":
        f = Field(self, prec, nrhs)
        if data is not None:
            if nrhs == 1:
                f.upload(data)
            else:
                for r in range(nrhs):
                    f.upload(data[r], r)
        return f

    # operator
    def hop(self, out, inp, U, dagger: bool, cb: int):
        self.check(self.lib.wd_hop(self.h, out.h, inp.h, U.h, int(dagger), cb), "wd_hop")

    def diag(self, out, inp, mass: float):
        self.check(self.lib.wd_diag(self.h, out.h, inp.h, mass), "wd_diag")

    def apply_M(self, out, inp, U, mass: float, dagger: bool):
        self.check(self.lib.wd_apply_M(self.h, out.h, inp.h, U.h, mass, int(dagger)),
                   "wd_apply_M")

    def apply_MdagM(self, out, inp, tmp, U, mass: float):
        self.check(self.lib.wd_apply_MdagM(self.h, out.h, inp.h, tmp.h, U.h, mass),
                   "wd_apply_MdagM")

    def time_op(self, op: int, out, inp, U, mass: float = 0.0, dagger: bool = False,
                cb: int = 0, iters: int = 10) -> float:
        """Average launch duration in ms (HIP events on the context stream)."""
        ms = ctypes.c_double()
        self.check(self.lib.wd_time_op(self.h, op, out.h, inp.h, U.h, mass, int(dagger), cb,
                                       iters, ctypes.byref(ms)), "wd_time_op")
        return ms.value

    def apply_M_sloppy(self, out, inp, Us, masses, dagger: bool = False, rhs_per_thread: int = 1,
                       site_chunk: int = 0):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        assert m.size == inp.nrhs
        self.check(self.lib.wd_apply_M_sloppy(self.h, out.h, inp.h, Us.h, _ptr(m), int(dagger),
                                              rhs_per_thread, site_chunk), "wd_apply_M_sloppy")

    def time_sloppy(self, out, inp, Us, masses, rhs_per_thread: int = 1, site_chunk: int = 0,
                    cb: int = -1, iters: int = 10) -> float:
        """Average launch duration in ms of the sloppy M (cb = -1) or its
        one-parity launch (cb = 0, 1), HIP events on the context stream."""
        m = np.ascontiguousarray(masses, dtype=np.float64)
        assert m.size == inp.nrhs
        ms = ctypes.c_double()
        self.check(self.lib.wd_time_sloppy(self.h, out.h, inp.h, Us.h, _ptr(m), rhs_per_thread,
                                           site_chunk, cb, iters, ctypes.byref(ms)),
                   "wd_time_sloppy")
        return ms.value

    def link_product(self, impl: int, U: np.ndarray, h: np.ndarray, inner: int = 1,
                     iters: int = 1):
        """MFMA-vs-VALU link-product experiment (wd_link_product).  U: complex
        [nsites, 3, 3], h: complex [nsites, nsys, 2, 3] (rounded to F16).
        Returns (U h summed over `inner` repetitions as complex64 [nsites,
        nsys, 2, 3], average launch ms)."""
        nsites, nsys = h.shape[0], h.shape[1]
        assert U.shape == (nsites, 3, 3) and h.shape == (nsites, nsys, 2, 3)

        def pack(z):
            re = np.ascontiguousarray(z.real, dtype=np.float16).view(np.uint16).astype(np.uint32)
            im = np.ascontiguousarray(z.imag, dtype=np.float16).view(np.uint16).astype(np.uint32)
            return np.ascontiguousarray(re | (im << 16))
        Up, hp = pack(U), pack(h)
        out = np.empty((nsites, nsys, 2, 3, 2), dtype=np.float32)
        ms = ctypes.c_double()
        self.check(self.lib.wd_link_product(self.h, impl, nsites, nsys, _ptr(Up), _ptr(hp), inner,
                                            iters, _ptr(out), ctypes.byref(ms)),
                   "wd_link_product")
        return out[..., 0] + 1j * out[..., 1], ms.value

    def tune_report(self, f, Us):
        """Times every (rhs per thread, site chunk) candidate of the sloppy M over
        f's systems; returns (best_R, best_chunk, {(R, chunk): ms}) without the
        candidates that were not timed."""
        best, chunk = ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(TUNE_NR * TUNE_NC)
        self.check(self.lib.wd_tune_report(self.h, f.h, Us.h, ctypes.byref(best),
                                           ctypes.byref(chunk), _ptr(ms)), "wd_tune_report")
        table = {(R, C): ms[ir * TUNE_NC + ic]
                 for ir, R in enumerate(TUNE_R) for ic, C in enumerate(TUNE_CHUNK)
                 if ms[ir * TUNE_NC + ic] >= 0}
        return best.value, chunk.value, table

    def cg_mixed(self, x, b, U, Us, masses, tol: float = 1e-10, delta: float = 2e-2,
                 spinor_prec: int = F16, rhs_per_thread: int = 0, site_chunk: int = -1,
                 max_iter: int = 100000, check_every: int = 16, use_graph: bool = True,
                 even_odd: bool = False, reliable_every: int = 0):
        return self._mixed(self.lib.wd_cg_mixed, "wd_cg_mixed", x, b, U, Us, masses, tol, delta,
                           spinor_prec, rhs_per_thread, site_chunk, max_iter, check_every,
                           use_graph, even_odd, reliable_every)

    def bicgstab_mixed(self, x, b, U, Us, masses, tol: float = 1e-10, delta: float = 2e-2,
                       spinor_prec: int = F16, rhs_per_thread: int = 0, site_chunk: int = -1,
                       max_iter: int = 100000, check_every: int = 16, use_graph: bool = True,
                       reliable_every: int = 0):
        """wd_bicgstab_mixed: BiCGStab on the Schur system of every system."""
        return self._mixed(self.lib.wd_bicgstab_mixed, "wd_bicgstab_mixed", x, b, U, Us, masses,
                           tol, delta, spinor_prec, rhs_per_thread, site_chunk, max_iter,
                           check_every, use_graph, True, reliable_every)

    def _mixed(self, fn, name, x, b, U, Us, masses, tol, delta, spinor_prec, rhs_per_thread,
               site_chunk, max_iter, check_every, use_graph, even_odd, reliable_every):
        p = MixedParams()
        self.lib.wd_mixed_params_default(ctypes.byref(p))
        p.max_iter, p.tol, p.delta = max_iter, tol, delta
        p.spinor_prec, p.rhs_per_thread, p.site_chunk = spinor_prec, rhs_per_thread, site_chunk
        p.check_every, p.use_graph, p.even_odd = check_every, int(use_graph), int(even_odd)
        p.reliable_every = reliable_every
        m = np.ascontiguousarray(masses, dtype=np.float64)
        assert m.size == b.nrhs
        reps = (MixedReport * b.nrhs)()
        sec = ctypes.c_double()
        rc = fn(self.h, x.h, b.h, U.h, Us.h, _ptr(m), ctypes.byref(p), reps, ctypes.byref(sec))
        self.check(rc, name)
        return {
            "status": rc,
            "seconds": sec.value,
            "systems": [{"iterations": r.iterations, "reliable_updates": r.reliable_updates,
                         "converged": bool(r.converged), "true_rel_rsd": r.true_rel_rsd,
                         "rhs_norm": r.rhs_norm, "rhs_per_thread": r.rhs_per_thread,
                         "site_chunk": r.site_chunk, "block_tile": r.block_tile}
                        for r in reps],
        }

    def pion_correlator(self, S, t0: int, rhs_first: int = 0, rhs_count: Optional[int] = None):
        """C(dt), dt = 0 .. Lt-1, of the point-source propagators S[rhs_first :
        rhs_first + rhs_count] (the 12 spin x colour solutions of one source)."""
        if rhs_count is None:
            rhs_count = S.nrhs - rhs_first
        Ct = np.zeros(self.L[3])
        self.check(self.lib.wd_pion_correlator(self.h, S.h, rhs_first, rhs_count, t0, _ptr(Ct)),
                   "wd_pion_correlator")
        return Ct

    def pion_correlator_noalloc(self, S, t0: int, rhs_first: int = 0,
                                rhs_count: Optional[int] = None):
        """As pion_correlator, through wd_pion_correlator_noalloc (no allocation;
        rhs_count * 2 * Lt <= 16384, Lt <= 512)."""
        if rhs_count is None:
            rhs_count = S.nrhs - rhs_first
        Ct = np.zeros(self.L[3])
        self.check(self.lib.wd_pion_correlator_noalloc(self.h, S.h, rhs_first, rhs_count, t0,
                                                       _ptr(Ct)), "wd_pion_correlator_noalloc")
        return Ct

    def set_option(self, option: int, value: int):
        self.check(self.lib.wd_set_option(self.h, option, value), "wd_set_option")

    def get_option(self, option: int) -> int:
        v = ctypes.c_int()
        self.check(self.lib.wd_get_option(self.h, option, ctypes.byref(v)), "wd_get_option")
        return v.value

    def release_workspace(self):
        """wd_release_workspace: free the solver work memory cached by the context."""
        self.check(self.lib.wd_release_workspace(self.h), "wd_release_workspace")

    def workspace_bytes(self) -> int:
        v = ctypes.c_size_t()
        self.check(self.lib.wd_workspace_bytes(self.h, ctypes.byref(v)), "wd_workspace_bytes")
        return v.value

    def norm2(self, f) -> float:
        v = ctypes.c_double()
        self.check(self.lib.wd_norm2(self.h, f.h, ctypes.byref(v)), "wd_norm2")
        return v.value

    def dot(self, a, b) -> complex:
        o = np.zeros(2)
        self.check(self.lib.wd_dot(self.h, a.h, b.h, _ptr(o)), "wd_dot")
        return complex(o[0], o[1])

    def cg(self, x, b, U, mass: float, max_iter: int = 500, tol: float = 1e-9,
           tol_relative: bool = False, use_x0: bool = True, check_every: int = 8,
           use_graph: bool = True, history: bool = True, even_odd: bool = False):
        """CGNR on the full M^dag M (wd_cg, the reference's algorithm) or, with
        even_odd, on the Schur complement of the even sites (wd_cg_eo)."""
        p = CGParams()
        self.lib.wd_cg_params_default(ctypes.byref(p))
        p.max_iter, p.tol, p.tol_relative = max_iter, tol, int(tol_relative)
        p.use_x0, p.check_every, p.use_graph = int(use_x0), check_every, int(use_graph)
        hist = np.zeros(max(max_iter, 1))
        if history:
            p.rsd_history = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rep = CGReport()
        fn, name = ((self.lib.wd_cg_eo, "wd_cg_eo") if even_odd else (self.lib.wd_cg, "wd_cg"))
        rc = fn(self.h, x.h, b.h, U.h, mass, ctypes.byref(p), ctypes.byref(rep))
        self.check(rc, name)
        return {
            "status": rc,
            "iterations": rep.iterations,
            "converged": bool(rep.converged),
            "rsd": rep.rsd,
            "true_rsd": rep.true_rsd,
            "rhs_norm": rep.rhs_norm,
            "seconds": rep.seconds,
            "operator_applies": rep.operator_applies,
            "history": hist[: rep.iterations].copy(),
        }


class Gauge:
    def __init__(self, ctx: Context, prec: int, recon: int):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.wd_gauge_create(ctx.h, ctypes.byref(h), prec, recon), "wd_gauge_create")
        self.h = h
        self.prec, self.recon = prec, recon

    def upload(self, A):
        A = [np.ascontiguousarray(a, dtype=np.complex128) for a in A]
        for a in A:
            assert a.size == self.ctx.V * 9
        self.ctx.check(self.ctx.lib.wd_gauge_upload_ref(self.ctx.h, self.h, _gauge_ptrs(A)),
                       "wd_gauge_upload_ref")

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.wd_gauge_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field:
    def __init__(self, ctx: Context, prec: int, nrhs: int):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.wd_field_create(ctx.h, ctypes.byref(h), prec, nrhs), "wd_field_create")
        self.h = h
        self.prec, self.nrhs = prec, nrhs

    def upload(self, a: np.ndarray, rhs: int = 0):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.size == self.ctx.V * 12
        self.ctx.check(self.ctx.lib.wd_field_upload_ref(self.ctx.h, self.h, _ptr(a), rhs),
                       "wd_field_upload_ref")

    def download(self, rhs: int = 0) -> np.ndarray:
        a = np.empty(self.ctx.V * 12, dtype=np.complex128)
        self.ctx.check(self.ctx.lib.wd_field_download_ref(self.ctx.h, self.h, _ptr(a), rhs),
                       "wd_field_download_ref")
        return a

    def zero(self):
        self.ctx.check(self.ctx.lib.wd_field_zero(self.ctx.h, self.h), "wd_field_zero")

    def copy_from(self, src: "Field"):
        self.ctx.check(self.ctx.lib.wd_field_copy(self.ctx.h, self.h, src.h), "wd_field_copy")

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.wd_field_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
