"""Gauge configuration files (NERSC, ILDG) and the gauge observables, host code
of libipcc_api.so (host/include/gauge_io.h); no GPU needed.

No configuration file ships with the reference, so the readers are pinned by
round trips through the writers (every storage variant: 2 or 3 rows, 32 or
64 bit, either byte order), by the reference writer's labelling quirk (3x3
payloads marked 4D_SU3_GAUGE), by the header checks (plaquette, link trace,
dimensions, truncation, checksum), and the plaquette by mathematics: unit
links (1), a U(1) flux through the xy planes (closed form), invariance under
random gauge transformations, and an independent NumPy evaluation.
"""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
API = os.path.join(ROOT, "ipcc2021-final_amd", "libipcc_api.so")
NERSC, ILDG = 0, 1
OK, EOPEN, EFORMAT, EDIMS, ESHORT, ECHECK, ECHECKSUM = 0, -1, -2, -3, -4, -5, -6


@pytest.fixture(scope="module")
def api():
    if not os.path.exists(API):
        pytest.skip("libipcc_api.so not built")
    lib = ctypes.CDLL(API)
    dpp = ctypes.POINTER(ctypes.c_void_p)
    lib.ipcc_gauge_read.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int]
    lib.ipcc_gauge_write.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.ipcc_gauge_plaquette.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.ipcc_gauge_plaquette.restype = ctypes.c_double
    lib.ipcc_gauge_link_trace.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.ipcc_gauge_link_trace.restype = ctypes.c_double
    return lib


def _L(L):
    return (ctypes.c_int * 4)(*L)


def _ptrs(A):
    return (ctypes.c_void_p * 4)(*[a.ctypes.data for a in A])


def write(api, fmt, path, L, A, rows=3, bpr=8, big=True):
    return api.ipcc_gauge_write(fmt, path.encode(), _L(L), _ptrs(A), rows, bpr, int(big))


def read(api, fmt, path, L, check_checksum=False):
    A = [np.zeros(int(np.prod(L)) * 9, dtype=np.complex128) for _ in range(4)]
    vals = np.zeros(4)
    meta = (ctypes.c_int * 8)()
    rc = api.ipcc_gauge_read(fmt, path.encode(), _L(L), _ptrs(A), vals.ctypes.data, meta,
                             int(check_checksum))
    return rc, A, vals, list(meta)


def plaquette(api, L, A):
    return api.ipcc_gauge_plaquette(_L(L), _ptrs(A))


# ---- NumPy gauge geometry (eo layout of include/operator.h:56-67) ----------------------
def _geometry(L):
    V = int(np.prod(L))
    lex = np.arange(V)
    x = [lex % L[0], (lex // L[0]) % L[1], (lex // (L[0] * L[1])) % L[2], lex // (L[0] * L[1] * L[2])]
    eo = lex // 2 + ((x[0] + x[1] + x[2] + x[3]) % 2) * (V // 2)
    return x, eo


def _eo_shift(L, mu):
    """eo index of x + mu for every site, ordered by the eo index of x."""
    x, eo = _geometry(L)
    y = [c.copy() for c in x]
    y[mu] = (y[mu] + 1) % L[mu]
    lex_y = y[0] + L[0] * (y[1] + L[1] * (y[2] + L[2] * y[3]))
    out = np.empty_like(eo)
    out[eo] = eo[lex_y]
    return out


def numpy_plaquette(L, A):
    V = int(np.prod(L))
    U = [a.reshape(V, 3, 3) for a in A]
    s = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            sm, sn = _eo_shift(L, mu), _eo_shift(L, nu)
            a = U[mu] @ U[nu][sm]
            b = U[nu] @ U[mu][sn]
            s += np.sum(np.real(np.einsum("vij,vij->v", a, np.conj(b))))
    return s / (18.0 * V)


def _su3(w, L, seed):
    return w.gauge_random_su3(L, seed)


# ---- observables --------------------------------------------------------------------------
def test_plaquette_unit_and_flux_closed_forms(api, w):
    L = [6, 4, 4, 4]
    A = w.gauge_unit(L)
    assert plaquette(api, L, A) == pytest.approx(1.0, abs=1e-15)
    assert api.ipcc_gauge_link_trace(_L(L), _ptrs(A)) == pytest.approx(1.0, abs=1e-15)
    # U_y(x) = diag(e^{i phi x0}, e^{-i phi x0}, 1), phi = 2 pi / L0: every xy plaquette is
    # diag(e^{i phi}, e^{-i phi}, 1), every other plaquette 1
    phi = 2 * np.pi / L[0]
    x, eo = _geometry(L)
    Uy = np.zeros((int(np.prod(L)), 3, 3), dtype=np.complex128)
    Uy[eo, 0, 0] = np.exp(1j * phi * x[0])
    Uy[eo, 1, 1] = np.exp(-1j * phi * x[0])
    Uy[eo, 2, 2] = 1.0
    A[1] = Uy.reshape(-1)
    expect = (5.0 + (2.0 * np.cos(phi) + 1.0) / 3.0) / 6.0
    assert plaquette(api, L, A) == pytest.approx(expect, abs=1e-14)


def test_plaquette_gauge_invariant_and_matches_numpy(api, w):
    L = [4, 4, 6, 4]
    A = _su3(w, L, 7)
    p0 = plaquette(api, L, A)
    assert p0 == pytest.approx(numpy_plaquette(L, A), abs=1e-13)
    V = int(np.prod(L))
    g = _su3(w, L, 8)[0].reshape(V, 3, 3)          # one SU(3) matrix per site
    B = []
    for mu in range(4):
        U = A[mu].reshape(V, 3, 3)
        gs = g[_eo_shift(L, mu)]
        B.append((g @ U @ np.conj(np.transpose(gs, (0, 2, 1)))).reshape(-1).copy())
    assert plaquette(api, L, B) == pytest.approx(p0, abs=1e-13)
    assert abs(api.ipcc_gauge_link_trace(_L(L), _ptrs(B)) -
               api.ipcc_gauge_link_trace(_L(L), _ptrs(A))) > 1e-3  # not invariant


# ---- NERSC ----------------------------------------------------------------------------------
@pytest.mark.parametrize("rows", [2, 3])
@pytest.mark.parametrize("bpr", [4, 8])
@pytest.mark.parametrize("big", [True, False])
def test_nersc_round_trip(api, w, tmp_path, rows, bpr, big):
    L = [4, 4, 4, 6]
    A = _su3(w, L, 11)
    path = str(tmp_path / "conf.nersc")
    assert write(api, NERSC, path, L, A, rows, bpr, big) == OK
    rc, B, vals, meta = read(api, NERSC, path, L, check_checksum=True)
    assert rc == OK
    assert meta[:3] == [rows, bpr, int(big)] and meta[3] and meta[4] and meta[5]
    assert meta[6] == meta[7]  # checksum of the payload = the header's
    tol = 1e-14 if bpr == 8 else 2e-6
    for a, b in zip(A, B):
        assert np.max(np.abs(a - b)) < tol
    if bpr == 8 and rows == 3:
        assert all(np.array_equal(a, b) for a, b in zip(A, B))
    assert vals[0] == pytest.approx(plaquette(api, L, A), abs=1e-14 if bpr == 8 else 1e-6)


def test_nersc_reference_writer_labels(api, w, tmp_path):
    """The reference's writer labels 3x3 payloads 4D_SU3_GAUGE
    (refs/utils_nersc_lat.cpp:100, :135); the reader goes by the payload size."""
    L = [4, 4, 4, 4]
    A = _su3(w, L, 12)
    path = str(tmp_path / "ref_style.nersc")
    assert write(api, NERSC, path, L, A, 3, 8, True) == OK
    data = open(path, "rb").read().replace(b"4D_SU3_GAUGE_3x3", b"4D_SU3_GAUGE")
    open(path, "wb").write(data)
    rc, B, _, meta = read(api, NERSC, path, L)
    assert rc == OK and meta[0] == 3
    assert all(np.array_equal(a, b) for a, b in zip(A, B))


def test_nersc_header_checks(api, w, tmp_path):
    L = [4, 4, 4, 4]
    A = _su3(w, L, 13)
    path = str(tmp_path / "c.nersc")
    assert write(api, NERSC, path, L, A, 2, 8, True) == OK
    good = open(path, "rb").read()
    head, body = good.split(b"END_HEADER\n", 1)

    def variant(h, b=body):
        p = str(tmp_path / "v.nersc")
        open(p, "wb").write(h + b"END_HEADER\n" + b)
        return p
    lines = head.split(b"\n")
    bad_plaq = b"\n".join(l if not l.startswith(b"PLAQUETTE") else b"PLAQUETTE = 0.123" for l in lines)
    assert read(api, NERSC, variant(bad_plaq), L)[0] == ECHECK
    bad_sum = b"\n".join(l if not l.startswith(b"CHECKSUM") else b"CHECKSUM = 452c11d8" for l in lines)
    assert read(api, NERSC, variant(bad_sum), L)[0] == OK          # not checked by default
    assert read(api, NERSC, variant(bad_sum), L, check_checksum=True)[0] == ECHECKSUM
    assert read(api, NERSC, variant(head, body[:-1000]), L)[0] == ESHORT
    assert read(api, NERSC, path, [4, 4, 4, 6])[0] == EDIMS
    assert read(api, NERSC, str(tmp_path / "missing"), L)[0] == EOPEN
    assert read(api, NERSC, variant(b"NOT A HEADER"), L)[0] == EFORMAT


# ---- ILDG -------------------------------------------------------------------------------------
def _lime_record(rtype, payload, mb, me):
    h = bytearray(144)
    h[0:4] = (0x456789AB).to_bytes(4, "big")
    h[4:6] = (1).to_bytes(2, "big")
    h[6] = (0x80 if mb else 0) | (0x40 if me else 0)
    h[8:16] = len(payload).to_bytes(8, "big")
    h[16:16 + len(rtype)] = rtype.encode()
    return bytes(h) + payload + b"\0" * ((8 - len(payload) % 8) % 8)


def _lime_records(path):
    data, out, pos = open(path, "rb").read(), [], 0
    while pos < len(data):
        assert int.from_bytes(data[pos:pos + 4], "big") == 0x456789AB
        n = int.from_bytes(data[pos + 8:pos + 16], "big")
        rtype = data[pos + 16:pos + 144].split(b"\0")[0].decode()
        out.append((rtype, data[pos + 144:pos + 144 + n]))
        pos += 144 + n + (8 - n % 8) % 8
    return out


@pytest.mark.parametrize("bpr", [4, 8])
def test_ildg_round_trip_and_records(api, w, tmp_path, bpr):
    L = [4, 6, 4, 4]
    A = _su3(w, L, 21)
    path = str(tmp_path / "conf.lime")
    assert write(api, ILDG, path, L, A, 3, bpr) == OK
    recs = _lime_records(path)
    assert [r[0] for r in recs] == ["ildg-format", "ildg-binary-data"]
    assert f"<precision>{8 * bpr}</precision>".encode() in recs[0][1]
    assert len(recs[1][1]) == int(np.prod(L)) * 4 * 9 * 2 * bpr
    # payload: big-endian, sites x fastest, mu, row-major 3x3 (check one link)
    x, eo = _geometry(L)
    site, mu = 5, 2
    dt = ">f8" if bpr == 8 else ">f4"
    raw = np.frombuffer(recs[1][1], dtype=dt).reshape(-1, 4, 9, 2)
    link = raw[site, mu, :, 0] + 1j * raw[site, mu, :, 1]
    assert np.max(np.abs(link - A[mu].reshape(-1, 9)[eo[site]])) < (1e-15 if bpr == 8 else 1e-6)
    rc, B, vals, meta = read(api, ILDG, path, L)
    assert rc == OK and meta[:3] == [3, bpr, 1]
    tol = 0 if bpr == 8 else 2e-6
    for a, b in zip(A, B):
        assert np.max(np.abs(a - b)) <= tol
    # extra records before and after the binary data (as in files of other writers)
    extra = str(tmp_path / "extra.lime")
    open(extra, "wb").write(_lime_record("scidac-private-file-xml", b"<x/>", True, False)
                            + b"".join(_lime_record(t, p, False, False) for t, p in recs)
                            + _lime_record("ildg-data-lfn", b"lfn://some/where", False, True))
    rc2, C, _, _ = read(api, ILDG, extra, L)
    assert rc2 == OK and all(np.array_equal(b, c) for b, c in zip(B, C))


def test_ildg_errors(api, w, tmp_path):
    L = [4, 4, 4, 4]
    path = str(tmp_path / "c.lime")
    assert write(api, ILDG, path, L, _su3(w, L, 22)) == OK
    assert read(api, ILDG, path, [4, 4, 4, 6])[0] == EDIMS
    data = open(path, "rb").read()
    bad = str(tmp_path / "bad.lime")
    open(bad, "wb").write(b"\0" * 4 + data[4:])
    assert read(api, ILDG, bad, L)[0] == EFORMAT
    open(bad, "wb").write(data[:-4096])
    assert read(api, ILDG, bad, L)[0] in (ESHORT, EFORMAT)
