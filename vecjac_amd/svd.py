"""Host-side interface of the GPU Jacobi SVD, mirroring the reference's calls.

`dvjsvd(m, n, G, ldG, V, ldV, eS, fS, js, stp, swp, work, iwork)` and its
z/s/c twins take numpy arrays laid out exactly as the reference's arguments
(column-major, G[j * ldG + i]); they call the drop-in C entry points of
libvecjac_hip.so.  `device_vjsvd` runs the same driver on torch CUDA
tensors.  `pjs(id, n)` returns the strategy table as a uint32 array.
`nssvj(kind, joba, jobu, jobv, m, n, A, lda, sva, mv, V, ldv, ...)` calls the
LAPACK xGESVJ-style drivers ?nssvj_ with LAPACK's argument list.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_DT = {"d": np.float64, "z": np.float64, "s": np.float32, "c": np.float32}


def pjs(id_: int, n: int):
    """Strategy table (pjs.c:7): (table of shape (stp, n), stp)."""
    L = _lib.lib()
    stp = C.c_uint(0)
    p = L.pjs(id_, n, C.byref(stp))
    if not p:
        raise ValueError(f"pjs({id_}, {n}) failed")
    s = stp.value
    t = np.ctypeslib.as_array(p, shape=(s * n,)).copy().reshape(s, n)
    L.pjs_free(p)
    return t.astype(np.uint32), s


def aligned(shape, dtype, fill=0) -> np.ndarray:
    """64-byte aligned numpy array (the reference's alignment rule)."""
    dtype = np.dtype(dtype)
    cnt = int(np.prod(shape))
    raw = np.empty(cnt * dtype.itemsize + 64, dtype=np.uint8)
    off = (-raw.ctypes.data) % 64
    a = raw[off: off + cnt * dtype.itemsize].view(dtype).reshape(shape)
    a[...] = fill
    return a


def _u64(x):
    return C.byref(C.c_uint64(int(x)))


def _u32(x):
    return C.byref(C.c_uint(int(x)))


def vjsvd(kind: str, m, n, Gr, ldGr, Vr, ldVr, eS, fS, js, stp, swp, work, iwork, Gi=None, ldGi=None, Vi=None,
          ldVi=None) -> int:
    """Call ?vjsvd_ with the reference's argument list; arrays are numpy."""
    L = _lib.lib()
    f = getattr(L, f"{kind}vjsvd_")
    if kind in "dz" and Gr.dtype != np.float64 or kind in "sc" and Gr.dtype != np.float32:
        raise TypeError("dtype does not match the routine")
    if kind in "ds":
        rc = f(_u64(m), _u64(n), Gr.ctypes.data, _u64(ldGr), Vr.ctypes.data, _u64(ldVr), eS.ctypes.data,
               fS.ctypes.data, js.ctypes.data, _u32(stp), _u32(swp), work.ctypes.data, iwork.ctypes.data)
    else:
        rc = f(_u64(m), _u64(n), Gr.ctypes.data, _u64(ldGr), Gi.ctypes.data, _u64(ldGi), Vr.ctypes.data,
               _u64(ldVr), Vi.ctypes.data, _u64(ldVi), eS.ctypes.data, fS.ctypes.data, js.ctypes.data, _u32(stp),
               _u32(swp), work.ctypes.data, iwork.ctypes.data)
    return _lib.check(rc, f"{kind}vjsvd_")


def sigma(eS, fS) -> np.ndarray:
    """sigma_j = fS_j * 2^eS_j as float64 (0 where eS = -inf)."""
    e = np.asarray(eS, dtype=np.float64)
    f = np.asarray(fS, dtype=np.float64)
    out = np.zeros_like(f)
    ok = np.isfinite(e)
    out[ok] = np.ldexp(f[ok], e[ok].astype(np.int64))
    return out


def dvjsvd_host(A_cols: np.ndarray, m: int, n: int, strategy: int = 2, swp: int = 999):
    """Plain SVD of the m x n matrix whose columns are the rows of A_cols
    (shape (n, m)) through dvjsvd_.  Returns (sigma, U_cols, V_cols, rc)."""
    js, stp = pjs(strategy, n)
    G = aligned((n, m), np.float64)
    G[...] = A_cols
    V = aligned((n, n), np.float64)
    eS = aligned(n, np.float64)
    fS = aligned(n, np.float64)
    work = aligned(max(8, 5 * n), np.float64)
    iwork = aligned(max(1, n // 8), np.uint32)
    rc = vjsvd("d", m, n, G, m, V, n, eS, fS, js, stp, swp, work, iwork)
    return sigma(eS, fS), G, V, rc


def device_vjsvd(kind: str, G, V, eS, fS, js_dev, stp: int, swp: int = 999, Gi=None, Vi=None, vpreset: bool = False,
                 stream=None):
    """Run the plain driver on torch CUDA tensors.  G (and Gi) have shape
    (n, ldG) (columns as rows), V (and Vi) shape (n, ldV); js_dev is a
    uint32/int32 CUDA tensor with stp*n entries.  Returns (rc, stats) where
    stats = (sweeps, launched sweeps, per-sweep transformation counts)."""
    import torch
    L = _lib.lib()
    n, ldG = G.shape
    ldV = V.shape[1]
    m = ldG
    stats = (C.c_uint * 66)()
    s = (stream or torch.cuda.current_stream()).cuda_stream
    f = getattr(L, f"vjh_{kind}vjsvd")
    if kind in "ds":
        rc = f(m, n, G.data_ptr(), ldG, V.data_ptr(), ldV, eS.data_ptr(), fS.data_ptr(), js_dev.data_ptr(), stp, swp,
               int(vpreset), s, stats)
    else:
        rc = f(m, n, G.data_ptr(), Gi.data_ptr(), ldG, V.data_ptr(), Vi.data_ptr(), ldV, eS.data_ptr(),
               fS.data_ptr(), js_dev.data_ptr(), stp, swp, int(vpreset), s, stats)
    _lib.check(rc, f"vjh_{kind}vjsvd")
    return rc, (stats[0], stats[1], list(stats[2:2 + 64]))


class secants:
    """Context manager for the plain drivers' rotation form
    (vjh_set_secants): on = the reference's USE_SECANTS build (the 2x2 EVD
    returns secants and the rotations divide by them)."""

    def __init__(self, on: bool = True):
        self.on = int(bool(on))

    def __enter__(self):
        L = _lib.lib()
        self.prev = L.vjh_get_secants()
        L.vjh_set_secants(self.on)
        return self

    def __exit__(self, *exc):
        _lib.lib().vjh_set_secants(self.prev)
        return False


def _i64(x):
    return C.byref(C.c_int64(int(x)))


def nssvj(kind: str, joba: str, jobu: str, jobv: str, m, n, A, lda, sva, mv, V, ldv, work, lwork, nsweep,
          rwork=None, lrwork=None) -> int:
    """Call ?nssvj_ (dnssvj.f:335, znssvj.f:349) with LAPACK's argument list
    and return INFO.  Arrays are numpy: A and V column-major with leading
    dimensions lda / ldv (complex128 / complex64 for z / c), work the real
    WORK (d, s) or the complex CWORK (z, c) with rwork the real RWORK."""
    L = _lib.lib()
    f = getattr(L, f"{kind}nssvj_")
    want = {"d": np.float64, "s": np.float32, "z": np.complex128, "c": np.complex64}[kind]
    if A.dtype != want:
        raise TypeError("dtype does not match the routine")
    info = C.c_int64(0)
    ptr = lambda x: None if x is None else x.ctypes.data  # noqa: E731
    args = [joba.encode(), jobu.encode(), jobv.encode(), _i64(m), _i64(n), ptr(A), _i64(lda), ptr(sva), _i64(mv),
            ptr(V), _i64(ldv), ptr(work), _i64(lwork)]
    if kind in "zc":
        args += [ptr(rwork), _i64(lrwork)]
    args += [_i64(nsweep), C.byref(info)]
    f(*args)
    return _lib.check(info.value, f"{kind}nssvj_")


def gesvj(A: np.ndarray, joba: str = "G", jobu: str = "U", jobv: str = "V", nsweep: int = 30, ctol=None,
          V0: np.ndarray = None):
    """xGESVJ-style SVD of the m x n matrix A (m >= n) through ?nssvj_.
    Returns (sva, U, V, stat, info) with sigma = stat[0] * sva; U is A on
    exit (m x n), V the n x n right singular vectors (JOBV 'V'), V0 times the
    rotations (JOBV 'A', V0 of shape (mv, n)), or None (JOBV 'N')."""
    A = np.asarray(A)
    kind = {np.dtype(np.float64): "d", np.dtype(np.float32): "s", np.dtype(np.complex128): "z",
            np.dtype(np.complex64): "c"}[A.dtype]
    real_t = np.float64 if kind in "dz" else np.float32
    m, n = A.shape
    Af = np.asfortranarray(A.copy())
    sva = np.zeros(max(n, 1), real_t)
    if jobv.upper() == "A":
        mv = V0.shape[0]
        Vf = np.asfortranarray(np.asarray(V0, dtype=A.dtype).copy())
        ldv = max(mv, 1)
    else:
        mv = 0
        Vf = np.asfortranarray(np.zeros((max(n, 1), max(n, 1)), A.dtype))
        ldv = max(n, 1)
    if kind in "ds":
        lwork = max(m + n, 6)
        work = np.zeros(lwork, real_t)
        if ctol is not None:
            work[0] = ctol
        info = nssvj(kind, joba, jobu, jobv, m, n, Af, max(m, 1), sva, mv, Vf, ldv, work, lwork, nsweep)
        stat = work[:6].copy()
    else:
        lwork = max(m + n, 1)
        cwork = np.zeros(lwork, A.dtype)
        lrwork = max(n, 6)
        rwork = np.zeros(lrwork, real_t)
        if ctol is not None:
            rwork[0] = ctol
        info = nssvj(kind, joba, jobu, jobv, m, n, Af, max(m, 1), sva, mv, Vf, ldv, cwork, lwork, nsweep,
                     rwork, lrwork)
        stat = rwork[:6].copy()
    V = None if jobv.upper() == "N" else Vf[:mv, :n] if jobv.upper() == "A" else Vf[:n, :n]
    return sva[:n], Af, V, stat, info
