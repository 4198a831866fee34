"""Oracle (test infrastructure only): ctypes binding of liboracle_vj.so, the
CPU restatement of the reference, and of oracle/_ref/libvjref.so, the
reference's own hot-path sources compiled in place (when they were present
at build time).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only to check results.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle_vj.so"
REF_SO = HERE / "_ref" / "libvjref.so"

_fint = C.c_int64
_fnat = C.c_uint64
_vp = C.c_void_p


class Cplx(C.Structure):
    """double complex return value (two SSE-class doubles: xmm0, xmm1)."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("sweeps", C.c_uint), ("sT", C.c_int), ("sV", C.c_int), ("rescales", C.c_uint),
                ("swt", C.c_uint64 * 64)]


_oracle = None
_ref = None


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(str(ORACLE_SO))
        P = C.POINTER
        for pre, t in (("d", C.c_double), ("s", C.c_float)):
            for nm in (f"vjo_{pre}bjac2", f"vjo_{pre}bjac2i"):
                f = getattr(L, nm)
                f.restype, f.argtypes = _fint, [P(_fint)] + [_vp] * 8
        for pre in ("z", "c"):
            for nm in (f"vjo_{pre}bjac2", f"vjo_{pre}bjac2i"):
                f = getattr(L, nm)
                f.restype, f.argtypes = _fint, [P(_fint)] + [_vp] * 10
        for pre, t in (("d", C.c_double), ("z", C.c_double), ("s", C.c_float), ("c", C.c_float)):
            f = getattr(L, f"vjo_{pre}vjsvd_core")
            f.restype = _fint
            f.argtypes = [C.c_int, P(_fnat), P(_fnat), _vp, P(_fnat), _vp, P(_fnat), _vp, P(_fnat), _vp, P(_fnat),
                          _vp, _vp, _vp, P(C.c_uint), P(C.c_uint), _vp, _vp, C.POINTER(Stats)]
        L.vjo_ddpscl.restype, L.vjo_ddpscl.argtypes = C.c_double, [_fnat, _vp, _vp, _vp, _vp]
        L.vjo_zdpscl.restype, L.vjo_zdpscl.argtypes = None, [_fnat, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.vjo_dnorm2.restype, L.vjo_dnorm2.argtypes = C.c_double, [_fnat, _vp, _vp, _vp]
        L.vjo_znorm2.restype, L.vjo_znorm2.argtypes = C.c_double, [_fnat, _vp, _vp, _vp, _vp]
        L.vjo_djrot.restype, L.vjo_djrot.argtypes = C.c_double, [_fint, _vp, _vp, C.c_double, C.c_double]
        L.vjo_djrotf.restype, L.vjo_djrotf.argtypes = C.c_double, [_fint, _vp, _vp, C.c_double, C.c_double]
        L.vjo_zjrot.restype = C.c_double
        L.vjo_zjrot.argtypes = [_fint, _vp, _vp, _vp, _vp, C.c_double, C.c_double, C.c_double]
        L.vjo_dgsscl.restype, L.vjo_dgsscl.argtypes = C.c_double, [_fint, C.c_double, _vp, _vp, _vp, _vp]
        L.vjo_zgsscl.restype = C.c_double
        L.vjo_zgsscl.argtypes = [_fint, C.c_double, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp]
        L.vjo_sdpscl.restype, L.vjo_sdpscl.argtypes = C.c_float, [_fnat, _vp, _vp, _vp, _vp]
        L.vjo_set_norm_mode.restype, L.vjo_set_norm_mode.argtypes = None, [C.c_int]
        L.vjo_get_norm_mode.restype, L.vjo_get_norm_mode.argtypes = C.c_int, []
        L.vjo_set_secants.restype, L.vjo_set_secants.argtypes = None, [C.c_int]
        L.vjo_get_secants.restype, L.vjo_get_secants.argtypes = C.c_int, []
        for nm, t in (("vjo_dvjsvd_batch", C.c_double), ("vjo_svjsvd_batch", C.c_float)):
            f = getattr(L, nm)
            f.restype = C.c_int64
            f.argtypes = [C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_uint, _vp]
        _oracle = L
    return _oracle


_ref_sec = None


def ref() -> C.CDLL | None:
    """The reference's own kernels (built from its sources), or None."""
    global _ref
    if _ref is None and REF_SO.exists():
        _ref = _declare_ref(C.CDLL(str(REF_SO)))
    return _ref


def ref_sec() -> C.CDLL | None:
    """The reference built with USE_SECANTS (its rotations and dvjsvd_), or
    None."""
    global _ref_sec
    so = HERE / "_ref" / "libvjref_sec.so"
    if _ref_sec is None and so.exists():
        _ref_sec = _declare_ref(C.CDLL(str(so)))
    return _ref_sec


def _declare_ref(L: C.CDLL) -> C.CDLL:
    """Prototypes of the reference build's entry points."""
    P = C.POINTER
    for nm in ("dbjac2_", "dbjac2i", "sbjac2_", "sbjac2i"):
        f = getattr(L, nm)
        f.restype, f.argtypes = _fint, [P(_fint)] + [_vp] * 8
    for nm in ("zbjac2_", "zbjac2i", "cbjac2_", "cbjac2i"):
        f = getattr(L, nm)
        f.restype, f.argtypes = _fint, [P(_fint)] + [_vp] * 10
    L.ddpscl_.restype, L.ddpscl_.argtypes = C.c_double, [P(_fnat), _vp, _vp, _vp, _vp]
    L.sdpscl_.restype, L.sdpscl_.argtypes = C.c_float, [P(_fnat), _vp, _vp, _vp, _vp]
    L.zdpscl_.restype, L.zdpscl_.argtypes = Cplx, [P(_fnat), _vp, _vp, _vp, _vp, _vp, _vp]
    L.dnormf_.restype, L.dnormf_.argtypes = C.c_double, [P(_fnat), _vp, _vp, _vp, _vp, _vp]
    L.djrot_.restype, L.djrot_.argtypes = C.c_double, [P(_fint), _vp, _vp, _vp, _vp]
    L.djrotf_.restype, L.djrotf_.argtypes = C.c_double, [P(_fint), _vp, _vp, _vp, _vp]
    L.zjrot_.restype, L.zjrot_.argtypes = C.c_double, [P(_fint), _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.dgsscl_.restype, L.dgsscl_.argtypes = C.c_double, [P(_fint), _vp, _vp, _vp, _vp, _vp]
    L.zgsscl_.restype, L.zgsscl_.argtypes = C.c_double, [P(_fint), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.dvjsvd_.restype = _fint
    L.dvjsvd_.argtypes = [P(_fnat), P(_fnat), _vp, P(_fnat), _vp, P(_fnat), _vp, _vp, _vp, P(C.c_uint),
                          P(C.c_uint), _vp, _vp]
    return L


def aligned(shape, dtype, fill=0) -> np.ndarray:
    """64-byte aligned array (the reference's alignment requirement)."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape))
    raw = np.empty(n * dtype.itemsize + 64, dtype=np.uint8)
    off = (-raw.ctypes.data) % 64
    a = raw[off: off + n * dtype.itemsize].view(dtype).reshape(shape)
    a[...] = fill
    return a


def bjac2(kind: str, n_signed: int, ins, internal: bool = False, use_ref: bool = False, p_init=None):
    """Run the 2x2 EVD of `kind` in {'d','z','s','c'} on the oracle (default)
    or on the reference build.  Returns (rc, [c, at(/cat, sat), l1, l2], p)."""
    dtype = np.float64 if kind in "dz" else np.float32
    if use_ref:
        L = ref()
        if L is None:
            raise RuntimeError("reference build not available")
        name = f"{kind}bjac2{'i' if internal else '_'}"
        f = getattr(L, name)
    else:
        L = oracle()
        f = getattr(L, f"vjo_{kind}bjac2{'i' if internal else ''}")
    nn = abs(n_signed)
    vl = 8 if dtype == np.float64 else 16
    ins_al = []
    for a in ins:
        b = aligned(nn, dtype)
        b[...] = np.asarray(a, dtype=dtype)
        ins_al.append(b)
    outs = [aligned(nn, dtype) for _ in range(5 if kind in "zc" else 4)]
    p = aligned(max(1, nn // vl), np.uint32)
    if p_init is not None:
        p[...] = p_init
    rc = f(C.byref(_fint(n_signed)), *[_ptr(a) for a in ins_al], *[_ptr(a) for a in outs], _ptr(p))
    return rc, outs, p


def vjsvd_core(kind: str, G, V=None, js=None, stp=0, swp=999, strict=False, Gi=None, Vi=None, iwork0=0,
               ldG=None, ldV=None):
    """Relaxed-core SVD on the oracle.  G (and Gi for complex) are (n, ld)
    arrays holding columns as rows (column-major storage); they are updated
    in place (G -> U).  Returns (rc, eS, fS, V(, Vi), stats)."""
    L = oracle()
    dtype = np.float64 if kind in "dz" else np.float32
    n, ld = G.shape
    m = ld if ldG is None else ldG
    if V is None:
        V = aligned((n, n), dtype)
    if kind in "zc" and Vi is None:
        Vi = aligned((n, n), dtype)
    eS = aligned(n, dtype)
    fS = aligned(n, dtype)
    iw = aligned(max(1, n // 8), np.uint32)
    iw[0] = iwork0
    work = aligned(8 * n + 64, dtype)
    st = Stats()
    f = getattr(L, f"vjo_{kind}vjsvd_core")
    mm, nn = _fnat(m), _fnat(n)
    lG, lV = _fnat(ld), _fnat(V.shape[1])
    rc = f(int(strict), C.byref(mm), C.byref(nn), _ptr(G), C.byref(lG),
           _ptr(Gi) if Gi is not None else None, C.byref(lG),
           _ptr(V), C.byref(lV), _ptr(Vi) if Vi is not None else None, C.byref(lV),
           _ptr(eS), _ptr(fS), _ptr(js), C.byref(C.c_uint(stp)), C.byref(C.c_uint(swp)),
           _ptr(work), _ptr(iw), C.byref(st))
    return rc, eS, fS, V, Vi, st


class norm_mode:
    """Context manager selecting the oracle's column-norm definition:
    0 = the reference's long-double dnormf_, 1 = the GPU's correctly rounded
    scaled norm."""

    def __init__(self, mode: int):
        self.mode = mode

    def __enter__(self):
        L = oracle()
        self.prev = L.vjo_get_norm_mode()
        L.vjo_set_norm_mode(self.mode)
        return self

    def __exit__(self, *exc):
        oracle().vjo_set_norm_mode(self.prev)
        return False


class secants:
    """Context manager selecting the oracle's rotation form: on = the
    reference's USE_SECANTS build (the 2x2 EVD returns secants and the
    rotations divide by them), off = cosines (the default build)."""

    def __init__(self, on: bool = True):
        self.on = int(bool(on))

    def __enter__(self):
        L = oracle()
        self.prev = L.vjo_get_secants()
        L.vjo_set_secants(self.on)
        return self

    def __exit__(self, *exc):
        oracle().vjo_set_secants(self.prev)
        return False


def sigma(eS: np.ndarray, fS: np.ndarray) -> np.ndarray:
    """sigma_j = fS_j * 2^eS_j (exact; 0 for eS = -inf)."""
    e = np.asarray(eS, dtype=np.float64)
    f = np.asarray(fS, dtype=np.float64)
    out = np.zeros_like(f)
    ok = np.isfinite(e)
    out[ok] = np.ldexp(f[ok], e[ok].astype(np.int64))
    return out
