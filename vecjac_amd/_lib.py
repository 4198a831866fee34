"""ctypes binding of libvecjac_hip.so (the C ABI declared in include/vecjac.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object has not been built.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import os

# VJ_LIB selects another build of the same library (for comparing builds)
LIB_PATH = Path(os.environ.get("VJ_LIB") or Path(__file__).resolve().with_name("libvecjac_hip.so"))

_fint = C.c_int64
_fnat = C.c_uint64
_pd = C.c_void_p  # every array argument is passed as a raw address
_pu = C.c_void_p

_lib = None


def _decl(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def lib() -> C.CDLL:
    """Load (once) and return the HIP library with its prototypes declared."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m vecjac_amd.build`")
    # One HIP runtime per process: when torch is importable its bundled
    # libamdhip64.so.7 is loaded first, and the library's own dependency on
    # that soname then binds to the same copy (device memory and streams are
    # shared with torch instead of being split over two runtimes).
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    L = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    # reference-compatible 2x2 EVD (host arrays)
    for name in ("dbjac2_", "sbjac2_", "dbjac2i", "sbjac2i"):
        _decl(L, name, _fint, [P(_fint), _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pu])
    for name in ("zbjac2_", "cbjac2_", "zbjac2i", "cbjac2i"):
        _decl(L, name, _fint, [P(_fint), _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pu])
    # device-memory 2x2 EVD
    for name in ("vjh_dbjac2", "vjh_sbjac2", "vjh_dbjac2i", "vjh_sbjac2i"):
        _decl(L, name, _fint, [_fint, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pu, C.c_void_p])
    for name in ("vjh_zbjac2", "vjh_cbjac2", "vjh_zbjac2i", "vjh_cbjac2i"):
        _decl(L, name, _fint, [_fint, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pu, C.c_void_p])
    _decl(L, "pjs", P(C.c_uint), [C.c_long, C.c_uint, P(C.c_uint)])
    _decl(L, "pjs_free", None, [P(C.c_uint)])
    _decl(L, "vjh_last_error", C.c_char_p, [])
    _decl(L, "vjh_abi_version", C.c_int, [])
    _decl(L, "vjh_set_secants", None, [C.c_int])
    _decl(L, "vjh_get_secants", C.c_int, [])
    _declare_svd(L)
    _lib = L
    return L


def _declare_svd(L) -> None:
    """Prototypes of the SVD entry points (present once built)."""
    P = C.POINTER
    sv = [P(_fnat), P(_fnat), _pd, P(_fnat), _pd, P(_fnat), _pd, _pd, _pu, P(C.c_uint), P(C.c_uint), _pd, _pu]
    cv = [P(_fnat), P(_fnat), _pd, P(_fnat), _pd, P(_fnat), _pd, P(_fnat), _pd, P(_fnat), _pd, _pd, _pu,
          P(C.c_uint), P(C.c_uint), _pd, _pu]
    for name, args in (("dvjsvd_", sv), ("svjsvd_", sv), ("zvjsvd_", cv), ("cvjsvd_", cv)):
        _decl(L, name, _fint, args)
    U = C.c_uint
    dsv = [_fnat, _fnat, _pd, _fnat, _pd, _fnat, _pd, _pd, _pu, U, U, U, C.c_void_p, _pu]
    dcv = [_fnat, _fnat, _pd, _pd, _fnat, _pd, _pd, _fnat, _pd, _pd, _pu, U, U, U, C.c_void_p, _pu]
    for name, args in (("vjh_dvjsvd", dsv), ("vjh_svjsvd", dsv), ("vjh_zvjsvd", dcv), ("vjh_cvjsvd", dcv)):
        _decl(L, name, _fint, args)
    # LAPACK xGESVJ-style drivers (subroutines: INFO is the last argument)
    F = P(_fint)
    ch = [C.c_char_p] * 3
    rn = ch + [F, F, _pd, F, _pd, F, _pd, F, _pd, F, F, F]
    cn = ch + [F, F, _pd, F, _pd, F, _pd, F, _pd, F, _pd, F, F, F]
    for name, args in (("dnssvj_", rn), ("snssvj_", rn), ("znssvj_", cn), ("cnssvj_", cn)):
        _decl(L, name, None, args)


def last_error() -> str:
    return lib().vjh_last_error().decode(errors="replace")


def check(rc: int, what: str) -> int:
    """Raise on a HIP-runtime failure (<= -100000); other codes are returned."""
    if rc <= -100000:
        raise RuntimeError(f"{what}: {last_error()} (status {rc})")
    return rc
