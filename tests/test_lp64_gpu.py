"""The LP64 twin of the library (libvecjac_hip_lp64.so: the reference's
integers are 32-bit, as in the reference built without MKL_ILP64,
common.h:186-190) against the ILP64 library on the same inputs: dvjsvd_,
zvjsvd_, dbjac2_, zbjac2_ and dnssvj_ called with 32-bit integers return
the same codes and the same bits, and the argument checks report the same
codes."""
import ctypes as C

import numpy as np
import pytest

from vecjac_amd import _lib
from vecjac_amd import svd as S

pytestmark = pytest.mark.gpu

VP = C.c_void_p


def _libs():
    _lib.lib()  # the HIP runtime torch uses is loaded first
    ilp = C.CDLL(str(_lib.LIB_PATH))
    lp = C.CDLL(str(_lib.LIB_PATH.with_name("libvecjac_hip_lp64.so")))
    assert ilp.vjh_fint_bits() == 64 and lp.vjh_fint_bits() == 32
    for L, it in ((ilp, C.c_int64), (lp, C.c_int32)):
        L.dvjsvd_.restype = it
        L.dvjsvd_.argtypes = [VP] * 13
        L.zvjsvd_.restype = it
        L.zvjsvd_.argtypes = [VP] * 17
        L.dbjac2_.restype = it
        L.dbjac2_.argtypes = [VP] * 9
        L.zbjac2_.restype = it
        L.zbjac2_.argtypes = [VP] * 11
        L.dnssvj_.restype = None
        L.dnssvj_.argtypes = [VP] * 15
    return ilp, lp


def _dvjsvd(L, it, nt, A, js, swp=999):
    n, m = A.shape
    G = S.aligned((n, m), np.float64)
    G[...] = A
    V = S.aligned((n, n), np.float64)
    eS = S.aligned(n, np.float64)
    fS = S.aligned(n, np.float64)
    work = S.aligned(8 * n, np.float64)
    iw = S.aligned(max(1, n // 8), np.uint32)
    jsa = S.aligned(js.shape, np.uint32)
    jsa[...] = js
    u = lambda v: C.byref(nt(v))  # noqa: E731
    rc = L.dvjsvd_(u(m), u(n), G.ctypes.data, u(m), V.ctypes.data, u(n), eS.ctypes.data, fS.ctypes.data,
                   jsa.ctypes.data, C.byref(C.c_uint(js.shape[0])), C.byref(C.c_uint(swp)), work.ctypes.data,
                   iw.ctypes.data)
    return rc, G, V, eS, fS


def test_lp64_dvjsvd_equals_ilp64():
    ilp, lp = _libs()
    g = np.random.default_rng(64)
    n, m = 64, 200
    A = g.standard_normal((n, m))
    js, _ = S.pjs(2, n)
    print("dvjsvd_ ILP64 and LP64", flush=True)
    r64 = _dvjsvd(ilp, C.c_int64, C.c_uint64, A, js)
    r32 = _dvjsvd(lp, C.c_int32, C.c_uint32, A, js)
    assert r64[0] > 0 and r32[0] == r64[0]
    for a, b in zip(r64[1:], r32[1:]):
        assert np.array_equal(a, b)
    # argument checks with 32-bit integers: odd n (-2 as in dvjsvd.c:45) and ldG < m (-4)
    A_odd = g.standard_normal((7, 20))
    assert _dvjsvd(lp, C.c_int32, C.c_uint32, A_odd, np.zeros((6, 7), np.uint32))[0] == \
        _dvjsvd(ilp, C.c_int64, C.c_uint64, A_odd, np.zeros((6, 7), np.uint32))[0] < 0


def test_lp64_zvjsvd_equals_ilp64():
    ilp, lp = _libs()
    g = np.random.default_rng(65)
    n, m = 32, 96
    Ar, Ai = g.standard_normal((n, m)), g.standard_normal((n, m))
    js, _ = S.pjs(2, n)
    js_a = S.aligned(js.shape, np.uint32)
    js_a[...] = js
    outs = []
    for L, nt in ((ilp, C.c_uint64), (lp, C.c_uint32)):
        Gr, Gi = S.aligned((n, m), np.float64), S.aligned((n, m), np.float64)
        Gr[...], Gi[...] = Ar, Ai
        Vr, Vi = S.aligned((n, n), np.float64), S.aligned((n, n), np.float64)
        eS, fS = S.aligned(n, np.float64), S.aligned(n, np.float64)
        work = S.aligned(16 * n, np.float64)
        iw = S.aligned(max(1, n // 8), np.uint32)
        u = lambda v: C.byref(nt(v))  # noqa: E731
        print(f"zvjsvd_ {nt.__name__}", flush=True)
        rc = L.zvjsvd_(u(m), u(n), Gr.ctypes.data, u(m), Gi.ctypes.data, u(m), Vr.ctypes.data, u(n), Vi.ctypes.data,
                       u(n), eS.ctypes.data, fS.ctypes.data, js_a.ctypes.data, C.byref(C.c_uint(js.shape[0])),
                       C.byref(C.c_uint(999)), work.ctypes.data, iw.ctypes.data)
        outs.append((rc, Gr, Gi, Vr, Vi, eS, fS))
    assert outs[0][0] > 0 and outs[1][0] == outs[0][0]
    for a, b in zip(outs[0][1:], outs[1][1:]):
        assert np.array_equal(a, b)


def test_lp64_bjac2_equals_ilp64():
    from vecjac_amd import workloads as W
    ilp, lp = _libs()
    n = 1 << 12
    for kind in ("d", "z"):
        ins = [a for a in W.jac2_batch(n, 77, kind == "z") if a is not None]
        res = []
        for L, it in ((ilp, C.c_int64), (lp, C.c_int32)):
            arrs = [S.aligned(n, np.float64) for _ in ins]
            for a, b in zip(arrs, ins):
                a[...] = b
            outs = [S.aligned(n, np.float64) for _ in range(5 if kind == "z" else 4)]
            p = S.aligned(n // 8, np.uint32)  # one word per 8 problems
            print(f"{kind}bjac2_ {it.__name__}", flush=True)
            f = L.zbjac2_ if kind == "z" else L.dbjac2_
            rc = f(C.byref(it(n)), *[a.ctypes.data for a in arrs], *[o.ctypes.data for o in outs], p.ctypes.data)
            res.append((rc, outs, p))
        assert res[0][0] >= 0 and res[1][0] == res[0][0]
        for a, b in zip(res[0][1], res[1][1]):
            assert np.array_equal(a, b, equal_nan=True)
        assert np.array_equal(res[0][2], res[1][2])


def test_lp64_dnssvj_equals_ilp64():
    ilp, lp = _libs()
    g = np.random.default_rng(66)
    m, n = 120, 40
    A0 = np.asfortranarray(g.standard_normal((m, n)))
    res = []
    for L, it in ((ilp, C.c_int64), (lp, C.c_int32)):
        A = A0.copy(order="F")
        sva = np.zeros(n)
        V = np.zeros((n, n), order="F")
        work = np.zeros(max(m + n, 6))
        info = it(0)
        i = lambda v: C.byref(it(v))  # noqa: E731
        print(f"dnssvj_ {it.__name__}", flush=True)
        L.dnssvj_(b"G", b"U", b"V", i(m), i(n), A.ctypes.data, i(m), sva.ctypes.data, i(n), V.ctypes.data, i(n),
                  work.ctypes.data, i(work.size), i(30), C.byref(info))
        res.append((info.value, A, sva, V, work[:6].copy()))
    assert res[0][0] == 0 and res[1][0] == 0
    for a, b in zip(res[0][1:], res[1][1:]):
        assert np.array_equal(a, b)
