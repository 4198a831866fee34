"""Pins the oracle (CPU restatement) to the reference.

1. Against the committed golden vectors (tests/golden/*.npz, written by
   make_golden.py from the reference's own sources compiled in place): the
   2x2 EVD in four precisions and two forms, and the full dvjsvd_ driver on
   Gaussian, rank-deficient, sweep-limited and graded (etc/dgsex.c) inputs.
   Bar: bit-for-bit equality.
2. Against the reference build itself, when oracle/_ref/libvjref.so is
   present: the level-1 kernels (ddpscl_/zdpscl_/sdpscl_, djrot_/djrotf_/
   zjrot_, dgsscl_/zgsscl_), the internal 2x2 forms and dvjsvd_ on fresh
   random inputs; and the rotations and dvjsvd_ of the reference built with
   USE_SECANTS (oracle/_ref/libvjref_sec.so) against the oracle's secant
   form.  Bar: bit-for-bit equality.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from oracle import vjo
from tests.util import bit_mismatch
from vecjac_amd import workloads as W

GOLD = Path(__file__).resolve().parent / "golden"


def _pjs_table(n):
    """ME table without the GPU library (same construction as csrc/pjs.cpp)."""
    from vecjac_amd import svd as S
    return S.pjs(2, n)


# ------------------------------------------------------------- golden vectors
@pytest.mark.parametrize("kind", ["d", "z", "s", "c"])
@pytest.mark.parametrize("form", ["tan", "sin"])
def test_bjac2_golden(kind, form):
    g = np.load(GOLD / "bjac2.npz")
    ins = [g[f"{kind}_in{k}"] for k in range(4 if kind in "zc" else 3)]
    nn = ins[0].size
    rc, outs, p = vjo.bjac2(kind, nn if form == "tan" else -nn, ins)
    assert rc == 0
    for k, a in enumerate(outs):
        assert bit_mismatch(a, g[f"{kind}_{form}_out{k}"]) == 0, f"output {k}"
    assert np.array_equal(p, g[f"{kind}_{form}_p"])


@pytest.mark.parametrize("name", ["g64x32", "g128x48", "dgsex16", "rankdef48x32", "g64x32_swp2"])
def test_dvjsvd_golden(name):
    g = np.load(GOLD / "dvjsvd.npz")
    A = g[f"{name}_A"]
    n, m = A.shape
    js = g[f"{name}_js"]
    stp = js.shape[0]
    G = vjo.aligned((n, m), np.float64)
    G[...] = A
    with vjo.norm_mode(0):
        rc, eS, fS, V, _, _ = vjo.vjsvd_core("d", G, js=js, stp=stp, swp=int(g[f"{name}_swp"]), strict=True)
    assert rc == int(g[f"{name}_rc"])
    assert bit_mismatch(G, g[f"{name}_U"]) == 0
    assert bit_mismatch(V, g[f"{name}_V"]) == 0
    assert bit_mismatch(eS, g[f"{name}_eS"]) == 0
    assert bit_mismatch(fS, g[f"{name}_fS"]) == 0


# -------------------------------------------------------- the reference build
ref_only = pytest.mark.skipif(vjo.ref() is None, reason="reference build (oracle/_ref) not present")


def _ef(x):
    nrm = np.sqrt(np.sum(np.asarray(x, dtype=np.longdouble) ** 2))
    e = float(np.floor(np.log2(float(nrm))))
    return e, float(nrm) / 2.0 ** e


@ref_only
@pytest.mark.parametrize("kind", ["d", "z", "s", "c"])
def test_bjac2_internal_forms_vs_ref(kind):
    dt = np.float64 if kind in "dz" else np.float32
    vl = 8 if dt == np.float64 else 16
    n = 4096
    ins = [a for a in W.jac2_batch(n, 31, kind in "zc", dtype=dt) if a is not None]
    g = W.rng(32)
    p0 = g.integers(0, 1 << 16, n // vl).astype(np.uint32)
    p0[::3] &= np.uint32(0xFFFFFF00 if vl == 8 else 0xFFFF0000)  # some blocks gated off
    for ns in (n, -n):
        r1, o1, q1 = vjo.bjac2(kind, ns, ins, internal=True, p_init=p0)
        r2, o2, q2 = vjo.bjac2(kind, ns, ins, internal=True, use_ref=True, p_init=p0)
        # lanes of gated-off blocks are not written by either: compare the rest
        live = np.repeat((p0 & np.uint32(0xFF if vl == 8 else 0xFFFF)) != 0, vl)
        for a, b in zip(o1, o2):
            assert bit_mismatch(a[live], b[live]) == 0
        assert np.array_equal(q1, q2)


@ref_only
@pytest.mark.parametrize("m", [8, 64, 1000, 4096])
def test_dot_kernels_vs_ref(m):
    R = vjo.ref()
    L = vjo.oracle()
    g = W.rng(m)
    for trial in range(20):
        sx, sy = np.ldexp(1.0, g.integers(-600, 600)), np.ldexp(1.0, g.integers(-600, 600))
        x = vjo.aligned(m, np.float64)
        y = vjo.aligned(m, np.float64)
        x[...] = g.standard_normal(m) * sx
        y[...] = g.standard_normal(m) * sy + (0.5 * x / sx * sy if trial % 2 else 0)
        ex, fx = _ef(x)
        ey, fy = _ef(y)
        e = vjo.aligned(2, np.float64)
        f = vjo.aligned(2, np.float64)
        e[...] = (ex, ey)
        f[...] = (fx, fy)
        mm = C.c_uint64(m)
        a = L.vjo_ddpscl(m, x.ctypes.data, y.ctypes.data, e.ctypes.data, f.ctypes.data)
        b = R.ddpscl_(C.byref(mm), x.ctypes.data, y.ctypes.data, e.ctypes.data, f.ctypes.data)
        assert np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64)
        # complex: x^H y / (|x| |y|)
        xi = vjo.aligned(m, np.float64)
        yi = vjo.aligned(m, np.float64)
        xi[...] = g.standard_normal(m) * sx
        yi[...] = g.standard_normal(m) * sy
        zr = C.c_double()
        zi = C.c_double()
        L.vjo_zdpscl(m, x.ctypes.data, xi.ctypes.data, y.ctypes.data, yi.ctypes.data, e.ctypes.data,
                     f.ctypes.data, C.byref(zr), C.byref(zi))
        z = R.zdpscl_(C.byref(mm), x.ctypes.data, xi.ctypes.data, y.ctypes.data, yi.ctypes.data, e.ctypes.data,
                      f.ctypes.data)
        assert zr.value == z.re and zi.value == z.im
        # FP32
        xs = vjo.aligned(m, np.float32)
        ys = vjo.aligned(m, np.float32)
        xs[...] = x / sx
        ys[...] = y / sy
        exs, fxs = _ef(xs)
        eys, fys = _ef(ys)
        es = vjo.aligned(2, np.float32)
        fs = vjo.aligned(2, np.float32)
        es[...] = (exs, eys)
        fs[...] = (fxs, fys)
        if m % 16 == 0:
            a = L.vjo_sdpscl(m, xs.ctypes.data, ys.ctypes.data, es.ctypes.data, fs.ctypes.data)
            b = R.sdpscl_(C.byref(mm), xs.ctypes.data, ys.ctypes.data, es.ctypes.data, fs.ctypes.data)
            assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)


sec_only = pytest.mark.skipif(vjo.ref_sec() is None, reason="USE_SECANTS reference build not present")


@ref_only
@pytest.mark.parametrize("sec", [False, pytest.param(True, marks=sec_only)])
@pytest.mark.parametrize("n", [8, 64, 512])
def test_rotation_and_gs_kernels_vs_ref(n, sec):
    R = vjo.ref_sec() if sec else vjo.ref()
    L = vjo.oracle()
    g = W.rng(n + 1)
    ctx = vjo.secants(sec)
    ctx.__enter__()
    for trial in range(12):
        x0 = g.standard_normal(n) * np.ldexp(1.0, g.integers(-50, 50))
        y0 = g.standard_normal(n) * np.ldexp(1.0, g.integers(-50, 50))
        # a cosine (<= 1), or a secant (>= 1) in the USE_SECANTS form
        c = [1.0, np.sqrt(1.0 + 0.3 ** 2), 2.0][trial % 3] if sec else [1.0, 1.0 / np.sqrt(1.0 + 0.3 ** 2), 0.5][trial % 3]
        t = [0.0, 0.3, -1.7][(trial // 3) % 3]
        for sgn in (1, -1):
            xs = [vjo.aligned(n, np.float64) for _ in range(2)]
            ys = [vjo.aligned(n, np.float64) for _ in range(2)]
            for a, b in zip(xs, ys):
                a[...] = x0
                b[...] = y0
            cc, tt = C.c_double(c), C.c_double(t)
            r1 = L.vjo_djrot(sgn * n, xs[0].ctypes.data, ys[0].ctypes.data, c, t)
            r2 = R.djrot_(C.byref(C.c_int64(sgn * n)), xs[1].ctypes.data, ys[1].ctypes.data, C.byref(cc), C.byref(tt))
            assert r1 == r2 and bit_mismatch(xs[0], xs[1]) == 0 and bit_mismatch(ys[0], ys[1]) == 0
            # complex rotation
            arrs = [[vjo.aligned(n, np.float64) for _ in range(4)] for _ in range(2)]
            for grp in arrs:
                grp[0][...] = x0
                grp[1][...] = y0[::-1]
                grp[2][...] = y0
                grp[3][...] = x0[::-1]
            sat = [0.0, 0.2, -0.9][trial % 3]
            cat = t
            r1 = L.vjo_zjrot(sgn * n, *[a.ctypes.data for a in arrs[0]], c, cat, sat)
            r2 = R.zjrot_(C.byref(C.c_int64(sgn * n)), *[a.ctypes.data for a in arrs[1]], C.byref(C.c_double(c)),
                          C.byref(C.c_double(cat)), C.byref(C.c_double(sat)))
            assert r1 == r2
            for a, b in zip(arrs[0], arrs[1]):
                assert bit_mismatch(a, b) == 0
            # Gram-Schmidt
            xs2 = [vjo.aligned(n, np.float64) for _ in range(2)]
            ys2 = [vjo.aligned(n, np.float64) for _ in range(2)]
            for a, b in zip(xs2, ys2):
                a[...] = x0
                b[...] = y0
            ex, fx = _ef(x0)
            ey, fy = _ef(y0)
            e = vjo.aligned(2, np.float64)
            f = vjo.aligned(2, np.float64)
            e[...] = (ex, ey)
            f[...] = (fx, fy)
            tcos = 0.37 * (1 if trial % 2 else -1)
            r1 = L.vjo_dgsscl(sgn * n, tcos, xs2[0].ctypes.data, ys2[0].ctypes.data, e.ctypes.data, f.ctypes.data)
            r2 = R.dgsscl_(C.byref(C.c_int64(sgn * n)), C.byref(C.c_double(tcos)), xs2[1].ctypes.data,
                           ys2[1].ctypes.data, e.ctypes.data, f.ctypes.data)
            assert r1 == r2 and bit_mismatch(xs2[0], xs2[1]) == 0 and bit_mismatch(ys2[0], ys2[1]) == 0
    ctx.__exit__(None, None, None)


@ref_only
@pytest.mark.parametrize("m,n,seed", [(64, 16, 1), (96, 64, 2), (200, 80, 3), (256, 256, 4)])
def test_dvjsvd_vs_ref(m, n, seed):
    from tests.golden.make_golden import ref_dvjsvd
    A = W.gaussian_matrix(m, n, seed)
    if seed == 3:  # graded columns: high dynamic range
        A *= np.ldexp(1.0, (np.arange(n) * 11 - 400))[:, None]
    rc_r, G_r, V_r, eS_r, fS_r = ref_dvjsvd(A)
    js, stp = _pjs_table(n)
    G = vjo.aligned((n, m), np.float64)
    G[...] = A
    with vjo.norm_mode(0):
        rc, eS, fS, V, _, _ = vjo.vjsvd_core("d", G, js=js, stp=stp, swp=999, strict=True)
    assert rc == rc_r
    assert bit_mismatch(G, G_r) == 0 and bit_mismatch(V, V_r) == 0
    assert bit_mismatch(eS, eS_r) == 0 and bit_mismatch(fS, fS_r) == 0


@sec_only
@pytest.mark.parametrize("m,n,seed", [(64, 16, 1), (96, 64, 2), (200, 80, 3)])
def test_dvjsvd_secants_vs_ref(m, n, seed):
    """The reference's dvjsvd_ built with USE_SECANTS (the 2x2 EVD returns
    secants, dvjsvd.c:411-415, and the rotations divide by them) against the
    oracle's secant form, bit for bit; the two forms do differ."""
    from tests.golden.make_golden import ref_dvjsvd
    A = W.gaussian_matrix(m, n, seed)
    if seed == 3:
        A *= np.ldexp(1.0, (np.arange(n) * 11 - 400))[:, None]
    rc_r, G_r, V_r, eS_r, fS_r = ref_dvjsvd(A, R=vjo.ref_sec())
    rc_c, G_c, V_c, _, _ = ref_dvjsvd(A)
    js, stp = _pjs_table(n)
    G = vjo.aligned((n, m), np.float64)
    G[...] = A
    with vjo.norm_mode(0), vjo.secants(True):
        rc, eS, fS, V, _, _ = vjo.vjsvd_core("d", G, js=js, stp=stp, swp=999, strict=True)
    assert rc == rc_r
    assert bit_mismatch(G, G_r) == 0 and bit_mismatch(V, V_r) == 0
    assert bit_mismatch(eS, eS_r) == 0 and bit_mismatch(fS, fS_r) == 0
    assert bit_mismatch(V_c, V_r) > 0
