"""Writes the golden fixtures of tests/golden/ from the reference build.

    make -C oracle           # builds oracle/_ref/libvjref.so from the reference sources
    python tests/golden/make_golden.py

The fixtures are inputs and the reference's outputs for them:
  bjac2.npz   dbjac2_/zbjac2_/sbjac2_/cbjac2_ (sine and tangent forms) on
              seeded inputs of the 2^26-batch distribution plus special values;
  dvjsvd.npz  dvjsvd_ (built with its own long-double norm dnormf_) on small
              seeded matrices and on the graded 16x16 matrix of etc/dgsex.c,
              driven by this package's ME step table, which is stored with
              each case.
The reference itself is never copied: only these input/output vectors are.
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import vjo  # noqa: E402
from tests.util import special_jac2  # noqa: E402
from vecjac_amd import svd as S  # noqa: E402
from vecjac_amd import workloads as W  # noqa: E402

OUT = Path(__file__).resolve().parent


def golden_bjac2() -> None:
    R = vjo.ref()
    assert R is not None, "reference build missing (make -C oracle)"
    data = {}
    for kind in "dzsc":
        dt = np.float64 if kind in "dz" else np.float32
        rnd = [a for a in W.jac2_batch(2048, 77, kind in "zc", dtype=dt) if a is not None]
        spc = special_jac2(kind in "zc", dt)
        ins = [np.concatenate([a, b]) for a, b in zip(rnd, spc)]
        nn = ins[0].size
        for k, a in enumerate(ins):
            data[f"{kind}_in{k}"] = a
        for form, ns in (("tan", nn), ("sin", -nn)):
            rc, outs, p = vjo.bjac2(kind, ns, ins, use_ref=True)
            for k, a in enumerate(outs):
                data[f"{kind}_{form}_out{k}"] = np.array(a)
            data[f"{kind}_{form}_p"] = np.array(p)
    np.savez_compressed(OUT / "bjac2.npz", **data)


def dgsex(N: int = 16) -> np.ndarray:
    """etc/dgsex.c:11-16, columns as rows."""
    A = np.zeros((N, N))
    for j in range(N):
        for i in range(N):
            e = (1023 - j - abs(i - j)) if j < N // 2 else (-51 - j - abs(i - j))
            A[j, i] = np.ldexp(1.0, e)
    return A


def ref_dvjsvd(A: np.ndarray, swp: int = 999, js=None, R=None):
    R = R or vjo.ref()
    n, m = A.shape
    if js is None:
        js, _ = S.pjs(2, n)
    stp = js.shape[0]
    jsa = vjo.aligned(js.shape, np.uint32)
    jsa[...] = js
    G = vjo.aligned((n, m), np.float64)
    G[...] = A
    V = vjo.aligned((n, n), np.float64)
    eS = vjo.aligned(n, np.float64)
    fS = vjo.aligned(n, np.float64)
    work = vjo.aligned(8 * n, np.float64)
    iw = vjo.aligned(max(1, n // 8), np.uint32)
    mm, nn = C.c_uint64(m), C.c_uint64(n)
    rc = R.dvjsvd_(C.byref(mm), C.byref(nn), G.ctypes.data, C.byref(mm), V.ctypes.data, C.byref(nn),
                   eS.ctypes.data, fS.ctypes.data, jsa.ctypes.data, C.byref(C.c_uint(stp)), C.byref(C.c_uint(swp)),
                   work.ctypes.data, iw.ctypes.data)
    return rc, np.array(G), np.array(V), np.array(eS), np.array(fS)


def golden_dvjsvd() -> None:
    cases = {
        "g64x32": W.gaussian_matrix(64, 32, 1),
        "g128x48": W.gaussian_matrix(128, 48, 2),
        "dgsex16": dgsex(16),
        "rankdef48x32": None,
        "g64x32_swp2": W.gaussian_matrix(64, 32, 3),
    }
    g = W.rng(4)
    B = g.standard_normal((48, 16))
    Cm = g.standard_normal((16, 32))
    cases["rankdef48x32"] = (B @ Cm).T.copy()  # rank 16
    data = {}
    for name, A in cases.items():
        swp = 2 if name.endswith("swp2") else 999
        js, _ = S.pjs(2, A.shape[0])
        rc, G, V, eS, fS = ref_dvjsvd(A, swp, js)
        data[f"{name}_A"] = A
        data[f"{name}_js"] = js
        data[f"{name}_swp"] = np.array(swp)
        data[f"{name}_rc"] = np.array(rc)
        data[f"{name}_U"] = G
        data[f"{name}_V"] = V
        data[f"{name}_eS"] = eS
        data[f"{name}_fS"] = fS
    np.savez_compressed(OUT / "dvjsvd.npz", **data)


if __name__ == "__main__":
    golden_bjac2()
    golden_dvjsvd()
    for f in sorted(OUT.glob("*.npz")):
        print(f.name, f.stat().st_size)
