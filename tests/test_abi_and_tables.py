"""CPU checks of the C ABI and of the strategy tables (no GPU needed).

* libvecjac_hip.so and its LP64 twin libvecjac_hip_lp64.so load and export
  every function include/vecjac.h declares, with 64- and 32-bit reference
  integers respectively (vjh_fint_bits); the LP64 twin's host zsplit_ /
  zmerge_ take 32-bit arguments; liboracle_vj.so exports every function
  oracle/vjo.h declares.
* pjs(): every row is a permutation of 0..n-1 (disjoint pairs, race-free
  column updates); id 2 (ME) covers every unordered pair exactly once per
  sweep in n-1 steps; id 4 (MM) covers every pair at least once in n steps;
  odd n and unknown ids are refused, as pjs.c:14-26 refuses them.
"""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared(header: Path):
    txt = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)
    skip = {"if", "for", "while", "return", "sizeof", "defined"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_hip_library_exports_header():
    import ctypes as C
    from vecjac_amd import _lib
    L = C.CDLL(str(_lib.LIB_PATH))
    names = _declared(ROOT / "include" / "vecjac.h")
    assert "dvjsvd_" in names and "dbjac2_" in names and "pjs" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_lp64_library_exports_header():
    import ctypes as C
    from vecjac_amd import _lib
    lp = _lib.LIB_PATH.with_name("libvecjac_hip_lp64.so")
    L = C.CDLL(str(lp))
    names = _declared(ROOT / "include" / "vecjac.h")
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.vjh_fint_bits() == 32
    assert C.CDLL(str(_lib.LIB_PATH)).vjh_fint_bits() == 64


def test_lp64_host_zsplit_zmerge():
    """zsplit_/zmerge_ of the LP64 library read 32-bit m, n and leading
    dimensions (zsplit.c:3, zmerge.c:3 under LP64)."""
    import ctypes as C
    from vecjac_amd import _lib
    L = C.CDLL(str(_lib.LIB_PATH.with_name("libvecjac_hip_lp64.so")))
    for nm in ("zsplit_", "zmerge_"):
        getattr(L, nm).restype = C.c_int32
        getattr(L, nm).argtypes = [C.c_void_p] * 8
    U = lambda v: C.byref(C.c_uint32(v))  # noqa: E731
    g = np.random.default_rng(3)
    m, n, ld, ldr = 37, 11, 40, 39
    A = (g.standard_normal((n, ld)) + 1j * g.standard_normal((n, ld))).astype(np.complex128)
    Ar = np.zeros((n, ldr))
    Ai = np.zeros((n, ldr))
    assert L.zsplit_(U(m), U(n), A.ctypes.data, U(ld), Ar.ctypes.data, U(ldr), Ai.ctypes.data, U(ldr)) == 0
    assert np.array_equal(Ar[:, :m], A.real[:, :m]) and np.array_equal(Ai[:, :m], A.imag[:, :m])
    B = np.zeros((n, ld), dtype=np.complex128)
    assert L.zmerge_(U(m), U(n), Ar.ctypes.data, U(ldr), Ai.ctypes.data, U(ldr), B.ctypes.data, U(ld)) == 0
    assert np.array_equal(B[:, :m], A[:, :m])
    assert L.zsplit_(U(m), U(n), A.ctypes.data, U(m - 1), Ar.ctypes.data, U(ldr), Ai.ctypes.data, U(ldr)) == -4


def test_oracle_library_exports_header():
    import ctypes as C
    from oracle import vjo
    L = C.CDLL(str(vjo.ORACLE_SO))
    names = _declared(ROOT / "oracle" / "vjo.h")
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("n", [2, 4, 8, 16, 30, 64, 128, 250])
def test_pjs_me(n):
    from vecjac_amd import svd as S
    t, stp = S.pjs(2, n)
    assert stp == n - 1 and t.shape == (n - 1, n)
    cover = np.zeros((n, n), dtype=int)
    for r in t:
        assert np.array_equal(np.sort(r), np.arange(n))
        p, q = r[0::2], r[1::2]
        assert np.all(p < q)
        cover[p, q] += 1
    iu = np.triu_indices(n, 1)
    assert np.all(cover[iu] == 1)


@pytest.mark.parametrize("n", [2, 4, 8, 16, 30, 64, 128])
def test_pjs_mm(n):
    from vecjac_amd import svd as S
    t, stp = S.pjs(4, n)
    assert stp == n and t.shape == (n, n)
    cover = np.zeros((n, n), dtype=int)
    for r in t:
        assert np.array_equal(np.sort(r), np.arange(n))
        cover[r[0::2], r[1::2]] += 1
    iu = np.triu_indices(n, 1)
    assert np.all(cover[iu] >= 1)
    assert int(cover[iu].sum()) == n * n // 2


@pytest.mark.parametrize("n", [2, 4, 6, 8, 10, 16, 30, 64, 126, 128])
def test_pjs_block_modulus(n):
    """id 6 against its definition: round s meets block I = {2I, 2I+1} with
    block J = s - I mod n/2, crossing (2I, 2J), (2I+1, 2J+1) at step 2s and
    (2I, 2J+1), (2I+1, 2J) at step 2s+1, a block met by itself pairing its
    own columns at both steps."""
    from vecjac_amd import svd as S
    t, stp = S.pjs(6, n)
    assert stp == n and t.shape == (n, n)
    nb = n // 2
    cover = np.zeros((n, n), dtype=int)
    for st, r in enumerate(t):
        assert np.array_equal(np.sort(r), np.arange(n))
        p, q = r[0::2], r[1::2]
        assert np.all(p < q)
        cover[p, q] += 1
        s, h = divmod(st, 2)
        want = set()
        for I in range(nb):
            J = (s - I) % nb
            if I == J:
                want.add((2 * I, 2 * I + 1))
            elif I < J:
                want |= {(2 * I, 2 * J + h), (2 * I + 1, 2 * J + 1 - h)}
        assert set(zip(p.tolist(), q.tolist())) == want
    iu = np.triu_indices(n, 1)
    inblock = (iu[1] == iu[0] + 1) & (iu[0] % 2 == 0)
    assert np.all(cover[iu][inblock] == 2) and np.all(cover[iu][~inblock] == 1)


def test_pjs_refusals():
    from vecjac_amd import svd as S
    for args in ((2, 7), (4, 1), (3, 8), (2, 0), (6, 5), (5, 8)):
        with pytest.raises(ValueError):
            S.pjs(*args)
