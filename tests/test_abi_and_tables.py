"""CPU checks of the C ABI and of the strategy tables (no GPU needed).

* libvecjac_hip.so loads and exports every function include/vecjac.h
  declares; liboracle_vj.so exports every function oracle/vjo.h declares.
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
