"""The allocation-free entry point vjh_dvjsvd_noalloc, through the C ABI,
against vjh_dvjsvd (which allocates its scratch per call): two calls in a
row on the same GPU, the second with a smaller input, each bit-identical to
vjh_dvjsvd on the same input (U, V, eS, fS, the return code and the
per-sweep statistics).  The second call runs on the scratch the first one
left behind, so a value that must start cleared but is not shows up here.
Also: an n above the scratch's limit is refused with -2 and a message, and
a table whose row is not a permutation with -9, as vjh_dvjsvd does."""
import ctypes as C

import numpy as np
import pytest

from vecjac_amd import _lib
from vecjac_amd import svd as S
from vecjac_amd import workloads as W

pytestmark = pytest.mark.gpu

NOALLOC_MAX_N = 16384  # VJH_NOALLOC_MAX_N in include/vecjac.h


def _declare():
    L = _lib.lib()
    f = L.vjh_dvjsvd_noalloc
    f.restype = C.c_int64
    f.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_uint, C.c_uint, C.c_uint, C.c_void_p, C.c_void_p]
    return L


def _call(name, A, js, swp=999, V0=None):
    import torch
    L = _declare()
    n, m = A.shape
    G = torch.from_numpy(A.copy()).cuda()
    V = torch.from_numpy(V0.copy()).cuda() if V0 is not None else torch.full((n, n), np.nan, dtype=torch.float64,
                                                                              device="cuda")
    eS = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    fS = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    jsd = torch.from_numpy(np.ascontiguousarray(js).reshape(-1).view(np.int32)).cuda()
    stats = (C.c_uint * 66)()
    s = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    print(f"{name} m={m} n={n} swp={swp}", flush=True)
    rc = getattr(L, name)(m, n, G.data_ptr(), m, V.data_ptr(), n, eS.data_ptr(), fS.data_ptr(), jsd.data_ptr(),
                          js.shape[0], swp, int(V0 is not None), s, stats)
    torch.cuda.synchronize()
    return dict(rc=rc, G=G.cpu().numpy(), V=V.cpu().numpy(), eS=eS.cpu().numpy(), fS=fS.cpu().numpy(),
                stats=list(stats))


def _same(a, b):
    assert a["rc"] == b["rc"], (a["rc"], b["rc"], _lib.last_error())
    # sweeps and per-sweep transformations; not the number of sweeps
    # launched, which for vjh_dvjsvd includes the one it queues ahead
    assert a["stats"][0] == b["stats"][0] and a["stats"][2:] == b["stats"][2:]
    for k in ("G", "V", "eS", "fS"):
        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), k


@pytest.mark.parametrize("sizes", [((1024, 512), (200, 64)), ((300, 96), (40, 16))])
def test_noalloc_matches_allocating_entry(sizes):
    for i, (m, n) in enumerate(sizes):
        A = W.gaussian_matrix(m, n, 71 + i)
        js, _ = S.pjs(2, n)
        ref = _call("vjh_dvjsvd", A, js)
        got = _call("vjh_dvjsvd_noalloc", A, js)
        assert ref["rc"] > 0
        _same(got, ref)


def test_noalloc_sweep_limit_and_preset_v():
    """A sweep-limited call (no epilogue, scaled working G) and then a call
    resumed from its G and V with a preset V, on the scratch the previous
    calls left."""
    m, n = 256, 64
    A = W.gaussian_matrix(m, n, 5)
    js, _ = S.pjs(2, n)
    ref = _call("vjh_dvjsvd", A, js, swp=2)
    got = _call("vjh_dvjsvd_noalloc", A, js, swp=2)
    assert ref["rc"] == 2
    _same(got, ref)
    ref2 = _call("vjh_dvjsvd", ref["G"], js, V0=ref["V"])
    got2 = _call("vjh_dvjsvd_noalloc", ref["G"], js, V0=ref["V"])
    assert ref2["rc"] > 0
    _same(got2, ref2)


def test_noalloc_limits_and_errors():
    L = _declare()
    A = W.gaussian_matrix(64, 16, 9)
    js, _ = S.pjs(2, 16)
    bad = js.copy()
    bad[1, 0] = bad[1, 1]
    assert _call("vjh_dvjsvd_noalloc", A, bad)["rc"] == -9
    n = NOALLOC_MAX_N + 2
    rc = L.vjh_dvjsvd_noalloc(n, n, None, n, None, n, None, None, None, 0, 1, 0, None, None)
    assert rc == -2 and "exceeds" in _lib.last_error()
    A[2, 3] = np.inf
    r = _call("vjh_dvjsvd_noalloc", A, js)
    assert r["rc"] == _call("vjh_dvjsvd", A, js)["rc"] == -80
