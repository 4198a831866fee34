"""GPU parity of the plain Jacobi SVD drivers (dvjsvd_/zvjsvd_/svjsvd_/cvjsvd_
and their device-memory forms) against the oracle.

Every parity case is checked against the oracle in both norm modes: mode 0
is the reference's own arithmetic (its long-double column norm and 2Sum
dots, bit-identical to the reference build), mode 1 the GPU's correctly
rounded norm and dot with everything else the reference's operation
sequence.  Bars:
  * sigma: relative error <= 8 n eps against the oracle on the same input and
    the same step table, in each mode; the number
    of bitwise identical sigma is reported;
  * U, V: orthonormal columns and ||A - U S V^H||_F / ||A||_F, both
    evaluated in long double, within a small multiple of n eps;
  * return codes: identical to the oracle (sweep counts within one sweep).
"""
import numpy as np
import pytest

from oracle import vjo
from vecjac_amd import svd as S
from vecjac_amd import workloads as W

pytestmark = pytest.mark.gpu

EPS = {"d": np.finfo(np.float64).eps, "z": np.finfo(np.float64).eps,
       "s": np.finfo(np.float32).eps, "c": np.finfo(np.float32).eps}


def _mat(kind, m, n, seed):
    dt = np.float64 if kind in "dz" else np.float32
    A = W.gaussian_matrix(m, n, seed).astype(dt)
    Ai = W.gaussian_matrix(m, n, seed + 1000).astype(dt) if kind in "zc" else None
    return A, Ai


def _run_gpu(kind, A, Ai, strategy=2, swp=999):
    import torch
    n, m = A.shape
    js, stp = S.pjs(strategy, n)
    dt = torch.float64 if kind in "dz" else torch.float32
    G = torch.from_numpy(A.copy()).cuda()
    V = torch.empty((n, n), dtype=dt, device="cuda")
    Gi = torch.from_numpy(Ai.copy()).cuda() if Ai is not None else None
    Vi = torch.empty((n, n), dtype=dt, device="cuda") if Ai is not None else None
    eS = torch.empty(n, dtype=dt, device="cuda")
    fS = torch.empty(n, dtype=dt, device="cuda")
    jsd = torch.from_numpy(js.reshape(-1).view(np.int32)).cuda()
    print(f"gpu {kind}vjsvd m={m} n={n}", flush=True)
    rc, stats = S.device_vjsvd(kind, G, V, eS, fS, jsd, stp, swp, Gi=Gi, Vi=Vi)
    torch.cuda.synchronize()
    out = dict(rc=rc, stats=stats, G=G.cpu().numpy(), V=V.cpu().numpy(), eS=eS.cpu().numpy(), fS=fS.cpu().numpy())
    if Gi is not None:
        out["Gi"] = Gi.cpu().numpy()
        out["Vi"] = Vi.cpu().numpy()
    return out


def _run_oracle(kind, A, Ai, strategy=2, swp=999, mode=1):
    n, m = A.shape
    js, stp = S.pjs(strategy, n)
    dt = A.dtype
    G = vjo.aligned((n, m), dt)
    G[...] = A
    Gi = None
    if Ai is not None:
        Gi = vjo.aligned((n, m), dt)
        Gi[...] = Ai
    with vjo.norm_mode(mode):
        rc, eS, fS, V, Vi, st = vjo.vjsvd_core(kind, G, js=js, stp=stp, swp=swp, Gi=Gi)
    return dict(rc=rc, G=G, Gi=Gi, V=V, Vi=Vi, eS=eS, fS=fS, st=st)


def _checks(kind, A, Ai, r, tolmul=8.0):
    """Long-double orthogonality and reconstruction of one result."""
    n, m = A.shape
    eps = EPS[kind]
    ld = np.longdouble
    U = r["G"].astype(ld).T
    V = r["V"].astype(ld).T
    Ac = A.astype(ld).T
    sig = S.sigma(r["eS"], r["fS"]).astype(ld)
    if Ai is None:
        orthU = np.max(np.abs(U.T @ U - np.eye(n, dtype=ld)))
        orthV = np.max(np.abs(V.T @ V - np.eye(n, dtype=ld)))
        R = Ac - (U * sig) @ V.T
        rec = np.sqrt(np.sum(R * R)) / np.sqrt(np.sum(Ac * Ac))
    else:
        Ui = r["Gi"].astype(ld).T
        Vi = r["Vi"].astype(ld).T
        Aci = Ai.astype(ld).T
        # complex products in long double via real arithmetic
        def cmul(ar, ai, br, bi):
            return ar @ br - ai @ bi, ar @ bi + ai @ br
        gr, gi = cmul(U.T, -Ui.T, U, Ui)
        orthU = max(np.max(np.abs(gr - np.eye(n, dtype=ld))), np.max(np.abs(gi)))
        hr, hi = cmul(V.T, -Vi.T, V, Vi)
        orthV = max(np.max(np.abs(hr - np.eye(n, dtype=ld))), np.max(np.abs(hi)))
        pr, pi = cmul(U * sig, Ui * sig, V.T, -Vi.T)
        Rr, Ri = Ac - pr, Aci - pi
        rec = np.sqrt(np.sum(Rr * Rr + Ri * Ri)) / np.sqrt(np.sum(Ac * Ac + Aci * Aci))
    lim = tolmul * n * eps
    assert orthU <= lim, f"U orthogonality {orthU} > {lim}"
    assert orthV <= lim, f"V orthogonality {orthV} > {lim}"
    assert rec <= lim, f"reconstruction {rec} > {lim}"


@pytest.mark.parametrize("kind,m,n", [
    ("d", 40, 16), ("d", 100, 32), ("d", 250, 64), ("d", 500, 48), ("d", 1000, 32), ("d", 2000, 16),
    ("d", 3000, 16), ("d", 128, 128),
    ("z", 96, 32), ("z", 600, 24), ("z", 1500, 16),
    ("s", 100, 32), ("s", 700, 32),
    ("c", 80, 32), ("c", 1100, 16),
    # columns streamed in chunks (longer than the register-held pairs)
    ("z", 5000, 16), ("d", 9000, 16), ("c", 4500, 8), ("s", 8500, 8), ("z", 3000, 16), ("c", 2100, 8),
    # one pair: a single step per sweep, folded by the end-of-sweep launch alone
    ("d", 7, 2), ("z", 5, 2), ("d", 300, 2),
])
def test_plain_vs_oracle(kind, m, n):
    A, Ai = _mat(kind, m, n, seed=m * 7 + n)
    g = _run_gpu(kind, A, Ai)
    assert g["rc"] >= 0, g["rc"]
    sg = S.sigma(g["eS"], g["fS"])
    for mode in (1, 0):
        o = _run_oracle(kind, A, Ai, mode=mode)
        assert o["rc"] >= 0, o["rc"]
        assert abs(g["rc"] - o["rc"]) <= 1, (mode, g["rc"], o["rc"])
        so = S.sigma(o["eS"], o["fS"])
        # compare the sorted spectra (ties in the column order are allowed)
        sg_s, so_s = np.sort(sg), np.sort(so)
        rel = np.max(np.abs(sg_s - so_s) / so_s)
        same = int(np.sum(sg == so))
        print(f"{kind} m={m} n={n} mode {mode}: sweeps gpu={g['rc']} oracle={o['rc']} max rel err={rel:.3e} "
              f"({rel / (8 * n * EPS[kind]):.3f} x 8n eps) bit-identical sigma {same}/{n}")
        assert rel <= 8 * n * EPS[kind], (mode, rel)
    _checks(kind, A, Ai, g)


def test_plain_identity_and_zero():
    n = 16
    A = np.eye(n)
    g = _run_gpu("d", A, None)
    assert g["rc"] == 0
    assert np.array_equal(S.sigma(g["eS"], g["fS"]), np.ones(n))
    assert np.array_equal(g["G"], np.eye(n)) and np.array_equal(g["V"], np.eye(n))
    Z = np.zeros((n, 24))
    g = _run_gpu("d", Z, None)
    assert g["rc"] == 0 and np.all(np.isneginf(g["eS"])) and np.all(g["fS"] == 1.0)
    assert np.array_equal(g["V"], np.eye(n))


def test_plain_nonfinite_input():
    A, _ = _mat("d", 64, 16, 3)
    A[3, 7] = np.nan
    g = _run_gpu("d", A, None)
    assert g["rc"] == -80  # dvjsvd.c:80
    A[3, 7] = np.inf
    g = _run_gpu("d", A, None)
    assert g["rc"] == -80
    Ar, Ai = _mat("z", 64, 16, 3)
    Ai[2, 2] = np.inf
    assert _run_gpu("z", Ar, Ai)["rc"] == -92  # zvjsvd.c:92


def test_plain_sweep_limit():
    A, _ = _mat("d", 64, 32, 5)
    g = _run_gpu("d", A, None, swp=2)
    for mode in (1, 0):
        assert _run_oracle("d", A, None, swp=2, mode=mode)["rc"] == 2
    assert g["rc"] == 2


def test_plain_graded_gram_schmidt():
    # the graded 16x16 stress matrix of etc/dgsex.c:11-16: huge and tiny
    # column groups (exercises the Gram-Schmidt codes and the scaling)
    N = 16
    A = np.zeros((N, N))
    for j in range(N):
        for i in range(N):
            e = (1023 - j - abs(i - j)) if j < N // 2 else (-51 - j - abs(i - j))
            A[j, i] = np.ldexp(1.0, e)
    g = _run_gpu("d", A, None)
    assert g["rc"] >= 0
    sg = np.sort(S.sigma(g["eS"], g["fS"]))
    for mode in (1, 0):
        o = _run_oracle("d", A, None, mode=mode)
        assert o["rc"] >= 0 and abs(g["rc"] - o["rc"]) <= 1, (mode, g["rc"], o["rc"])
        so = np.sort(S.sigma(o["eS"], o["fS"]))
        assert np.max(np.abs(sg - so) / so) <= 8 * N * EPS["d"], mode


def test_plain_host_entry_matches_device():
    m, n = 256, 64
    A, _ = _mat("d", m, n, 21)
    sig, U, V, rc = S.dvjsvd_host(A, m, n)
    g = _run_gpu("d", A, None)
    assert rc == g["rc"]
    assert np.array_equal(sig, S.sigma(g["eS"], g["fS"]))
    assert np.array_equal(U, g["G"]) and np.array_equal(V, g["V"])


def test_plain_host_argument_codes():
    import ctypes as C
    m, n = 64, 16
    js, stp = S.pjs(2, n)
    G = S.aligned((n, m), np.float64, 1.0)
    V = S.aligned((n, n), np.float64)
    eS = S.aligned(n, np.float64)
    fS = S.aligned(n, np.float64)
    work = S.aligned(5 * n, np.float64)
    iw = S.aligned(2, np.uint32)
    call = lambda **kw: S.vjsvd("d", kw.get("m", m), kw.get("n", n), kw.get("G", G), kw.get("ldG", m), V,
                                kw.get("ldV", n), eS, fS, kw.get("js", js), stp, 999, work, iw)
    assert call(m=60) == -1           # m % 8
    assert call(n=14) == -2           # (n/2) % 8
    assert call(ldG=60) == -4         # ldG < m
    assert call(ldV=8) == -6          # ldV < n
    raw = np.empty(n * m + 1)
    mis = raw[1:] if raw.ctypes.data % 64 == 0 else raw[:-1]
    mis = mis.reshape(n, m)
    if mis.ctypes.data % 64:
        assert call(G=mis) == -3      # alignment
    bad = js.copy()
    bad[0, 0] = bad[0, 1]
    assert call(js=bad) == -9         # not a permutation


@pytest.mark.parametrize("kind", ["d", "z"])
def test_streamed_wide_range_and_zero_column(kind):
    """Streamed columns (m > 8192) whose elements span more than 2^440: the
    one-pass sums under the predicted exponent fail the range check there and
    the sums are taken again; a zero column has no predicted exponent."""
    m, n = 9000, 8
    A, Ai = _mat(kind, m, n, seed=77)
    A[0, :100] *= 2.0 ** -600       # tiny elements in a column with normal ones
    A[3, :] = 0.0                   # a zero column
    if Ai is not None:
        Ai[0, :100] *= 2.0 ** -600
        Ai[3, :] = 0.0
    g = _run_gpu(kind, A, Ai)
    sg = np.sort(S.sigma(g["eS"], g["fS"]))
    for mode in (1, 0):
        o = _run_oracle(kind, A, Ai, mode=mode)
        assert g["rc"] >= 0 and abs(g["rc"] - o["rc"]) <= 1, (mode, g["rc"], o["rc"])
        so = np.sort(S.sigma(o["eS"], o["fS"]))
        nz = so > 0
        assert np.all(sg[~nz] == 0) and np.sum(~nz) == 1
        assert np.max(np.abs(sg[nz] - so[nz]) / so[nz]) <= 8 * n * EPS[kind], mode
