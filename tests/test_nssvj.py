"""LAPACK xGESVJ-style drivers dnssvj_/snssvj_/znssvj_/cnssvj_ (dnssvj.f:335,
snssvj.f:321, znssvj.f:349, cnssvj.f:349).

What pins them:
  * the argument checks and INFO codes, in the reference's order
    (dnssvj.f:402-431, znssvj.f:416-446) and the complex workspace query
    (znssvj.f:451-455): CPU tests, no kernel runs;
  * bitwise identity with the plain driver (the dvjsvd_ kernels, the same
    strategy table and threshold) for shapes that need no padding: sigma,
    U and V, permuted into non-increasing order;
  * sigma against the oracle's dvjsvd port on the same input and table
    (relative error <= 8 n eps, the plain driver's bar) and, for padded
    shapes, against LAPACK's SVD through numpy (error <= 8 n eps sigma_max);
  * U orthonormal, V orthogonal, ||A - U diag(sigma) V^H|| / ||A|| small, in
    long double;
  * LAPACK's documented outputs: SVA non-increasing with sigma = WORK(1) SVA,
    WORK(2) = nonzero count, WORK(3) = count above the underflow threshold,
    WORK(4) = sweep index at exit, WORK(5) < TOL and WORK(6) = 0 after
    convergence, INFO = NSWEEP - 1 without; the zero-matrix and one-column
    quick returns (dnssvj.f:599-621); JOBA 'L'/'U', JOBU 'N'/'C', JOBV 'A'.
"""
import numpy as np
import pytest

from oracle import vjo
from vecjac_amd import svd as S
from vecjac_amd import workloads as W

RT = {"d": np.float64, "s": np.float32, "z": np.float64, "c": np.float32}
CT = {"d": np.float64, "s": np.float32, "z": np.complex128, "c": np.complex64}


def _mat(kind, m, n, seed):
    A = W.gaussian_matrix(m, n, seed).T.astype(RT[kind])  # m x n
    if kind in "zc":
        A = (A + 1j * W.gaussian_matrix(m, n, seed + 1000).T.astype(RT[kind])).astype(CT[kind])
    return np.asfortranarray(A)


# ------------------------------------------------------------ CPU: arguments
def _call(kind, joba="G", jobu="U", jobv="V", m=4, n=3, lda=None, mv=0, ldv=None, lwork=None, nsweep=30,
          lrwork=None, w0=0.0):
    A = np.zeros((max(m, 1), max(n, 1)), CT[kind], order="F")
    V = np.zeros((8, 8), CT[kind], order="F")
    sva = np.zeros(8, RT[kind])
    lda = m if lda is None else lda
    ldv = max(n, 1) if ldv is None else ldv
    if kind in "ds":
        lwork = max(m + n, 6) if lwork is None else lwork
        work = np.zeros(max(lwork, 6), RT[kind])
        work[0] = w0
        return S.nssvj(kind, joba, jobu, jobv, m, n, A, lda, sva, mv, V, ldv, work, lwork, nsweep), work
    lwork = m + n if lwork is None else lwork
    lrwork = max(n, 6) if lrwork is None else lrwork
    cwork = np.zeros(max(lwork, 1), CT[kind])
    rwork = np.zeros(max(lrwork, 6), RT[kind])
    rwork[0] = w0
    info = S.nssvj(kind, joba, jobu, jobv, m, n, A, lda, sva, mv, V, ldv, cwork, lwork, nsweep, rwork, lrwork)
    return info, (cwork, rwork)


@pytest.mark.parametrize("kind", ["d", "s", "z", "c"])
def test_argument_codes(kind):
    assert _call(kind, joba="X")[0] == -1
    assert _call(kind, joba="x", jobu="Q")[0] == -1          # JOBA is checked first
    assert _call(kind, jobu="Q")[0] == -2
    assert _call(kind, jobv="Z")[0] == -3
    assert _call(kind, m=-1, n=0)[0] == -4
    assert _call(kind, m=2, n=3)[0] == -5
    assert _call(kind, n=-1)[0] == -5
    assert _call(kind, lda=3)[0] == -7
    assert _call(kind, jobv="A", mv=-1)[0] == -9
    assert _call(kind, jobv="V", ldv=2)[0] == -11
    assert _call(kind, jobv="A", mv=5, ldv=4)[0] == -11
    assert _call(kind, jobu="C", w0=1.0)[0] == -12             # CTOL must exceed one
    assert _call(kind, lwork=6)[0] == -13
    assert _call(kind, nsweep=0)[0] == (-14 if kind in "ds" else -16)
    # lower-case job letters are accepted (LSAME)
    assert _call(kind, joba="g", jobu="x")[0] == -2
    if kind in "zc":
        assert _call(kind, lrwork=5)[0] == -15
        assert _call(kind, jobu="F", jobv="J", nsweep=0)[0] == -16  # 'F' and 'J' are valid
    else:
        assert _call(kind, jobu="F")[0] == -2
        assert _call(kind, jobv="J")[0] == -3
        assert _call(kind, m=2, n=2, lwork=5)[0] == -13           # LWORK >= max(M+N, 6)


@pytest.mark.parametrize("kind", ["z", "c"])
def test_workspace_query(kind):
    info, (cwork, rwork) = _call(kind, m=7, n=3, lwork=-1)
    assert info == 0 and cwork[0] == 10 and rwork[0] == 6
    info, (cwork, rwork) = _call(kind, m=30, n=20, lrwork=-1)
    assert info == 0 and cwork[0] == 50 and rwork[0] == 20


@pytest.mark.parametrize("kind", ["d", "s", "z", "c"])
def test_void_matrix_quick_return(kind):
    for m, n in ((0, 0), (5, 0)):
        A = np.full((max(m, 1), 1), 7, CT[kind], order="F")
        sva = np.full(1, 3, RT[kind])
        if kind in "ds":
            work = np.full(8, 9, RT[kind])
            info = S.nssvj(kind, "G", "U", "V", m, n, A, max(m, 1), sva, 0, A, 1, work, 8, 30)
            assert info == 0 and np.all(work == 9)
        else:
            cwork = np.zeros(8, CT[kind])
            rwork = np.full(8, 9, RT[kind])
            info = S.nssvj(kind, "G", "U", "V", m, n, A, max(m, 1), sva, 0, A, 1, cwork, 8, 30, rwork, 8)
            assert info == 0 and np.all(rwork == 9)
        assert np.all(A == 7) and sva[0] == 3


def test_m_times_eps_limit():
    # single precision: M * eps >= 1 for M = 2^24 (the arrays are not read)
    info, _ = _call("s", m=1 << 24, n=1, lwork=6)
    assert info == -13
    A = np.zeros((1, 1), np.float32)
    work = np.zeros(8, np.float32)
    m = 1 << 24
    info = S.nssvj("s", "G", "U", "N", m, 1, A, m, np.zeros(1, np.float32), 0, A, 1, work, m + 1, 30)
    assert info == -4


# ------------------------------------------------------------------ GPU
def _long_checks(kind, A, sva, scale, U, V, k=None, tolmul=8.0):
    """U orthonormal (first k columns), V orthogonal, reconstruction."""
    m, n = A.shape
    k = n if k is None else k
    eps = np.finfo(RT[kind]).eps
    ld = np.clongdouble if kind in "zc" else np.longdouble
    Ul = U.astype(ld)[:, :k]
    Vl = V.astype(ld)
    sig = sva.astype(np.longdouble) * np.longdouble(scale)
    lim = tolmul * n * eps
    oU = np.max(np.abs(Ul.conj().T @ Ul - np.eye(k)))
    oV = np.max(np.abs(Vl.conj().T @ Vl - np.eye(n)))
    R = A.astype(ld) - (U.astype(ld) * sig) @ Vl.conj().T
    rec = np.sqrt(np.sum(np.abs(R) ** 2) / np.sum(np.abs(A.astype(ld)) ** 2))
    assert oU <= lim, f"U orthogonality {oU} > {lim}"
    assert oV <= lim, f"V orthogonality {oV} > {lim}"
    assert rec <= lim, f"reconstruction {rec} > {lim}"


def _device_plain(kind, A, swp=999):
    """The plain driver (vjh_?vjsvd) on the same input: (rc, sigma, U, V)."""
    import torch
    m, n = A.shape
    js, stp = S.pjs(2, n)
    dt = torch.float64 if kind in "dz" else torch.float32
    Gr = torch.from_numpy(np.ascontiguousarray(A.real.T)).cuda()
    V = torch.empty((n, n), dtype=dt, device="cuda")
    eS = torch.empty(n, dtype=dt, device="cuda")
    fS = torch.empty(n, dtype=dt, device="cuda")
    Gi = Vi = None
    if kind in "zc":
        Gi = torch.from_numpy(np.ascontiguousarray(A.imag.T)).cuda()
        Vi = torch.empty((n, n), dtype=dt, device="cuda")
    jsd = torch.from_numpy(js.reshape(-1).view(np.int32)).cuda()
    rc, _ = S.device_vjsvd(kind, Gr, V, eS, fS, jsd, stp, swp, Gi=Gi, Vi=Vi)
    torch.cuda.synchronize()
    U = Gr.cpu().numpy().T
    Vn = V.cpu().numpy().T
    if kind in "zc":
        U = U + 1j * Gi.cpu().numpy().T
        Vn = Vn + 1j * Vi.cpu().numpy().T
    return rc, S.sigma(eS.cpu().numpy(), fS.cpu().numpy()), U, Vn


def _bits(x):
    x = np.ascontiguousarray(x)
    return x.view(np.uint64 if x.dtype.itemsize in (8, 16) else np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,m,n", [("d", 64, 16), ("d", 300, 48), ("z", 96, 32), ("s", 120, 32), ("c", 72, 16)])
def test_bitwise_vs_plain_driver(kind, m, n):
    A = _mat(kind, m, n, seed=m + 3 * n)
    print(f"gpu {kind}nssvj vs plain m={m} n={n}", flush=True)
    sva, U, V, st, info = S.gesvj(A, nsweep=999)
    rc, sig, Up, Vp = _device_plain(kind, A)
    assert info == 0 and rc >= 0
    perm = np.argsort(-sig, kind="stable")
    assert st[0] == 1
    assert np.array_equal(_bits(sva), _bits(sig[perm].astype(RT[kind])))
    assert np.array_equal(_bits(U), _bits(Up[:, perm].astype(CT[kind])))
    assert np.array_equal(_bits(V), _bits(Vp[:, perm].astype(CT[kind])))
    assert st[1] == n and st[2] == n and st[3] == rc + 1
    tol = np.sqrt(RT[kind](m)) * np.finfo(RT[kind]).eps / 2
    assert 0 < st[4] < tol and st[5] == 0
    assert np.all(np.diff(sva) <= 0)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,m,n", [("d", 96, 32), ("z", 64, 16), ("s", 64, 16), ("c", 48, 16)])
def test_sigma_vs_oracle(kind, m, n):
    A = _mat(kind, m, n, seed=5 * m + n)
    print(f"gpu {kind}nssvj vs oracle m={m} n={n}", flush=True)
    sva, U, V, st, info = S.gesvj(A, nsweep=999)
    js, stp = S.pjs(2, n)
    G = vjo.aligned((n, m), RT[kind])
    G[...] = A.real.T
    Gi = None
    if kind in "zc":
        Gi = vjo.aligned((n, m), RT[kind])
        Gi[...] = A.imag.T
    G0, Gi0 = G.copy(), (Gi.copy() if Gi is not None else None)
    for mode in (1, 0):  # the GPU's norm and dot; the reference's own
        G[...] = G0
        if Gi is not None:
            Gi[...] = Gi0
        with vjo.norm_mode(mode):
            rc, eS, fS, _, _, _ = vjo.vjsvd_core(kind, G, js=js, stp=stp, swp=999, Gi=Gi)
        so = np.sort(vjo.sigma(eS, fS))[::-1]
        assert info == 0 and rc >= 0 and abs(st[3] - (rc + 1)) <= 1
        rel = np.max(np.abs(sva.astype(np.float64) * st[0] - so) / so)
        print(f"{kind} m={m} n={n} mode {mode} max rel err vs oracle {rel:.3e}")
        assert rel <= 8 * n * np.finfo(RT[kind]).eps
    _long_checks(kind, A, sva, st[0], U, V)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,m,n", [("d", 37, 37), ("d", 100, 33), ("d", 9, 5), ("d", 2, 2), ("d", 1000, 7),
                                      ("z", 21, 21), ("z", 50, 13), ("s", 80, 15), ("c", 31, 31)])
def test_padded_shapes_vs_lapack(kind, m, n):
    """Odd N (a zero column is appended) and M = N odd (a zero row too)."""
    A = _mat(kind, m, n, seed=11 * m + n)
    print(f"gpu {kind}nssvj padded m={m} n={n}", flush=True)
    sva, U, V, st, info = S.gesvj(A, nsweep=999)
    assert info == 0
    sl = np.linalg.svd(A.astype(np.complex128 if kind in "zc" else np.float64), compute_uv=False)
    err = np.max(np.abs(sva.astype(np.float64) * st[0] - sl)) / sl[0]
    print(f"{kind} m={m} n={n} sweeps={st[3]} max err / sigma_max {err:.3e}")
    assert err <= 8 * n * np.finfo(RT[kind]).eps
    assert np.all(np.diff(sva) <= 0) and st[1] == n and st[2] == n
    _long_checks(kind, A, sva, st[0], U, V)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["d", "z"])
def test_triangular_joba(kind):
    m, n = 70, 30
    A = _mat(kind, m, n, seed=77)
    L = np.tril(A)
    Uu = np.triu(A)
    print(f"gpu {kind}nssvj JOBA L/U m={m} n={n}", flush=True)
    rl = S.gesvj(A, joba="L")       # strict upper part of A is garbage
    rg = S.gesvj(L, joba="G")
    assert rl[4] == 0 and rg[4] == 0
    for a, b in zip(rl[:4], rg[:4]):
        assert np.array_equal(_bits(a), _bits(b))
    ru = S.gesvj(A, joba="u")
    rg = S.gesvj(Uu, joba="G")
    for a, b in zip(ru[:4], rg[:4]):
        assert np.array_equal(_bits(a), _bits(b))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["d", "c"])
def test_jobv_apply(kind):
    m, n, mv = 90, 24, 40
    A = _mat(kind, m, n, seed=31)
    V0 = _mat(kind, mv, n, seed=32)
    print(f"gpu {kind}nssvj JOBV=A m={m} n={n} mv={mv}", flush=True)
    sa, Ua, Va, sta, ia = S.gesvj(A, jobv="A", V0=V0)
    sv, Uv, Vv, stv, iv = S.gesvj(A, jobv="V")
    assert ia == 0 and iv == 0
    # G's rotations do not depend on V: same sigma and U bit for bit
    assert np.array_equal(_bits(sa), _bits(sv)) and np.array_equal(_bits(Ua), _bits(Uv))
    ref = V0.astype(np.clongdouble if kind in "zc" else np.longdouble) @ Vv
    err = np.max(np.abs(Va - ref)) / np.max(np.abs(ref))
    assert err <= 8 * n * np.finfo(RT[kind]).eps, err
    # an all-zero V stays zero, an empty one (MV = 0) is not touched
    sz, Uz, Vz, stz, iz = S.gesvj(A, jobv="A", V0=np.zeros((mv, n), CT[kind]))
    assert iz == 0 and np.all(Vz == 0) and np.array_equal(_bits(sz), _bits(sv))
    s0, U0, V0e, st0, i0 = S.gesvj(A, jobv="A", V0=np.zeros((0, n), CT[kind]))
    assert i0 == 0 and np.array_equal(_bits(s0), _bits(sv))


@pytest.mark.gpu
def test_jobu_n_and_c():
    m, n = 200, 40
    A = _mat("d", m, n, seed=41)
    print(f"gpu dnssvj JOBU=N/C m={m} n={n}", flush=True)
    sn, Un, _, stn, infon = S.gesvj(A, jobu="N", jobv="N")
    su, Uu, Vu, stu, infou = S.gesvj(A, jobu="U", jobv="V")
    assert infon == 0 and infou == 0
    eps = np.finfo(np.float64).eps / 2
    # singular values only: TOL = M eps, the columns of A are U * SVA
    assert stn[4] < m * eps and stn[3] <= stu[3] + 1
    assert np.max(np.abs(sn - su) / su) <= 8 * n * 2 * eps
    cn = np.linalg.norm(Un, axis=0)
    assert np.max(np.abs(cn - sn) / sn) <= 8 * n * 2 * eps
    # user-controlled threshold CTOL = 1e4: converged to 1e4 eps
    sc, Uc, Vc, stc, infoc = S.gesvj(A, jobu="C", ctol=1e4)
    assert infoc == 0 and stc[4] < 1e4 * eps and stc[3] <= stu[3] + 1
    _long_checks("d", A, sc, stc[0], Uc, Vc, tolmul=1e4)


@pytest.mark.gpu
def test_sweep_limit():
    m, n = 160, 64
    A = _mat("d", m, n, seed=51)
    print(f"gpu dnssvj NSWEEP=3 m={m} n={n}", flush=True)
    sva, U, V, st, info = S.gesvj(A, nsweep=3)
    # not converged: INFO = NSWEEP - 1, WORK(4) = NSWEEP + 1, rotations in
    # the last sweep, still a decomposition A = U diag(sigma) V^T with unit
    # columns of U and orthogonal V
    assert info == 2 and st[3] == 4 and st[5] > 0 and st[4] > 0
    assert np.all(np.diff(sva) <= 0)
    assert np.max(np.abs(np.linalg.norm(U, axis=0) - 1)) <= 8 * np.finfo(np.float64).eps * np.sqrt(m)
    ld = np.longdouble
    R = A.astype(ld) - (U.astype(ld) * (sva.astype(ld) * ld(st[0]))) @ V.astype(ld).T
    assert np.sqrt(np.sum(R * R) / np.sum(A.astype(ld) ** 2)) <= 8 * n * np.finfo(np.float64).eps
    assert np.max(np.abs(V.T @ V - np.eye(n))) <= 8 * n * np.finfo(np.float64).eps
    sf, Uf, Vf, stf, infof = S.gesvj(A, nsweep=999)
    assert infof == 0 and stf[3] > 4


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["d", "z"])
def test_zero_and_one_column(kind):
    print(f"gpu {kind}nssvj zero matrix / one column", flush=True)
    Z = np.zeros((12, 5), CT[kind], order="F")
    sva, U, V, st, info = S.gesvj(Z)
    assert info == 0 and np.all(sva == 0) and np.array_equal(U, np.eye(12, 5)) and np.array_equal(V, np.eye(5))
    assert list(st) == [1, 0, 0, 0, 0, 0]
    sva, U, V, st, info = S.gesvj(Z, jobu="N")   # A is left alone without U
    assert info == 0 and np.all(U == 0)
    a = _mat(kind, 30, 1, seed=61)
    sva, U, V, st, info = S.gesvj(a)
    nrm = np.linalg.norm(a)
    assert info == 0 and abs(sva[0] * st[0] - nrm) <= 4 * np.finfo(np.float64).eps * nrm
    assert abs(np.linalg.norm(U) - 1) <= 4 * np.finfo(np.float64).eps and V[0, 0] == 1
    assert list(st[1:]) == [1, 0, 0, 0, 0]
    sva, Uc, V, st, info = S.gesvj(a, jobu="C", ctol=10.0)  # one column: 'C' leaves A alone
    assert info == 0 and np.array_equal(Uc, a)


@pytest.mark.gpu
def test_rank_deficient_and_scaling():
    m, n = 64, 20
    A = _mat("d", m, n, seed=71)
    A[:, 3] = 0.0                        # a zero column
    A[:, 7] = A[:, 2]                    # a repeated column
    print(f"gpu dnssvj rank deficient m={m} n={n}", flush=True)
    sva, U, V, st, info = S.gesvj(A)
    assert info == 0 and st[1] == n - 1 and st[2] == n - 1   # one exact zero
    assert sva[-1] == 0 and np.all(U[:, -1] == 0)
    assert sva[-2] < 1e-12 * sva[0]                          # the repeated column
    # values beyond the double range: sigma = SCALE * SVA with SCALE = 2^k > 1
    B = _mat("d", 48, 16, seed=72)
    B = B / np.max(np.abs(B))
    big = np.ldexp(B, 1023)
    print("gpu dnssvj scaled beyond the double range", flush=True)
    sb, Ub, Vb, stb, infob = S.gesvj(big)
    sr, Ur, Vr, str_, infor = S.gesvj(B)
    assert infob == 0 and stb[0] > 1 and np.log2(stb[0]) == int(np.log2(stb[0]))
    k = int(np.log2(stb[0]))
    assert np.max(np.abs(np.ldexp(sb, k - 1023) - sr) / sr) <= 8 * 16 * np.finfo(np.float64).eps
    assert np.max(np.abs(Ub - Ur)) <= 1e-12
