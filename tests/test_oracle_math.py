"""Pins the oracle by what the mathematics fixes (independent of the
reference build), chosen so that a dropped term, a wrong sign or index or a
transposed operand fails at least one check.

2x2 EVD: problems built from known eigenpairs as the reference's generator
does (rnd.c:136-191, in extended precision); residual ||U L U^T - A||/||A||,
|c^2 + s^2 - 1| and eigenvalue errors measured in long double as the
reference's wnrme.c does (worr/wrer/wlam); the closed-form special cases
b = 0 and a = d; eigenvalues against LAPACK (numpy.linalg.eigvalsh).

SVD: sigma against LAPACK on well-conditioned matrices; high relative
accuracy on a graded matrix against a 50-digit mpmath SVD (which LAPACK's
bidiagonal SVD does not achieve); identity, diagonal and zero known answers;
the reference's error codes.
"""
import numpy as np
import pytest

from oracle import vjo
from vecjac_amd import workloads as W

LD = np.longdouble
EPS = np.finfo(np.float64).eps


def _known_pairs(n, seed, cplx=False):
    g = W.rng(seed)
    l1 = (g.uniform(-1, 1, n) * np.exp2(g.uniform(-30, 30, n))).astype(LD)
    l2 = (g.uniform(-1, 1, n) * np.exp2(g.uniform(-30, 30, n))).astype(LD)
    t = np.tan(g.uniform(-np.pi / 4, np.pi / 4, n)).astype(LD)
    c2 = 1 + t * t
    a = (l1 + l2 * t * t) / c2
    d = (l2 + l1 * t * t) / c2
    b = t * (l1 - l2) / c2
    if not cplx:
        return l1, l2, a.astype(np.float64), d.astype(np.float64), b.astype(np.float64)
    phi = g.uniform(-np.pi, np.pi, n)
    return l1, l2, a.astype(np.float64), d.astype(np.float64), (b * np.cos(phi)).astype(np.float64), \
        (b * np.sin(phi)).astype(np.float64)


def test_jac2_real_known_eigenpairs():
    n = 1 << 14
    l1, l2, a, d, b = _known_pairs(n, 1)
    rc, (c, s, L1, L2), p = vjo.bjac2("d", -n, [a, d, b])
    c, s, L1, L2 = (x.astype(LD) for x in (c, s, L1, L2))
    A11, A22, A21 = a.astype(LD), d.astype(LD), b.astype(LD)
    # A = U diag(L1, L2) U^T with U = [[c, -s], [s, c]] (wnrme.c:35-54)
    r11 = c * c * L1 + s * s * L2 - A11
    r22 = s * s * L1 + c * c * L2 - A22
    r21 = c * s * (L1 - L2) - A21
    res = np.sqrt(r11 ** 2 + r22 ** 2 + 2 * r21 ** 2) / np.sqrt(A11 ** 2 + A22 ** 2 + 2 * A21 ** 2)
    assert np.max(res) <= 8 * EPS
    assert np.max(np.abs(c * c + s * s - 1)) <= 4 * EPS
    # eigenvalues vs the generated ones: norm-wise relative error (wlam)
    Lmax = np.maximum(np.abs(L1), np.abs(L2))
    lmax = np.maximum(np.abs(l1), np.abs(l2))
    assert np.max(np.abs(Lmax - lmax) / lmax) <= 8 * EPS
    # L1 < L2 bits
    bits = np.unpackbits(p.view(np.uint8), bitorder="little").reshape(-1, 32)[:, :8].reshape(-1)
    assert np.array_equal(bits.astype(bool), (L1 < L2))


def test_jac2_complex_known_eigenpairs():
    n = 1 << 13
    l1, l2, a, d, br, bi = _known_pairs(n, 2, cplx=True)
    rc, (c, ca, sa, L1, L2), p = vjo.bjac2("z", -n, [a, d, br, bi])
    c, ca, sa, L1, L2 = (x.astype(LD) for x in (c, ca, sa, L1, L2))
    s2 = ca * ca + sa * sa
    assert np.max(np.abs(c * c + s2 - 1)) <= 4 * EPS
    A11, A22, Br, Bi = a.astype(LD), d.astype(LD), br.astype(LD), bi.astype(LD)
    r11 = c * c * L1 + s2 * L2 - A11
    r22 = s2 * L1 + c * c * L2 - A22
    r21r = c * ca * (L1 - L2) - Br
    r21i = c * sa * (L1 - L2) - Bi
    res = np.sqrt(r11 ** 2 + r22 ** 2 + 2 * (r21r ** 2 + r21i ** 2)) / \
        np.sqrt(A11 ** 2 + A22 ** 2 + 2 * (Br ** 2 + Bi ** 2))
    assert np.max(res) <= 8 * EPS


def test_jac2_special_cases():
    a = np.array([3.0, -2.0, 5.0, 5.0, 0.0, 7.0, 7.0, 1e-300])
    d = np.array([1.0, 4.0, 5.0, 5.0, 0.0, 7.0, -7.0, 2e-300])
    b = np.array([0.0, -0.0, 2.0, -3.0, 0.0, 1e-10, 0.0, 0.0])
    rc, (c, t, L1, L2), p = vjo.bjac2("d", 8, [a, d, b])
    # b = 0: identity rotation, eigenvalues are the diagonal
    for k in (0, 1, 4, 6, 7):
        assert c[k] == 1.0 and t[k] == 0.0 and L1[k] == a[k] and L2[k] == d[k]
    # a = d, b != 0: tan(2 phi) clipped to sqrt(huge), tan(phi) = +-1 up to rounding
    for k in (2, 3):
        assert abs(abs(t[k]) - 1.0) <= 2 * EPS
        assert abs(max(L1[k], L2[k]) - (a[k] + abs(b[k]))) <= 4 * EPS * abs(a[k])
        assert abs(min(L1[k], L2[k]) - (a[k] - abs(b[k]))) <= 4 * EPS * abs(a[k])
    assert np.sign(t[2]) == 1 and np.sign(t[3]) == -1


def test_jac2_vs_lapack_eigvalsh():
    n = 4096
    a, d, b = W.jac2_batch(n, 3, False, frac_b0=0.0, frac_ad=0.0)
    rc, (c, t, L1, L2), p = vjo.bjac2("d", n, [a, d, b])
    M = np.stack([np.stack([a, b], -1), np.stack([b, d], -1)], -2)
    ev = np.linalg.eigvalsh(M)
    ours = np.sort(np.stack([L1, L2], -1), axis=-1)
    scale = np.max(np.abs(ev), axis=-1, keepdims=True)
    assert np.max(np.abs(ours - ev) / scale) <= 8 * EPS


def _svd(A, swp=999, strict=False, iwork0=0):
    from vecjac_amd import svd as S
    n, m = A.shape
    js, stp = S.pjs(2, n)
    G = vjo.aligned((n, m), np.float64)
    G[...] = A
    rc, eS, fS, V, _, st = vjo.vjsvd_core("d", G, js=js, stp=stp, swp=swp, strict=strict, iwork0=iwork0)
    return rc, vjo.sigma(eS, fS), G, V


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("m,n", [(40, 16), (120, 64), (64, 64)])
def test_svd_vs_lapack(mode, m, n):
    A = W.gaussian_matrix(m, n, m + n)
    with vjo.norm_mode(mode):
        rc, sig, U, V = _svd(A)
    assert rc > 0
    sv = np.linalg.svd(A.T, compute_uv=False)
    assert np.max(np.abs(np.sort(sig)[::-1] - sv) / sv) <= 8 * n * EPS
    Ul, Vl = U.astype(LD), V.astype(LD)
    assert np.max(np.abs(Ul @ Ul.T - np.eye(n, dtype=LD))) <= 8 * n * EPS
    assert np.max(np.abs(Vl @ Vl.T - np.eye(n, dtype=LD))) <= 8 * n * EPS
    R = A.T.astype(LD) - (Ul.T * sig.astype(LD)) @ Vl
    assert np.sqrt(np.sum(R * R)) / np.sqrt(np.sum(A.astype(LD) ** 2)) <= 8 * n * EPS


def test_svd_high_relative_accuracy_graded():
    mpmath = pytest.importorskip("mpmath")
    m, n = 12, 8
    g = W.rng(9)
    B = g.standard_normal((m, n))
    D = np.exp2(-40.0 * np.arange(n))  # column scales down to 2^-280
    A = (B * D).T.copy()  # columns as rows
    rc, sig, _, _ = _svd(A)
    assert rc >= 0
    mpmath.mp.dps = 60
    M = mpmath.matrix([[mpmath.mpf(float(A[j, i])) for j in range(n)] for i in range(m)])
    s_mp = sorted([float(x) for x in mpmath.svd_r(M, compute_uv=False)], reverse=True)
    ours = np.sort(sig)[::-1]
    kappa_B = np.linalg.cond(B)
    rel = np.abs(ours - np.array(s_mp)) / np.array(s_mp)
    assert np.max(rel) <= 8 * n * EPS * kappa_B, (rel, kappa_B)


def test_svd_known_answers():
    n = 16
    rc, sig, U, V = _svd(np.eye(n))
    assert rc == 0 and np.array_equal(sig, np.ones(n)) and np.array_equal(U, np.eye(n))
    dg = np.diag(np.array([3.0, -0.5, 8.0, 1e-200, 2.0, 4.0, -7.0, 1.0, 6.0, 5.0, 0.25, 9.0, 1e100, -3.0, 2.5, 1.5]))
    rc, sig, U, V = _svd(dg)
    assert rc == 0 and np.array_equal(np.sort(sig), np.sort(np.abs(np.diag(dg))))
    rc, sig, U, V = _svd(np.zeros((n, 24)))
    assert rc == 0 and np.all(sig == 0) and np.array_equal(V, np.eye(n))


def test_svd_error_codes():
    A = W.gaussian_matrix(64, 16, 1)
    A[2, 3] = np.nan
    assert _svd(A)[0] == -80
    A[2, 3] = -np.inf
    assert _svd(A)[0] == -80
    A = W.gaussian_matrix(64, 16, 1)
    assert _svd(A, iwork0=1)[0] == -13          # preset V that is zero
    assert _svd(W.gaussian_matrix(60, 16, 1), strict=True)[0] == -1    # m % 8
    assert _svd(W.gaussian_matrix(64, 14, 1), strict=True)[0] == -2    # (n/2) % 8
    assert _svd(W.gaussian_matrix(64, 16, 1), swp=1)[0] == 1          # not converged
