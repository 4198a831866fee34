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


# ---------------------------------------------------------------------------
# The complex and FP32 drivers (zvjsvd.c:40-694, svjsvd.c, cvjsvd.c) and the
# correctly rounded column norm / scaled dot of norm mode 1.
# ---------------------------------------------------------------------------

_DT = {"d": np.float64, "z": np.float64, "s": np.float32, "c": np.float32}
_EPS = {"d": EPS, "z": EPS, "s": float(np.finfo(np.float32).eps), "c": float(np.finfo(np.float32).eps)}


def _svd_kind(kind, A, strategy=2, swp=999, iwork0=0, V0=None):
    """Oracle driver of `kind` on A (m x n numpy, complex for z/c); returns
    (rc, sigma, U (m x n complex/real), V (n x n), stats)."""
    from vecjac_amd import svd as S
    m, n = A.shape
    dt = _DT[kind]
    js, stp = S.pjs(strategy, n)
    G = vjo.aligned((n, m), dt)
    G[...] = A.real.T
    Gi = None
    if kind in "zc":
        Gi = vjo.aligned((n, m), dt)
        Gi[...] = A.imag.T
    V = Vi = None
    if V0 is not None:
        V = vjo.aligned((n, n), dt)
        V[...] = V0.real.T
        if kind in "zc":
            Vi = vjo.aligned((n, n), dt)
            Vi[...] = V0.imag.T
    rc, eS, fS, V, Vi, st = vjo.vjsvd_core(kind, G, V=V, js=js, stp=stp, swp=swp, Gi=Gi, Vi=Vi,
                                           iwork0=iwork0)
    U = G.T.astype(np.float64) + (1j * Gi.T.astype(np.float64) if Gi is not None else 0)
    Vm = V.T.astype(np.float64) + (1j * Vi.T.astype(np.float64) if Vi is not None else 0)
    return rc, vjo.sigma(eS, fS), U, Vm, st


def _ld_checks(A, sig, U, V, lim):
    """Orthogonality of U and V and ||A - U S V^H||_F / ||A||_F in long
    double (complex products through their real and imaginary parts, as the
    reference's wnrme does in extended precision)."""
    def split(X):
        return np.real(X).astype(LD), np.imag(X).astype(LD)

    def mul(ar, ai, br, bi):
        return ar @ br - ai @ bi, ar @ bi + ai @ br

    n = sig.size
    Ur, Ui = split(U)
    Vr, Vi = split(V)
    I = np.eye(n, dtype=LD)
    gr, gi = mul(Ur.T, -Ui.T, Ur, Ui)
    hr, hi = mul(Vr.T, -Vi.T, Vr, Vi)
    s = sig.astype(LD)
    pr, pi = mul(Ur * s, Ui * s, Vr.T, -Vi.T)
    Ar, Ai = split(A)
    Rr, Ri = Ar - pr, Ai - pi
    orthU = max(np.max(np.abs(gr - I)), np.max(np.abs(gi)))
    orthV = max(np.max(np.abs(hr - I)), np.max(np.abs(hi)))
    rec = np.sqrt(np.sum(Rr * Rr + Ri * Ri)) / np.sqrt(np.sum(Ar * Ar + Ai * Ai))
    assert orthU <= lim and orthV <= lim and rec <= lim, (orthU, orthV, rec, lim)


@pytest.mark.parametrize("kind", ["z", "s", "c"])
@pytest.mark.parametrize("m,n", [(40, 16), (96, 32), (64, 64)])
@pytest.mark.parametrize("mode", [0, 1])
def test_svd_kinds_vs_lapack(kind, m, n, mode):
    """z/s/c drivers: sigma against LAPACK (complex128 / float64 SVD of the
    same data) within 8 n eps of the working precision; U, V orthonormal and
    the reconstruction in long double."""
    g = W.rng(100 * m + n)
    A = g.standard_normal((m, n))
    if kind in "zc":
        A = A + 1j * g.standard_normal((m, n))
    A = A.astype(np.complex64 if kind == "c" else np.float32 if kind == "s" else A.dtype).astype(
        np.complex128 if kind in "zc" else np.float64)
    with vjo.norm_mode(mode):
        rc, sig, U, V, st = _svd_kind(kind, A)
    assert rc > 0
    sv = np.linalg.svd(A, compute_uv=False)
    eps = _EPS[kind]
    assert np.max(np.abs(np.sort(sig)[::-1] - sv) / sv) <= 8 * n * eps
    _ld_checks(A, sig, U, V, 8 * n * eps)


@pytest.mark.parametrize("kind", ["z", "c"])
def test_svd_complex_known_answers(kind):
    """Identity: no sweep, sigma = 1, U = V = I.  A diagonal with complex
    phases has orthogonal columns: no rotation (only swaps that order the
    columns, code -1 of dvjsvd.c:422-459), sigma = |d_i| to rounding and V
    a permutation matrix.  A zero matrix: sigma = 0."""
    n = 16
    rc, sig, U, V, _ = _svd_kind(kind, np.eye(n, dtype=np.complex128))
    assert rc == 0 and np.array_equal(sig, np.ones(n)) and np.array_equal(U, np.eye(n))
    g = W.rng(31)
    mag = np.exp2(g.integers(-60, 60, n)).astype(np.float64)
    ph = np.exp(1j * g.uniform(-np.pi, np.pi, n))
    d = (mag * ph).astype(np.complex64 if kind == "c" else np.complex128).astype(np.complex128)
    rc, sig, U, V, _ = _svd_kind(kind, np.diag(d))
    assert rc == 0
    assert np.allclose(np.sort(sig), np.sort(np.abs(d)), rtol=4 * _EPS[kind], atol=0)
    assert np.array_equal(np.sort(V, axis=0)[-1], np.ones(n)) and np.count_nonzero(V) == n
    rc, sig, U, V, _ = _svd_kind(kind, np.zeros((24, n), dtype=np.complex128))
    assert rc == 0 and np.all(sig == 0)


def test_svd_complex_conjugation_pinned():
    """A Hermitian-form slip (x^T y in place of x^H y, or a sign of the
    imaginary part of the rotation) leaves the columns of a matrix with
    purely imaginary couplings non-orthogonal: sigma, U^H U and the
    reconstruction are all checked against a matrix built with known
    singular values and complex singular vectors."""
    m, n = 48, 16
    g = W.rng(17)
    sig0 = np.linspace(4.0, 1.0, n)
    A = W.graded_svd_matrix(m, n, sig0, 17, True)
    rc, sig, U, V, _ = _svd_kind("z", A)
    assert rc > 0
    assert np.max(np.abs(np.sort(sig)[::-1] - sig0) / sig0) <= 8 * n * EPS
    _ld_checks(A, sig, U, V, 8 * n * EPS)


@pytest.mark.parametrize("kind", ["z", "c"])
def test_svd_complex_high_relative_accuracy_graded(kind):
    """A = B D with B complex Gaussian (well conditioned) and D a column
    scaling over many binades: one-sided Jacobi keeps every sigma to a
    relative error of order n eps kappa(B), far below what the scaling
    would allow a normwise-accurate method; checked against a 40-digit
    mpmath SVD of the exact stored matrix (svd_c)."""
    mpmath = pytest.importorskip("mpmath")
    m, n = 12, 8
    g = W.rng(23 if kind == "z" else 24)
    B = g.standard_normal((m, n)) + 1j * g.standard_normal((m, n))
    span = 200.0 if kind == "z" else 24.0
    D = np.exp2(-span / (n - 1) * np.arange(n))
    A = B * D
    if kind == "c":
        A = A.astype(np.complex64).astype(np.complex128)
    rc, sig, _, _, _ = _svd_kind(kind, A)
    assert rc >= 0
    mpmath.mp.dps = 40
    M = mpmath.matrix([[mpmath.mpc(float(A[i, j].real), float(A[i, j].imag)) for j in range(n)]
                       for i in range(m)])
    s_mp = np.array(sorted([float(x) for x in mpmath.svd_c(M, compute_uv=False)], reverse=True))
    rel = np.abs(np.sort(sig)[::-1] - s_mp) / s_mp
    assert np.max(rel) <= 8 * n * _EPS[kind] * np.linalg.cond(B), rel


def test_svd_float_high_relative_accuracy_graded():
    """svjsvd: the FP32 driver on a column-graded matrix against mpmath."""
    mpmath = pytest.importorskip("mpmath")
    m, n = 12, 8
    g = W.rng(29)
    B = g.standard_normal((m, n))
    A = (B * np.exp2(-3.0 * np.arange(n))).astype(np.float32).astype(np.float64)
    rc, sig, _, _, _ = _svd_kind("s", A)
    assert rc >= 0
    mpmath.mp.dps = 40
    M = mpmath.matrix([[mpmath.mpf(float(A[i, j])) for j in range(n)] for i in range(m)])
    s_mp = np.array(sorted([float(x) for x in mpmath.svd_r(M, compute_uv=False)], reverse=True))
    rel = np.abs(np.sort(sig)[::-1] - s_mp) / s_mp
    assert np.max(rel) <= 8 * n * _EPS["s"] * np.linalg.cond(B), rel


@pytest.mark.parametrize("kind", ["d", "z"])
def test_svd_vpreset_resume(kind):
    """iwork[0] = 1 (dvjsvd.c:86-96, 143-157): V is taken as given and
    updated.  A run stopped after two sweeps returns its working G (A V
    2^sT, no epilogue, dvjsvd.c:604) and V; resumed from those two with
    iwork[0] = 1 it reaches the sigma of an uninterrupted run (scaled by
    2^sT, within 8 n eps) in the remaining sweeps (+-1), with A = U S V^H
    and V orthonormal.  A preset V that is zero is refused (-13, complex
    -17)."""
    m, n = 64, 32
    g = W.rng(41)
    A = g.standard_normal((m, n))
    if kind == "z":
        A = A + 1j * g.standard_normal((m, n))
    rc_full, sig_full, _, _, _ = _svd_kind(kind, A)
    rc1, _, G1, V1, st1 = _svd_kind(kind, A, swp=2)
    assert rc1 == 2 and st1.sT != 0
    rc2, sig2, U2, V2, _ = _svd_kind(kind, G1, iwork0=1, V0=V1)
    assert rc2 >= 0 and abs((rc1 + rc2) - rc_full) <= 1, (rc1, rc2, rc_full)
    sig2 = np.ldexp(sig2, -st1.sT)
    assert np.max(np.abs(np.sort(sig2) - np.sort(sig_full)) / np.sort(sig_full)) <= 8 * n * EPS
    _ld_checks(A, sig2, U2, V2, 8 * n * EPS)
    rc, *_ = _svd_kind(kind, A, iwork0=1, V0=np.zeros((n, n)))
    assert rc == (-17 if kind == "z" else -13)


# ------------------------- exact norms and dots -------------------------

def _exact_sumsq(parts):
    """Sum of squares of binary floating-point values as an exact integer N
    and binary exponent E: sum = N * 2^(2E)."""
    fr = [np.frexp(np.asarray(p, dtype=np.float64)) for p in parts]
    E = min(int(np.min(e)) for _, e in fr) - 53
    N = 0
    for mant, e in fr:
        for mm, ee in zip(mant.tolist(), e.tolist()):
            if mm != 0.0:
                q = int(np.ldexp(mm, 53))  # exact 53-bit integer mantissa
                N += (q << (ee - 53 - E)) ** 2
    return N, E


def _cr_sqrt_ef(N, E, bits):
    """Correctly rounded sqrt(N * 2^(2E)) to `bits` significant bits, as
    (e, f) with f in [1, 2): q = round(sqrt(N) 2^s) with 2^(bits-1) <= q <
    2^bits, from the integer square root and an exact midpoint test."""
    import math
    if N == 0:
        return -np.inf, 1.0

    def floor_q(s):
        return math.isqrt(N << (2 * s)) if s >= 0 else math.isqrt(N >> (-2 * s))

    s = (2 * bits - N.bit_length()) // 2
    while floor_q(s) >= (1 << bits):
        s -= 1
    while floor_q(s) < (1 << (bits - 1)):
        s += 1
    q = floor_q(s)
    # sqrt(N) 2^s > q + 1/2  <=>  4 N 4^s > (2q + 1)^2
    lhs, rhs = 4 * N, (2 * q + 1) ** 2
    up = (lhs << (2 * s)) > rhs if s >= 0 else lhs > (rhs << (-2 * s))
    q += int(up)
    if q == (1 << bits):
        q >>= 1
        s -= 1
    return float(E - s + bits - 1), q / float(1 << (bits - 1))


@pytest.mark.parametrize("span", [(-1074, -900), (-1000, 1000), (900, 1023), (-30, 30)])
@pytest.mark.parametrize("cplx", [False, True])
def test_norm_cr_exact(span, cplx):
    """Norm mode 1 (the GPU's column norm) returns the correctly rounded
    ||x||_2 in (e, f) form on vectors whose elements span 2^-1074 .. 2^1023,
    subnormals and values whose squares overflow included, against the
    exact sum of squares in integer arithmetic."""
    import math
    L = vjo.oracle()
    L.vjo_znorm2.argtypes  # declared by the binding
    g = W.rng(abs(span[0]) + span[1] + (7 if cplx else 0))
    for trial in range(40):
        m = int(g.integers(1, 300))
        ex = g.integers(span[0], span[1], m)
        x = np.ldexp(g.uniform(0.5, 1.0, m) * np.where(g.integers(0, 2, m) == 1, -1, 1), ex)
        if trial % 5 == 0:
            x[g.integers(0, m)] = 0.0
        xa = vjo.aligned(m, np.float64)
        xa[...] = x
        e0 = np.zeros(1)
        f0 = np.zeros(1)
        with vjo.norm_mode(1):
            if cplx:
                y = np.ldexp(g.uniform(0.5, 1.0, m), g.integers(span[0], span[1], m))
                ya = vjo.aligned(m, np.float64)
                ya[...] = y
                L.vjo_znorm2(m, xa.ctypes.data, ya.ctypes.data, e0.ctypes.data, f0.ctypes.data)
                N, E = _exact_sumsq([x, y])
            else:
                L.vjo_dnorm2(m, xa.ctypes.data, e0.ctypes.data, f0.ctypes.data)
                N, E = _exact_sumsq([x])
        e, f = _cr_sqrt_ef(N, E, 53)
        assert (e0[0], f0[0]) == (e, f), (trial, m, e0[0], f0[0], e, f)
    assert math.isinf(-np.inf)


def test_norm_cr_exact_float():
    """The FP32 norm of mode 1: correctly rounded to 24 bits (the sum of
    squares is accumulated in double from exact products; a double rounding
    can only matter within 2^-29 ulp of a tie, which random data avoids)."""
    L = vjo.oracle()
    import ctypes as C
    L.vjo_snorm2.restype = C.c_float
    L.vjo_snorm2.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
    g = W.rng(5)
    for trial in range(60):
        m = int(g.integers(1, 200))
        x = np.ldexp(g.uniform(0.5, 1.0, m), g.integers(-149 + 24, 127, m)).astype(np.float32)
        xa = vjo.aligned(m, np.float32)
        xa[...] = x
        e0 = np.zeros(1, np.float32)
        f0 = np.zeros(1, np.float32)
        with vjo.norm_mode(1):
            L.vjo_snorm2(m, xa.ctypes.data, e0.ctypes.data, f0.ctypes.data)
        N, E = _exact_sumsq([x.astype(np.float64)])
        e, f = _cr_sqrt_ef(N, E, 24)
        assert (float(e0[0]), float(f0[0])) == (e, f), (trial, e0[0], f0[0], e, f)


@pytest.mark.parametrize("cplx", [False, True])
def test_dot_cr_vs_exact(cplx):
    """Norm mode 1's scaled dot, sum_i (x_i 2^-ex)(y_i 2^-ey) / (fx fy)
    (ddpscl.c:36-74; complex x^H y, zdpscl.c): against the exact dot
    product (rational arithmetic) divided by fx fy, within 2 ulp (one
    rounding of the sum, one of fx fy, one of the quotient)."""
    from fractions import Fraction as Fr
    L = vjo.oracle()
    g = W.rng(61 + cplx)
    for trial in range(30):
        m = int(g.integers(2, 200))
        x = g.standard_normal(m) * np.exp2(g.integers(-300, 300))
        y = g.standard_normal(m) * np.exp2(g.integers(-300, 300))
        xi = g.standard_normal(m) * np.exp2(g.integers(-5, 5)) * np.max(np.abs(x)) if cplx else None
        yi = g.standard_normal(m) * np.max(np.abs(y)) if cplx else None
        arr = [vjo.aligned(m, np.float64) for _ in range(4)]
        for a, v in zip(arr, (x, y, xi, yi)):
            if v is not None:
                a[...] = v
        e = vjo.aligned(2, np.float64)
        f = vjo.aligned(2, np.float64)
        with vjo.norm_mode(1):
            for k, (vr, vi) in enumerate(((arr[0], arr[2] if cplx else None), (arr[1], arr[3] if cplx else None))):
                ek, fk = np.zeros(1), np.zeros(1)
                if cplx:
                    L.vjo_znorm2(m, vr.ctypes.data, vi.ctypes.data, ek.ctypes.data, fk.ctypes.data)
                else:
                    L.vjo_dnorm2(m, vr.ctypes.data, ek.ctypes.data, fk.ctypes.data)
                e[k], f[k] = ek[0], fk[0]
            if cplx:
                zr, zi = np.zeros(1), np.zeros(1)
                L.vjo_zdpscl(m, arr[0].ctypes.data, arr[2].ctypes.data, arr[1].ctypes.data, arr[3].ctypes.data,
                             e.ctypes.data, f.ctypes.data, zr.ctypes.data, zi.ctypes.data)
                got = [zr[0], zi[0]]
            else:
                got = [L.vjo_ddpscl(m, arr[0].ctypes.data, arr[1].ctypes.data, e.ctypes.data, f.ctypes.data)]
        sc = Fr(2) ** int(-(e[0] + e[1]))
        de = Fr(float(f[0] * f[1]))
        if cplx:
            re = sum(Fr(a) * Fr(b) + Fr(c) * Fr(d) for a, b, c, d in zip(x, y, xi, yi))
            im = sum(Fr(a) * Fr(d) - Fr(c) * Fr(b) for a, b, c, d in zip(x, y, xi, yi))
            want = [float(re * sc / de), float(im * sc / de)]
        else:
            want = [float(sum(Fr(a) * Fr(b) for a, b in zip(x, y)) * sc / de)]
        for gv, wv in zip(got, want):
            assert abs(gv - wv) <= 2 * np.spacing(abs(wv)), (trial, gv, wv)
