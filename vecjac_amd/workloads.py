"""Seeded synthetic inputs for the SVD and 2x2-EVD workloads (numpy Philox).

Every generator takes an explicit seed and returns numpy arrays, so the GPU
path, the CPU baselines and the tests all see identical inputs.
"""
from __future__ import annotations

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def jac2_batch(n: int, seed: int, cplx: bool, dtype=np.float64, frac_b0: float = 0.01, frac_ad: float = 0.01):
    """2x2 Hermitian problems [[a, b], [conj(b), d]]: every component has a
    random sign and magnitude 2^e, e ~ U[-100, 100] (continuous); a fraction
    `frac_b0` has b = 0 and a fraction `frac_ad` has a = d."""
    g = rng(seed)

    def comp():
        e = g.uniform(-100.0, 100.0, n)
        s = np.where(g.integers(0, 2, n) == 1, -1.0, 1.0)
        return (s * np.exp2(e)).astype(dtype)

    a, d, br = comp(), comp(), comp()
    bi = comp() if cplx else None
    k = np.arange(n)
    kb0 = g.permutation(k)[: int(round(frac_b0 * n))]
    kad = g.permutation(k)[: int(round(frac_ad * n))]
    br[kb0] = 0
    if cplx:
        bi[kb0] = 0
    d[kad] = a[kad]
    return (a, d, br, bi) if cplx else (a, d, br)


def gaussian_matrix(m: int, n: int, seed: int, dtype=np.float64, ld: int | None = None):
    """m x n column-major N(0,1) matrix with leading dimension ld (>= m)."""
    ld = m if ld is None else ld
    g = rng(seed)
    A = np.zeros((n, ld), dtype=dtype)  # row j of this array = column j
    A[:, :m] = g.standard_normal((n, m))
    return A


def haar_unitary(m: int, n: int, g: np.random.Generator, cplx: bool):
    """m x n matrix with orthonormal columns, Haar distributed (QR of a
    Gaussian with the phases of diag(R) removed)."""
    Z = g.standard_normal((m, n))
    if cplx:
        Z = Z + 1j * g.standard_normal((m, n))
    Q, R = np.linalg.qr(Z)
    d = np.diag(R)
    return Q * (d / np.abs(d))


def graded_svd_matrix(m: int, n: int, sigma: np.ndarray, seed: int, cplx: bool):
    """A = Q1 diag(sigma) Q2^H with Haar Q1 (m x n) and Q2 (n x n)."""
    g = rng(seed)
    Q1 = haar_unitary(m, n, g, cplx)
    Q2 = haar_unitary(n, n, g, cplx)
    return (Q1 * sigma) @ Q2.conj().T


def batch_shapes(nb: int, seed: int, nmin: int = 8, nmax: int = 128):
    """Uneven batch: n ~ U{nmin..nmax}, m ~ U{n..4n}."""
    g = rng(seed)
    n = g.integers(nmin, nmax + 1, nb)
    m = np.array([g.integers(int(k), 4 * int(k) + 1) for k in n], dtype=np.int64)
    return m.astype(np.int64), n.astype(np.int64)


def gapped_sigma(n: int, g: np.random.Generator, lo: float = 1.0, hi: float = 10.0, gap: float = 1e-3):
    """n values ~ U[lo, hi] whose sorted neighbours differ by at least
    gap * (larger value) (rejection sampling)."""
    while True:
        s = np.sort(g.uniform(lo, hi, n))[::-1]
        if n < 2 or np.all((s[:-1] - s[1:]) >= gap * s[:-1]):
            return s


def grad_batch(nb: int, m: int, n: int, seed: int):
    """Differentiable-SVD workload: A_b = Q1 diag(sigma) Q2^T with Haar Q1
    (m x n), Q2 (n x n), sigma ~ U[1, 10] with minimum relative gap 1e-3;
    loss weights w (n), G_U (m x n), G_V (n x n) seeded Gaussian.
    Returns A (nb, m, n), w, GU, GV as float64 arrays."""
    g = rng(seed)
    A = np.empty((nb, m, n))
    for b in range(nb):
        s = gapped_sigma(n, g)
        Q1 = haar_unitary(m, n, g, False)
        Q2 = haar_unitary(n, n, g, False)
        A[b] = (Q1 * s) @ Q2.T
    w = g.standard_normal((nb, n))
    GU = g.standard_normal((nb, m, n))
    GV = g.standard_normal((nb, n, n))
    return A, w, GU, GV
