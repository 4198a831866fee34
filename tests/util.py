"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np


def bit_mismatch(a, b) -> int:
    """Number of positions whose bit patterns differ; two NaNs match (the
    default NaN of x86 and of the GPU differ in sign)."""
    a = np.asarray(a)
    b = np.asarray(b)
    ui = np.uint64 if a.dtype == np.float64 else np.uint32
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.sum((a.view(ui) != b.view(ui)) & ~both_nan))


def ulp_diff(a, b) -> np.ndarray:
    """|a - b| in units of the last place of b (0 where both are equal)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    sp = np.spacing(np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        d = np.abs(a - b) / sp
    d[a == b] = 0
    return d


def special_jac2(cplx: bool, dtype):
    """Hand-picked 2x2 inputs: zeros, equal diagonals, denormals, huge and
    tiny magnitudes, signs; a multiple of 16 rows."""
    f = np.finfo(dtype)
    tiny = f.smallest_subnormal
    vals = [0.0, -0.0, 1.0, -1.0, 2.5, tiny, -tiny, f.tiny, f.max, -f.max, f.max / 3, 1e-300 if dtype == np.float64 else 1e-38]
    rng = np.random.Generator(np.random.Philox(99))
    rows = []
    for a in vals:
        for d in vals:
            for b in vals[:8]:
                rows.append((a, d, b, rng.choice(vals) if cplx else 0.0))
    while len(rows) % 16:
        rows.append((1.0, 1.0, 0.0, 0.0))
    arr = np.array(rows, dtype=np.float64)
    with np.errstate(over="ignore"):
        cols = [arr[:, k].astype(dtype) for k in range(4)]
    return cols if cplx else cols[:3]


def nonfinite_jac2(cplx: bool, dtype):
    """2x2 inputs with infinities and NaNs among finite values (zeros,
    subnormals, the largest finite): every combination of the diagonal and
    the real part of a21, so that waves mix finite and non-finite lanes."""
    f = np.finfo(dtype)
    vals = [0.0, 1.0, -2.5, np.inf, -np.inf, np.nan, float(f.max), float(f.smallest_subnormal)]
    rng = np.random.Generator(np.random.Philox(7))
    rows = [(a, d, b, rng.choice(vals) if cplx else 0.0) for a in vals for d in vals for b in vals]
    arr = np.array(rows, dtype=np.float64)
    cols = [arr[:, k].astype(dtype) for k in range(4)]
    return cols if cplx else cols[:3]

