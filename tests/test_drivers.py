"""The command-line drivers (vecjac_amd/bin, built from vecjac_amd/drivers)
and the complex layout entry points.

CPU: argument checking of the drivers (exit codes of tdvjsvd.c / tdbjac2.c,
decided before any GPU call), host zsplit_/zmerge_ against numpy.
GPU: each driver on files in the reference's raw formats, its outputs read
back and compared with the library called directly; vjh_zsplit/vjh_zmerge.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from vecjac_amd import workloads as W

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "vecjac_amd" / "bin"


def _run(exe, *args, timeout=300):
    return subprocess.run([str(BIN / exe), *map(str, args)], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("exe", ["vj_tdvjsvd", "vj_tzvjsvd"])
def test_svd_driver_argument_codes(tmp_path, exe):
    assert _run(exe).returncode == 1                       # usage
    base = tmp_path / "A"
    assert _run(exe, 2, 0, 16, base).returncode == 3       # m
    assert _run(exe, 2, 64, 15, base).returncode == 4      # odd n
    assert _run(exe, 2, 64, 24, base).returncode == 4      # n/2 not a multiple of 8
    assert _run(exe, 2, 16, 32, base).returncode == 4      # n > m
    assert _run(exe, 3, 64, 16, base).returncode == 2      # strategy
    assert _run(exe, 2, 64, 16, base).returncode == 5      # missing BaseName.G


def test_bjac2_driver_argument_codes(tmp_path):
    assert _run("vj_tdbjac2").returncode != 0
    assert _run("vj_tdbjac2", tmp_path / "x", 2, 1).returncode != 0   # 4 problems: not a multiple of 8
    assert _run("vj_tdbjac2", tmp_path / "x", 10, 0).returncode == 0  # nothing to do
    assert _run("vj_tdbjac2", tmp_path / "x", 10, 1).returncode != 0  # missing files


def _zlib():
    from vecjac_amd import _lib
    L = _lib.lib()
    u = C.POINTER(C.c_uint64)
    for nm in ("zsplit_", "zmerge_"):
        f = getattr(L, nm)
        f.restype = C.c_int64
        f.argtypes = [u, u, C.c_void_p, u, C.c_void_p, u, C.c_void_p, u]
    for nm in ("vjh_zsplit", "vjh_zmerge"):
        f = getattr(L, nm)
        f.restype = C.c_int64
        f.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                      C.c_void_p]
    return L


def test_host_zsplit_zmerge():
    L = _zlib()
    g = W.rng(31)
    m, n, ld, ldr = 13, 5, 16, 14
    A = np.zeros((n, ld), np.complex128)
    A[:, :m] = g.standard_normal((n, m)) + 1j * g.standard_normal((n, m))
    Ar = np.full((n, ldr), 7.0)
    Ai = np.full((n, ldr), 7.0)
    U = lambda v: C.byref(C.c_uint64(v))
    assert L.zsplit_(U(m), U(n), A.ctypes.data, U(ld), Ar.ctypes.data, U(ldr), Ai.ctypes.data, U(ldr)) == 0
    assert np.array_equal(Ar[:, :m], A[:, :m].real) and np.array_equal(Ai[:, :m], A[:, :m].imag)
    assert np.all(Ar[:, m:] == 7.0)
    B = np.zeros_like(A)
    assert L.zmerge_(U(m), U(n), Ar.ctypes.data, U(ldr), Ai.ctypes.data, U(ldr), B.ctypes.data, U(ld)) == 0
    assert np.array_equal(B[:, :m], A[:, :m]) and np.all(B[:, m:] == 0)
    assert L.zsplit_(U(m), U(n), A.ctypes.data, U(m - 1), Ar.ctypes.data, U(ldr), Ai.ctypes.data, U(ldr)) == -4


@pytest.mark.gpu
def test_device_zsplit_zmerge():
    import torch
    L = _zlib()
    g = W.rng(32)
    m, n, ld = 1000, 37, 1003
    A = g.standard_normal((n, ld)) + 1j * g.standard_normal((n, ld))
    Ad = torch.from_numpy(A.view(np.float64).copy()).cuda()
    Ar = torch.zeros((n, m), dtype=torch.float64, device="cuda")
    Ai = torch.zeros((n, m), dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    assert L.vjh_zsplit(m, n, Ad.data_ptr(), ld, Ar.data_ptr(), m, Ai.data_ptr(), m, s) == 0
    torch.cuda.synchronize()
    assert np.array_equal(Ar.cpu().numpy(), A[:, :m].real) and np.array_equal(Ai.cpu().numpy(), A[:, :m].imag)
    Bd = torch.zeros_like(Ad)
    assert L.vjh_zmerge(m, n, Ar.data_ptr(), m, Ai.data_ptr(), m, Bd.data_ptr(), ld, s) == 0
    torch.cuda.synchronize()
    B = Bd.cpu().numpy().view(np.complex128)
    assert np.array_equal(B[:, :m], A[:, :m]) and np.all(B[:, m:] == 0)


def _sigma_file(path, n):
    raw = np.fromfile(path, dtype=np.longdouble)
    assert raw.size == n
    return raw.astype(np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("j", [2, 4])
def test_tdvjsvd_driver(tmp_path, j):
    from vecjac_amd import svd as S
    m, n = 96, 32
    A = W.gaussian_matrix(m, n, 41)  # (n, m): columns as rows
    base = tmp_path / "gauss"
    A.tofile(str(base) + ".G")
    r = _run("vj_tdvjsvd", j, m, n, base)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    f = r.stdout.strip().split(",")
    assert f[0] == f'"{base}"' and int(f[1]) == j and int(f[2]) == m and int(f[3]) == n
    sweeps = int(f[5])
    sig = _sigma_file(f"{base}.S{j}", n)
    U = np.fromfile(f"{base}.U{j}", dtype=np.float64).reshape(n, m)
    V = np.fromfile(f"{base}.V{j}", dtype=np.float64).reshape(n, n)
    sig2, U2, V2, rc = S.dvjsvd_host(A.copy(), m, n, strategy=j)
    assert rc == sweeps
    assert np.array_equal(sig, sig2) and np.array_equal(U, U2) and np.array_equal(V, V2)


@pytest.mark.gpu
def test_tzvjsvd_driver(tmp_path):
    m, n = 64, 16
    g = W.rng(42)
    A = g.standard_normal((n, m)) + 1j * g.standard_normal((n, m))
    base = tmp_path / "cplx"
    A.tofile(str(base) + ".G")
    r = _run("vj_tzvjsvd", 2, m, n, base)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    f = r.stdout.strip().split(",")
    assert int(f[6]) >= 1
    sig = np.sort(_sigma_file(f"{base}.S2", n))[::-1]
    U = np.fromfile(f"{base}.U2", dtype=np.complex128).reshape(n, m).T
    V = np.fromfile(f"{base}.V2", dtype=np.complex128).reshape(n, n).T
    ref = np.linalg.svd(A.T, compute_uv=False)
    assert np.max(np.abs(sig - ref) / ref) <= 8 * n * np.finfo(np.float64).eps
    s = _sigma_file(f"{base}.S2", n)
    assert np.max(np.abs((U * s) @ V.conj().T - A.T)) <= 1e-12 * ref[0]


@pytest.mark.gpu
def test_tdbjac2_driver(tmp_path):
    """Problems built from known eigenpairs: A = U diag(l1, l2) U^T."""
    lg, nb = 12, 2
    n = 1 << lg
    g = W.rng(43)
    base = tmp_path / "r2"
    cols = {k: [] for k in "fghkl"}
    for _ in range(nb):
        l1 = np.ldexp(g.uniform(1, 2, n), g.integers(-20, 20, n)) * np.where(g.random(n) < 0.5, -1, 1)
        l2 = np.ldexp(g.uniform(1, 2, n), g.integers(-20, 20, n)) * np.where(g.random(n) < 0.5, -1, 1)
        phi = g.uniform(-np.pi / 2, np.pi / 2, n)
        c, s = np.cos(phi), np.sin(phi)
        cols["f"].append(l1 * c * c + l2 * s * s)
        cols["g"].append(l2 * c * c + l1 * s * s)
        cols["h"].append((l1 - l2) * c * s)
        cols["k"].append(l1)
        cols["l"].append(l2)
    for k, v in cols.items():
        np.concatenate(v).tofile(f"{base}.{k}")
    r = _run("vj_tdbjac2", base, lg, nb)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = r.stdout.strip().splitlines()
    assert lines[0] == '"B","Ts","ORT","REN","RLN","RLX","RLM"' and len(lines) == nb + 1
    for ln in lines[1:]:
        B, Ts, ORT, REN, RLN, RLX, RLM = ln.split(",")
        assert float(ORT) <= 4 * np.finfo(np.float64).eps
        assert float(REN) <= 8 * np.finfo(np.float64).eps
        # the inputs were rounded to double after the construction: the
        # eigenvalues move by about eps |A| / |lambda|
        assert float(RLN) <= 1e-2


def _traces(tmp_path, A, name):
    import os
    from vecjac_amd import svd as S
    n, m = A.shape
    base = tmp_path / name
    A.tofile(str(base) + ".G")
    r = subprocess.run([str(BIN / "vj_tdvjsvd"), "2", str(m), str(n), str(base)], capture_output=True, text=True,
                       timeout=300, env={**os.environ, "VJ_JTRACE": "1"})
    assert r.returncode == 0, r.stderr
    ours = Path(f"{base}.2").read_text().splitlines()
    js, _ = S.pjs(2, n)
    js.astype(np.uint32).tofile(str(base) + ".js")
    rr = subprocess.run([str(ROOT / "oracle" / "_ref" / "ref_trace"), str(m), str(n), str(base) + ".G",
                         str(base) + ".js", str(base) + ".ref"], capture_output=True, text=True, timeout=300)
    assert rr.returncode == 0, rr.stderr
    ref = Path(f"{base}.ref").read_text().splitlines()
    print("\n".join(["ours:"] + ours + ["reference:"] + ref))
    return ours, ref


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gauss64x32", "dgsex16"])
def test_jtrace_vs_reference(tmp_path, case):
    """The trace file of vj_tdvjsvd with VJ_JTRACE set against the trace of
    the reference's own JTRACE build on the same input and table: the same
    lines (initial max, initial scaling, rescalings with their sweep and
    step, transformations per sweep, final shift and bound), except the
    per-phase times of the last line."""
    if not (ROOT / "oracle" / "_ref" / "ref_trace").exists():
        pytest.skip("reference build absent")
    if case == "dgsex16":
        from tests.golden.make_golden import dgsex
        A = dgsex(16)
    else:
        A = W.gaussian_matrix(64, 32, 44)
    ours, ref = _traces(tmp_path, A, case)
    assert len(ours) == len(ref)
    assert ours[:2] == ref[:2]                       # max |G| and the initial scaling: exact
    num = lambda s: float(s.split("M=")[1])
    for a, b in zip(ours[2:-1], ref[2:-1]):
        if a.startswith("sweep=") or a.startswith("sT="):
            # sweep, step, eM, sR, sN (or sT) exact; the bound M after the
            # shift is a max over updated elements: last-bit differences
            assert a.split("M=")[0] == b.split("M=")[0]
            assert abs(num(a) - num(b)) <= 1e-12 * abs(num(b))
        else:
            # "sweep transformations": the norms here are correctly rounded and
            # the reference's are its own, so a pair sitting on the tolerance
            # can go either way
            (sa, ta), (sb, tb) = map(int, a.split()), map(int, b.split())
            assert sa == sb and abs(ta - tb) <= max(2, tb // 100)
    assert ours[-1].startswith("Tn=") and ref[-1].startswith("Tn=")
