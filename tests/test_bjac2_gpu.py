"""GPU parity of the batched 2x2 Hermitian EVD against the oracle (which is
itself bit-identical to the reference's AVX-512 kernels, see
test_oracle_vs_ref.py).  Bar: bit-exact outputs and L1<L2 bit words."""
import numpy as np
import pytest

from oracle import vjo
from tests.util import bit_mismatch, nonfinite_jac2, special_jac2
from vecjac_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run_dev(hip, kind, n_signed, ins):
    import torch
    dt = torch.float64 if kind in "dz" else torch.float32
    nn = abs(n_signed)
    vl = 8 if kind in "dz" else 16
    dins = [torch.from_numpy(np.ascontiguousarray(a)).to("cuda") for a in ins]
    outs = [torch.empty(nn, dtype=dt, device="cuda") for _ in range(5 if kind in "zc" else 4)]
    p = torch.zeros(nn // vl, dtype=torch.int32, device="cuda")
    f = getattr(hip, f"vjh_{kind}bjac2")
    s = torch.cuda.current_stream().cuda_stream
    rc = f(n_signed, *[t.data_ptr() for t in dins], *[t.data_ptr() for t in outs], p.data_ptr(), s)
    torch.cuda.synchronize()
    return rc, [t.cpu().numpy() for t in outs], p.cpu().numpy().view(np.uint32)


def _run_host(hip, kind, n_signed, ins):
    import ctypes as C
    dt = np.float64 if kind in "dz" else np.float32
    nn = abs(n_signed)
    vl = 8 if kind in "dz" else 16
    ins = [np.ascontiguousarray(a, dtype=dt) for a in ins]
    outs = [np.empty(nn, dtype=dt) for _ in range(5 if kind in "zc" else 4)]
    p = np.zeros(nn // vl, dtype=np.uint32)
    f = getattr(hip, f"{kind}bjac2_")
    rc = f(C.byref(C.c_int64(n_signed)), *[a.ctypes.data for a in ins], *[a.ctypes.data for a in outs], p.ctypes.data)
    return rc, outs, p


@pytest.mark.parametrize("kind", ["d", "z", "s", "c"])
@pytest.mark.parametrize("sines", [False, True])
def test_bjac2_random_bitexact(hip, kind, sines):
    n = 1 << 18
    dt = np.float64 if kind in "dz" else np.float32
    ins = [a for a in W.jac2_batch(n, 1234, kind in "zc", dtype=dt) if a is not None]
    ns = -n if sines else n
    print(f"bjac2 kind={kind} n={ns} seed=1234", flush=True)
    rc, outs, p = _run_dev(hip, kind, ns, ins)
    rco, oo, po = vjo.bjac2(kind, ns, ins)
    assert rc == 0 and rco == 0
    for k, (a, b) in enumerate(zip(outs, oo)):
        assert bit_mismatch(a, b) == 0, f"output {k}"
    assert np.array_equal(p, po)


@pytest.mark.parametrize("kind", ["d", "z", "s", "c"])
def test_bjac2_special_values(hip, kind):
    dt = np.float64 if kind in "dz" else np.float32
    ins = special_jac2(kind in "zc", dt)
    n = ins[0].size
    for ns in (n, -n):
        print(f"bjac2 special kind={kind} n={ns}", flush=True)
        rc, outs, p = _run_dev(hip, kind, ns, ins)
        _, oo, po = vjo.bjac2(kind, ns, ins)
        for k, (a, b) in enumerate(zip(outs, oo)):
            assert bit_mismatch(a, b) == 0, f"output {k}"
        assert np.array_equal(p, po)


@pytest.mark.parametrize("kind", ["d", "z", "s", "c"])
def test_bjac2_nonfinite(hip, kind):
    """Lanes with an infinite or NaN input take the general scaling path,
    the others the integer-exponent one, within the same waves.

    Bit-exact except for one documented deviation: with an infinite input
    the reference forms a NaN whose sign bit (set for x86's invalid-operation
    NaN) reaches the tangent through sgn(a11 - a22); gfx950's NaNs do not
    carry it, so a zero tangent (at, or cat/sat) may differ in its sign.
    Only rows with a non-finite input, only those outputs and only zeros
    may differ."""
    dt = np.float64 if kind in "dz" else np.float32
    ins = nonfinite_jac2(kind in "zc", dt)
    n = ins[0].size
    finite_row = np.all([np.isfinite(a) for a in ins], axis=0)
    rot = (1,) if kind in "ds" else (1, 2)      # at; cat, sat
    ui = np.uint64 if dt == np.float64 else np.uint32
    for ns in (n, -n):
        print(f"bjac2 nonfinite kind={kind} n={ns}", flush=True)
        rc, outs, p = _run_dev(hip, kind, ns, ins)
        _, oo, po = vjo.bjac2(kind, ns, ins)
        assert rc == 0
        for k, (a, b) in enumerate(zip(outs, oo)):
            diff = (a.view(ui) != b.view(ui)) & ~(np.isnan(a) & np.isnan(b))
            if k not in rot:
                assert not diff.any(), f"output {k}"
                continue
            assert not (diff & finite_row).any(), f"output {k} on finite rows"
            assert np.all((a[diff] == 0) & (b[diff] == 0)), f"output {k}: a difference other than the sign of a zero"
        assert np.array_equal(p, po)


@pytest.mark.parametrize("kind", ["d", "z"])
def test_bjac2_host_entry(hip, kind):
    n = 4096
    ins = [a for a in W.jac2_batch(n, 5, kind == "z") if a is not None]
    rc, outs, p = _run_host(hip, kind, -n, ins)
    _, oo, po = vjo.bjac2(kind, -n, ins)
    assert rc == 1  # the reference's release (OpenMP) build returns 1 (dbjac2.c:113)
    for a, b in zip(outs, oo):
        assert bit_mismatch(a, b) == 0
    assert np.array_equal(p, po)


def test_bjac2_bad_length(hip):
    # |n| not a multiple of the lane count: the reference's argument check
    rc, _, _ = _run_host(hip, "d", 8, [np.zeros(8)] * 3)
    assert rc == 1
    import ctypes as C
    a = np.zeros(12)
    rc = hip.dbjac2_(C.byref(C.c_int64(12)), *([a.ctypes.data] * 7), np.zeros(2, np.uint32).ctypes.data)
    assert rc == -1


@pytest.mark.parametrize("kind", ["d", "z", "s", "c"])
@pytest.mark.parametrize("secants", [False, True])
def test_bjac2_internal_forms(hip, kind, secants):
    """dbjac2i/zbjac2i/sbjac2i/cbjac2i (dbjac2.c:178-251): n<0 secants, n>=0
    tangents, a block of VL problems solved only when the low VL bits of its
    p word are set; those bits become the block's l1 < l2 bits, the high bits
    of p stay, and a skipped block keeps its outputs.  Device and host entry
    points against the oracle's internal forms (bit-identical to the
    reference build), with outputs pre-filled with a sentinel."""
    import ctypes as C
    import torch
    dt = np.float64 if kind in "dz" else np.float32
    vl = 8 if kind in "dz" else 16
    n = 1 << 14
    ins = [a for a in W.jac2_batch(n, 4321, kind in "zc", dtype=dt) if a is not None]
    g = W.rng(11)
    lo = (1 << vl) - 1
    p0 = (g.integers(0, 1 << 16, n // vl).astype(np.uint32) << np.uint32(16)) | \
        np.where(g.integers(0, 3, n // vl) == 0, 0, g.integers(1, lo + 1, n // vl)).astype(np.uint32)
    ns = -n if secants else n
    nout = 5 if kind in "zc" else 4
    # oracle: outputs start as the sentinel too
    sent = dt(-7.25)
    rco, oo, po = vjo.bjac2(kind, ns, ins, internal=True, p_init=p0)
    skipped = np.repeat((p0 & lo) == 0, vl)
    for a in oo:
        a[skipped] = sent
    # device
    dins = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in ins]
    douts = [torch.full((n,), float(sent), dtype=torch.float64 if dt == np.float64 else torch.float32, device="cuda")
             for _ in range(nout)]
    dp = torch.from_numpy(p0.view(np.int32).copy()).cuda()
    print(f"bjac2i kind={kind} n={ns}", flush=True)
    rc = getattr(hip, f"vjh_{kind}bjac2i")(ns, *[t.data_ptr() for t in dins], *[t.data_ptr() for t in douts],
                                            dp.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == 0
    for k, (a, b) in enumerate(zip(douts, oo)):
        assert bit_mismatch(a.cpu().numpy(), b) == 0, f"device output {k}"
    assert np.array_equal(dp.cpu().numpy().view(np.uint32), po)
    # host
    outs = [np.full(n, sent, dtype=dt) for _ in range(nout)]
    ph = p0.copy()
    hins = [np.ascontiguousarray(a, dtype=dt) for a in ins]
    rc = getattr(hip, f"{kind}bjac2i")(C.byref(C.c_int64(ns)), *[a.ctypes.data for a in hins],
                                        *[a.ctypes.data for a in outs], ph.ctypes.data)
    assert rc == 1
    for k, (a, b) in enumerate(zip(outs, oo)):
        assert bit_mismatch(a, b) == 0, f"host output {k}"
    assert np.array_equal(ph, po)
