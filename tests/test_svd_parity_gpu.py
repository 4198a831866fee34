"""GPU parity of the plain SVD driver against the reference's own outputs and
arithmetic, on the inputs the rest of the suite leaves out:

  * the golden fixtures of the reference build (tests/golden/dvjsvd.npz:
    the reference's dvjsvd_ on stored inputs and step tables);
  * the graded construction A = Q1 diag(sigma) Q2^H with sigma geometric from
    1 down to 1e-12 (the 8192 x 2048 complex workload, at 512 x 128 and
    1024 x 256), and the column-graded A = B D, whose small singular values
    are determined to high relative accuracy;
  * full-width square matrices (512^2 and the 1024^2 workload itself);
  * a run resumed from a preset V (iwork[0] = 1, dvjsvd.c:86-96, 143-157);
  * the reference's USE_SECANTS rotation form (vjh_set_secants), against the
    oracle's secant form, which is pinned to the reference built with
    -DUSE_SECANTS.

Every comparison runs the oracle in both norm modes on the same input and
step table: mode 0 is the reference's arithmetic (bit-identical to the
reference build), mode 1 the GPU's correctly rounded norm and dot.

The bar on sigma is a relative error of 8 n eps wherever the input
determines sigma to that accuracy.  For A = Q1 diag(sigma) Q2^H it does not:
one-sided Jacobi's backward error is columnwise (||dA_j|| <= O(n eps)
||A_j||), the columns of that A all have norm ~ ||A||_F / sqrt(n), and a
singular value sigma_i moves by up to ||dA||_2 under such a perturbation,
i.e. its relative condition is sigma_1 / sigma_i (up to 1e12).  Two correct
runs that differ only in the last bit of a column norm (the oracle in mode 0
against the oracle in mode 1) already disagree on sigma = 1e-12 by about
1e-6 relative (5e7 x 8 n eps at n = 32).  There every sigma is held to
8 n eps against the oracle of the same arithmetic (mode 1), and against
mode 0 to 8 n eps where sigma_i >= 1e-4 sigma_1 and to 8 n eps sigma_i +
eps sigma_1 (one unit of ||A||_2 added) everywhere.
"""
import numpy as np
import pytest

from oracle import vjo
from vecjac_amd import svd as S
from vecjac_amd import workloads as W

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden" / "dvjsvd.npz"


def _gpu(kind, Ar, Ai, js, swp=999, V0=None, V0i=None):
    """Device driver on columns-as-rows arrays; returns numpy outputs."""
    import torch
    n, m = Ar.shape
    G = torch.from_numpy(np.ascontiguousarray(Ar)).cuda()
    Gi = torch.from_numpy(np.ascontiguousarray(Ai)).cuda() if Ai is not None else None
    if V0 is None:
        V = torch.empty((n, n), dtype=torch.float64, device="cuda")
        Vi = torch.empty((n, n), dtype=torch.float64, device="cuda") if Ai is not None else None
    else:
        V = torch.from_numpy(np.ascontiguousarray(V0)).cuda()
        Vi = torch.from_numpy(np.ascontiguousarray(V0i)).cuda() if Ai is not None else None
    eS = torch.empty(n, dtype=torch.float64, device="cuda")
    fS = torch.empty(n, dtype=torch.float64, device="cuda")
    jsd = torch.from_numpy(np.ascontiguousarray(js).reshape(-1).view(np.int32)).cuda()
    print(f"gpu {kind}vjsvd m={m} n={n} swp={swp} vpreset={V0 is not None}", flush=True)
    rc, st = S.device_vjsvd(kind, G, V, eS, fS, jsd, js.shape[0], swp, Gi=Gi, Vi=Vi, vpreset=V0 is not None)
    torch.cuda.synchronize()
    out = dict(rc=rc, hist=st[2], G=G.cpu().numpy(), V=V.cpu().numpy(), eS=eS.cpu().numpy(), fS=fS.cpu().numpy())
    if Ai is not None:
        out["Gi"], out["Vi"] = Gi.cpu().numpy(), Vi.cpu().numpy()
    return out


def _oracle(kind, Ar, Ai, js, mode, swp=999, V0=None, V0i=None):
    n, m = Ar.shape
    G = vjo.aligned((n, m), np.float64)
    G[...] = Ar
    Gi = None
    if Ai is not None:
        Gi = vjo.aligned((n, m), np.float64)
        Gi[...] = Ai
    V = Vi = None
    if V0 is not None:
        V = vjo.aligned((n, n), np.float64)
        V[...] = V0
        if Ai is not None:
            Vi = vjo.aligned((n, n), np.float64)
            Vi[...] = V0i
    with vjo.norm_mode(mode):
        rc, eS, fS, V, Vi, st = vjo.vjsvd_core(kind, G, V=V, Vi=Vi, js=np.ascontiguousarray(js), stp=js.shape[0],
                                               swp=swp, Gi=Gi, iwork0=int(V0 is not None))
    return dict(rc=rc, hist=list(st.swt[:64]), G=G, V=V, eS=eS, fS=fS, sT=st.sT)


def _split(A):
    """m x n (complex) matrix -> columns-as-rows real and imaginary planes."""
    Ar = np.ascontiguousarray(A.real.T)
    Ai = np.ascontiguousarray(A.imag.T) if np.iscomplexobj(A) else None
    return Ar, Ai


def _ld_checks(A, r, lim):
    """U, V orthonormal and ||A - U S V^H||_F / ||A||_F, in long double."""
    ld = np.longdouble

    def mul(ar, ai, br, bi):
        return ar @ br - ai @ bi, ar @ bi + ai @ br

    n = A.shape[1]
    Ur, Vr = r["G"].astype(ld).T, r["V"].astype(ld).T
    Ui = r["Gi"].astype(ld).T if "Gi" in r else np.zeros_like(Ur)
    Vi = r["Vi"].astype(ld).T if "Vi" in r else np.zeros_like(Vr)
    s = S.sigma(r["eS"], r["fS"]).astype(ld)
    I = np.eye(n, dtype=ld)
    gr, gi = mul(Ur.T, -Ui.T, Ur, Ui)
    hr, hi = mul(Vr.T, -Vi.T, Vr, Vi)
    pr, pi = mul(Ur * s, Ui * s, Vr.T, -Vi.T)
    Ar, Ai = np.real(A).astype(ld), np.imag(A).astype(ld)
    orthU = max(np.max(np.abs(gr - I)), np.max(np.abs(gi)))
    orthV = max(np.max(np.abs(hr - I)), np.max(np.abs(hi)))
    rec = np.sqrt(np.sum((Ar - pr) ** 2 + (Ai - pi) ** 2)) / np.sqrt(np.sum(Ar ** 2 + Ai ** 2))
    print(f"  orthU {orthU:.2e} orthV {orthV:.2e} rec {rec:.2e} (lim {lim:.2e})")
    assert orthU <= lim and orthV <= lim and rec <= lim, (orthU, orthV, rec)


def _sig(r):
    return np.sort(S.sigma(r["eS"], r["fS"]))[::-1]


def _report(tag, n, g, o, sg, so):
    rel = np.abs(sg - so) / so
    print(f"  {tag}: sweeps gpu {g['rc']} oracle {o['rc']}; max rel err {np.max(rel):.3e} "
          f"({np.max(rel) / (8 * n * EPS):.3g} x 8n eps) at sigma {so[np.argmax(rel)]:.3e}; "
          f"bit-identical {int(np.sum(sg == so))}/{n}")
    print(f"    hist gpu    {g['hist'][:g['rc'] + 1]}")
    print(f"    hist oracle {o['hist'][:o['rc'] + 1]}")
    return rel


# ------------------------------ golden vectors ------------------------------

_GOLD = ["g64x32", "g128x48", "dgsex16", "rankdef48x32", "g64x32_swp2"]


@pytest.mark.parametrize("name", _GOLD)
def test_golden_reference_outputs(name):
    """The reference build's own dvjsvd_ outputs on the stored input and
    step table: the sweep count within one, sigma within 8 n eps (the norms
    eS/fS of the working G for the sweep-limited case, which returns before
    the epilogue, dvjsvd.c:604).  The rank-deficient case's 16 singular
    values of rounding size (~eps ||A||) are only bounded: below
    8 n eps sigma_1 in both."""
    z = np.load(GOLDEN)
    A, js, swp = z[f"{name}_A"], z[f"{name}_js"].astype(np.uint32), int(z[f"{name}_swp"])
    n, m = A.shape
    g = _gpu("d", A, None, js, swp=swp)
    rc_ref = int(z[f"{name}_rc"])
    sg = _sig(g)
    sr = np.sort(vjo.sigma(z[f"{name}_eS"], z[f"{name}_fS"]))[::-1]
    big = sr >= 1e-8 * sr[0]
    rel = np.abs(sg[big] - sr[big]) / sr[big]
    print(f"{name}: sweeps gpu {g['rc']} reference {rc_ref}; max rel err {np.max(rel):.3e} "
          f"({np.max(rel) / (8 * n * EPS):.3g} x 8n eps)")
    if swp < 999:
        assert g["rc"] == rc_ref == swp
    else:
        assert g["rc"] >= 0 and abs(g["rc"] - rc_ref) <= 1
    assert np.max(rel) <= 8 * n * EPS
    assert np.all(sg[~big] <= 8 * n * EPS * sr[0]) and np.all(sr[~big] <= 8 * n * EPS * sr[0])
    assert np.sum(sg < 1e-8 * sg[0]) == np.sum(~big)
    if swp == 999:
        _ld_checks(A.T, g, 8 * n * EPS)


# --------------------------- graded constructions ---------------------------

def _graded_geometric_input(m, n, seed=7):
    sig = 10.0 ** (-12.0 * np.arange(n) / (n - 1))
    return W.graded_svd_matrix(m, n, sig, seed, True)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("m,n", [(512, 128), (1024, 256)])
def test_graded_geometric_sigma(m, n):
    """A = Q1 diag(sigma) Q2^H, complex, sigma = 10^(-12 i/(n-1)), the ME
    table; bars as in the module docstring."""
    A = _graded_geometric_input(m, n)
    Ar, Ai = _split(A)
    js, _ = S.pjs(2, n)
    g = _gpu("z", Ar, Ai, js)
    assert g["rc"] > 0
    sg = _sig(g)
    print(f"graded construction (sigma geometric 1 .. 1e-12) {m}x{n}")
    for mode in (1, 0):
        o = _oracle("z", Ar, Ai, js, mode)
        so = _sig(o)
        rel = _report(f"mode {mode}", n, g, o, sg, so)
        assert abs(g["rc"] - o["rc"]) <= 1, (mode, g["rc"], o["rc"])
        if mode == 1:
            assert np.max(rel) <= 8 * n * EPS
        else:
            well = so >= 1e-4 * so[0]
            assert np.max(rel[well]) <= 8 * n * EPS
            assert np.all(np.abs(sg - so) <= 8 * n * EPS * so + EPS * so[0])
    _ld_checks(A, g, 8 * n * EPS)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("kind,m,n", [("z", 1024, 256), ("d", 1024, 256), ("z", 2048, 128)])
def test_column_graded(kind, m, n):
    """A = B D: B Gaussian (kappa(B) ~ 3), D = diag(10^(-12 j/(n-1))), so
    sigma spans 12 decades and every sigma is determined to high relative
    accuracy: all of them, the smallest included, within 8 n eps of the
    oracle in both modes."""
    g_ = W.rng(5 + n)
    B = g_.standard_normal((m, n))
    if kind == "z":
        B = B + 1j * g_.standard_normal((m, n))
    A = B * (10.0 ** (-12.0 * np.arange(n) / (n - 1)))
    Ar, Ai = _split(A)
    js, _ = S.pjs(2, n)
    g = _gpu(kind, Ar, Ai, js)
    assert g["rc"] > 0
    sg = _sig(g)
    print(f"column-graded {kind} {m}x{n}, kappa(B) {np.linalg.cond(B):.2f}")
    for mode in (1, 0):
        o = _oracle(kind, Ar, Ai, js, mode)
        rel = _report(f"mode {mode}", n, g, o, sg, _sig(o))
        assert abs(g["rc"] - o["rc"]) <= 1, (mode, g["rc"], o["rc"])
        assert np.max(rel) <= 8 * n * EPS
    _ld_checks(A, g, 8 * n * EPS)


# ------------------------------- full width --------------------------------

@pytest.mark.timeout(400)
@pytest.mark.parametrize("n", [512, 1024])
def test_square_full_width(n):
    """n x n Gaussian with the ME table: n/2 pairs per step, so the steps
    run more pairs than a workgroup slot holds (n = 1024 is the plain
    benchmark's input itself)."""
    A = W.gaussian_matrix(n, n, 1)
    js, _ = S.pjs(2, n)
    g = _gpu("d", A, None, js)
    assert g["rc"] > 0
    sg = _sig(g)
    print(f"square {n}")
    for mode in (1, 0):
        o = _oracle("d", A, None, js, mode)
        rel = _report(f"mode {mode}", n, g, o, sg, _sig(o))
        assert abs(g["rc"] - o["rc"]) <= 1, (mode, g["rc"], o["rc"])
        assert np.max(rel) <= 8 * n * EPS
    _ld_checks(A.T, g, 8 * n * EPS)


# --------------------------------- V preset ---------------------------------

@pytest.mark.parametrize("kind,m,n", [("d", 256, 64), ("z", 256, 64),
                                    # columns streamed in chunks
                                    ("d", 9000, 8), ("z", 5000, 8)])
def test_vpreset_resume(kind, m, n):
    """A run stopped after two sweeps returns its working G (A V 2^sT) and
    V; the driver resumed from those with a preset V matches the oracle's
    iwork[0] = 1 path on the same G and V (sweeps within one, sigma within
    8 n eps, both modes) and an uninterrupted run (sigma 2^-sT within
    8 n eps), with A = U S V^H; also for columns long enough to be
    streamed."""
    g_ = W.rng(43)
    A = g_.standard_normal((m, n))
    if kind == "z":
        A = A + 1j * g_.standard_normal((m, n))
    Ar, Ai = _split(A)
    js, _ = S.pjs(2, n)
    full = _gpu(kind, Ar, Ai, js)
    part = _gpu(kind, Ar, Ai, js, swp=2)
    assert part["rc"] == 2
    res = _gpu(kind, part["G"], part.get("Gi"), js, V0=part["V"], V0i=part.get("Vi"))
    assert res["rc"] > 0 and abs(2 + res["rc"] - full["rc"]) <= 1, (res["rc"], full["rc"])
    sr = _sig(res)
    for mode in (1, 0):
        o = _oracle(kind, part["G"], part.get("Gi"), js, mode, V0=part["V"], V0i=part.get("Vi"))
        rel = _report(f"resumed, mode {mode}", n, res, o, sr, _sig(o))
        assert abs(res["rc"] - o["rc"]) <= 1
        assert np.max(rel) <= 8 * n * EPS
    sf = _sig(full)
    sT = int(np.rint(np.log2(sr[0] / sf[0])))
    rel = np.abs(np.ldexp(sr, -sT) - sf) / sf
    print(f"  resumed vs uninterrupted: 2^{sT}, max rel err {np.max(rel):.3e}")
    assert np.max(rel) <= 8 * n * EPS
    res["eS"] = res["eS"] - sT
    _ld_checks(A, res, 8 * n * EPS)


@pytest.mark.parametrize("kind", ["d", "z"])
def test_vpreset_zero_refused(kind):
    """A preset V that is zero: -13 (complex -17), before G is touched
    (dvjsvd.c:93-94), from the device entry point and from the reference's
    host entry point (iwork[0] = 1)."""
    import torch
    m, n = 64, 16
    A = W.gaussian_matrix(m, n, 2)
    js, stp = S.pjs(2, n)
    code = -17 if kind == "z" else -13
    r = _gpu(kind, A, A.copy() if kind == "z" else None, js, V0=np.zeros((n, n)), V0i=np.zeros((n, n)))
    assert r["rc"] == code
    assert np.array_equal(r["G"], A)
    G = S.aligned((n, m), np.float64)
    G[...] = A
    V = S.aligned((n, n), np.float64)
    eS, fS = S.aligned(n, np.float64), S.aligned(n, np.float64)
    work = S.aligned(8 * n, np.float64)
    iw = S.aligned(max(1, n // 8), np.uint32)
    iw[0] = 1
    if kind == "d":
        rc = S.vjsvd("d", m, n, G, m, V, n, eS, fS, js, stp, 999, work, iw)
    else:
        Gi = S.aligned((n, m), np.float64)
        Gi[...] = A
        Vi = S.aligned((n, n), np.float64)
        rc = S.vjsvd("z", m, n, G, m, V, n, eS, fS, js, stp, 999, work, iw, Gi=Gi, ldGi=m, Vi=Vi, ldVi=n)
    assert rc == code
    torch.cuda.synchronize()


@pytest.mark.parametrize("kind,m,n", [("d", 300, 96), ("d", 1024, 256), ("z", 400, 64)])
def test_secants_form(kind, m, n):
    """With vjh_set_secants(1) the driver equals the oracle's secant form in
    the GPU's arithmetic bit for bit, is within 8 n eps of the reference's
    arithmetic (mode 0) on a well-conditioned input, and differs from the
    cosine form (the switch acts)."""
    g = W.rng(5150 + n)
    A = g.standard_normal((m, n))
    if kind == "z":
        A = A + 1j * g.standard_normal((m, n))
    Ar, Ai = _split(A)
    js, _ = S.pjs(2, n)
    with S.secants(True):
        r = _gpu(kind, Ar, Ai, js)
    rc0 = _gpu(kind, Ar, Ai, js)
    with vjo.secants(True):
        o1 = _oracle(kind, Ar, Ai, js, 1)
        o0 = _oracle(kind, Ar, Ai, js, 0)
    assert r["rc"] == o1["rc"] and r["rc"] > 0
    for f in ("eS", "fS", "G", "V"):
        assert np.array_equal(r[f], o1[f], equal_nan=True), f
    sg = np.sort(S.sigma(r["eS"], r["fS"]))
    so = np.sort(S.sigma(o0["eS"], o0["fS"]))
    rel = float(np.max(np.abs(sg - so) / so))
    print(f"secants {kind} {m}x{n}: sweeps {r['rc']} (mode 0 {o0['rc']}), max rel err vs mode 0 {rel / (8 * n * EPS):.3f} x 8n eps")
    assert rel <= 8 * n * EPS and abs(r["rc"] - o0["rc"]) <= 1
    assert not np.array_equal(r["V"], rc0["V"])
