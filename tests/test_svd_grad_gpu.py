"""Differentiable batched SVD: GPU Jacobi forward + hand-written HIP adjoint.

Loss: sum_i w_i s_i + <G_U, U> + <G_V, V> with seeded Gaussian w, G_U, G_V, on
A = Q1 diag(s) Q2^T, s ~ U[1, 10] with minimum relative gap 1e-3 (the
differentiable-SVD workload at reduced batch).  Checks:
  * the closed form d(sum w_i s_i)/dA = U diag(w) V^T (relative 1e-12);
  * torch.linalg.svd FP64 autograd on the CPU with the same gauge (sorted
    singular values, largest-|.| entry of each U column positive): relative
    1e-8 of the gradient's max-norm;
  * central finite differences, h = 1e-6, evaluated with a 40-digit mpmath
    SVD (double-precision evaluation of the loss is too noisy for this bar):
    |<dA, E> - FD| <= 1e-8 max(1, |<dA, E>|) along random directions E.
"""
import numpy as np
import pytest

from vecjac_amd import workloads as W

pytestmark = pytest.mark.gpu


def _gauge(U, S, V):
    import torch
    order = torch.argsort(S, dim=-1, descending=True)
    S = torch.gather(S, -1, order)
    U = torch.gather(U, -1, order.unsqueeze(-2).expand_as(U))
    V = torch.gather(V, -1, order.unsqueeze(-2).expand_as(V))
    idx = U.abs().argmax(dim=-2, keepdim=True)
    sgn = torch.sign(torch.gather(U, -2, idx))
    return U * sgn, S, V * sgn


def _loss(U, S, V, w, GU, GV):
    return (w * S).sum() + (GU * U).sum() + (GV * V).sum()


def test_grad_closed_form_sigma_only():
    import torch
    from vecjac_amd.autograd import jacobi_svd
    A, w, GU, GV = W.grad_batch(8, 64, 32, 11)
    At = torch.from_numpy(A).cuda().requires_grad_(True)
    U, S, V = jacobi_svd(At)
    wt = torch.from_numpy(w).cuda()
    (wt * S).sum().backward()
    ref = (U * wt.unsqueeze(1)) @ V.transpose(1, 2)
    err = (At.grad - ref).abs().max() / ref.abs().max()
    print(f"closed-form rel err {float(err):.3e}")
    assert float(err) <= 1e-12


def test_grad_vs_torch_cpu():
    import torch
    from vecjac_amd.autograd import jacobi_svd
    A, w, GU, GV = W.grad_batch(16, 64, 32, 12)
    At = torch.from_numpy(A).cuda().requires_grad_(True)
    U, S, V = jacobi_svd(At)
    wt, GUt, GVt = (torch.from_numpy(x).cuda() for x in (w, GU, GV))
    _loss(U, S, V, wt, GUt, GVt).backward()
    Ac = torch.from_numpy(A).requires_grad_(True)
    Uc, Sc, Vh = torch.linalg.svd(Ac, full_matrices=False)
    Uc, Sc, Vc = _gauge(Uc, Sc, Vh.transpose(-2, -1))
    _loss(Uc, Sc, Vc, torch.from_numpy(w), torch.from_numpy(GU), torch.from_numpy(GV)).backward()
    g_gpu, g_cpu = At.grad.cpu(), Ac.grad
    err = float((g_gpu - g_cpu).abs().max() / g_cpu.abs().max())
    print(f"vs torch.linalg.svd autograd: rel err {err:.3e}")
    assert float((Uc - U.cpu()).abs().max()) <= 1e-10
    assert err <= 1e-8


def test_grad_vs_mpmath_finite_differences():
    mpmath = pytest.importorskip("mpmath")
    import torch
    from vecjac_amd.autograd import jacobi_svd
    mpmath.mp.dps = 40
    A, w, GU, GV = W.grad_batch(1, 24, 12, 13)
    At = torch.from_numpy(A).cuda().requires_grad_(True)
    U, S, V = jacobi_svd(At)
    _loss(U, S, V, *(torch.from_numpy(x).cuda() for x in (w, GU, GV))).backward()
    dA = At.grad.cpu().numpy()[0]

    def loss_mp(M):
        Um, Sm, Vm = mpmath.svd_r(mpmath.matrix(M), full_matrices=False)
        m, n = Um.rows, Um.cols
        s = [Sm[i] for i in range(n)]
        order = sorted(range(n), key=lambda i: -s[i])
        tot = mpmath.mpf(0)
        for jj, j in enumerate(order):
            col = [Um[r, j] for r in range(m)]
            k = max(range(m), key=lambda r: abs(col[r]))
            sg = 1 if col[k] > 0 else -1
            tot += w[0][jj] * s[j]
            tot += sum(GU[0][r][jj] * sg * col[r] for r in range(m))
            tot += sum(GV[0][r][jj] * sg * Vm[j, r] for r in range(n))  # Vm holds V^T
        return tot

    g = W.rng(14)
    h = mpmath.mpf("1e-6")
    for _ in range(2):
        E = g.standard_normal(A[0].shape)
        Am = mpmath.matrix(A[0].tolist())
        Em = mpmath.matrix(E.tolist())
        fd = (loss_mp(Am + h * Em) - loss_mp(Am - h * Em)) / (2 * h)
        an = float(np.sum(dA * E))
        err = abs(an - float(fd))
        print(f"directional derivative {an:.12e} FD {float(fd):.12e} err {err:.2e}")
        assert err <= 1e-8 * max(1.0, abs(an))


def test_failed_member_raises():
    """A member with an infinite entry fails in the forward driver (-80,
    dvjsvd.c:78-82): jacobi_svd raises and names it, instead of returning
    unwritten singular values; JacobiSVD.apply returns the per-member codes
    so that a caller can mask the member instead."""
    import torch
    from vecjac_amd.autograd import JacobiSVD, jacobi_svd
    A, _, _, _ = W.grad_batch(6, 32, 16, 13)
    A[4, 3, 5] = np.inf
    At = torch.from_numpy(A).cuda()
    print("jacobi_svd with an infinite entry in member 4", flush=True)
    with pytest.raises(RuntimeError, match=r"members \[4\]"):
        jacobi_svd(At)
    _, _, _, info = JacobiSVD.apply(At)
    info = info.cpu().numpy()
    assert info[4] == -80 and np.all(np.delete(info, 4) >= 0)


@pytest.mark.parametrize("dt,m,n,nb", [("d", 64, 32, 5), ("d", 50, 22, 5), ("d", 33, 17, 5), ("d", 600, 64, 5),
                                       ("d", 96, 96, 5), ("d", 7, 2, 5), ("s", 64, 32, 5), ("s", 130, 40, 5),
                                       # staged form, more matrices than workgroups (each workgroup
                                       # computes several, the next one's operands in flight)
                                       ("d", 64, 32, 1500), ("d", 45, 45, 1100), ("s", 33, 17, 1200),
                                       # wider than the LDS form: J/K in a global workspace
                                       ("d", 300, 128, 5), ("d", 112, 112, 5), ("s", 256, 128, 5)])
def test_backward_kernel_vs_closed_form(dt, m, n, nb):
    """vjh_dsvd_bwd / vjh_ssvd_bwd on their own, against the adjoint formula
    evaluated in numpy (float64) on random orthonormal U, V, distinct s and
    random cotangents: shapes that are not multiples of the 16 x 16 tiles,
    long columns, square cases, small matrices in the staged form (m n <=
    2048, n n <= 2048) over more matrices than workgroups, and widths past
    the LDS form (n > 96 FP64).  Relative error 1e-12 (FP64), 1e-4 (FP32) of
    the result's max-norm."""
    import torch
    from vecjac_amd import _lib
    from vecjac_amd import autograd  # noqa: F401  (declares the prototypes)
    g = W.rng(1000 * m + n)
    Us, Ss, Vs, dUs, dSs, dVs, ref = [], [], [], [], [], [], []
    for _ in range(nb):
        U = W.haar_unitary(m, n, g, False)
        V = W.haar_unitary(n, n, g, False)
        s = W.gapped_sigma(n, g)
        dU, dV, ds = g.standard_normal((m, n)), g.standard_normal((n, n)), g.standard_normal(n)
        J, K = U.T @ dU, V.T @ dV
        with np.errstate(divide="ignore"):
            F = 1.0 / (s[None, :] ** 2 - s[:, None] ** 2)
        np.fill_diagonal(F, 0.0)
        inner = np.diag(ds) + (F * (J - J.T)) * s[None, :] + s[:, None] * (F * (K - K.T))
        dA = U @ inner @ V.T + (np.eye(m) - U @ U.T) @ dU @ np.diag(1 / s) @ V.T
        for lst, x in zip((Us, Ss, Vs, dUs, dSs, dVs, ref), (U, s, V, dU, ds, dV, dA)):
            lst.append(x)
    tdt = torch.float64 if dt == "d" else torch.float32
    cm = lambda xs: torch.from_numpy(np.stack([x.T for x in xs]).copy()).to(tdt).cuda()  # noqa: E731
    vec = lambda xs: torch.from_numpy(np.stack(xs)).to(tdt).cuda()  # noqa: E731
    Ut, Vt, dUt, dVt = cm(Us), cm(Vs), cm(dUs), cm(dVs)
    St, dSt = vec(Ss), vec(dSs)
    out = torch.full((nb, n, m), float("nan"), dtype=tdt, device="cuda")
    L = _lib.lib()
    f = L.vjh_dsvd_bwd if dt == "d" else L.vjh_ssvd_bwd
    print(f"{dt}svd_bwd nb={nb} m={m} n={n}", flush=True)
    rc = f(nb, m, n, Ut.data_ptr(), St.data_ptr(), Vt.data_ptr(), dUt.data_ptr(), dSt.data_ptr(), dVt.data_ptr(),
           out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == 0, _lib.last_error()
    got = out.double().cpu().numpy().transpose(0, 2, 1)
    want = np.stack(ref)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"  max rel err {err:.3e}")
    assert err <= (1e-12 if dt == "d" else 1e-4)


def _complex_formula(U, s, V, dU, ds, dV):
    """The complex adjoint in numpy (complex128): J = U^H dU, K = V^H dV,
    dA = U [diag(ds) + (F o (J - J^H)) S + S (F o (K - K^H))
            + i diag(Im J_kk - Im K_kk) / (2 s)] V^H + (I - U U^H) dU S^-1 V^H."""
    m = U.shape[0]
    J, K = U.conj().T @ dU, V.conj().T @ dV
    with np.errstate(divide="ignore"):
        F = 1.0 / (s[None, :] ** 2 - s[:, None] ** 2)
    np.fill_diagonal(F, 0.0)
    inner = np.diag(ds).astype(complex) + (F * (J - J.conj().T)) * s[None, :] + s[:, None] * (F * (K - K.conj().T))
    inner += np.diag(1j * (np.diag(J).imag - np.diag(K).imag) / (2 * s))
    return U @ inner @ V.conj().T + (np.eye(m) - U @ U.conj().T) @ dU @ np.diag(1 / s) @ V.conj().T


@pytest.mark.parametrize("dt,m,n", [("z", 64, 32), ("z", 50, 22), ("z", 96, 64), ("z", 7, 2), ("c", 64, 32),
                                    # wider than the LDS form (complex n > 64)
                                    ("z", 200, 80), ("z", 160, 128)])
def test_complex_backward_kernel_vs_formula(dt, m, n):
    """vjh_zsvd_bwd / vjh_csvd_bwd on split planes against the complex
    adjoint evaluated in numpy on random unitary U, V, distinct s and random
    complex cotangents.  Relative error 1e-12 (FP64), 1e-4 (FP32)."""
    import torch
    from vecjac_amd import _lib
    from vecjac_amd import autograd  # noqa: F401  (declares the prototypes)
    nb = 4
    g = W.rng(2000 * m + n)
    Us, Ss, Vs, dUs, dSs, dVs, ref = [], [], [], [], [], [], []
    for _ in range(nb):
        U = W.haar_unitary(m, n, g, True)
        V = W.haar_unitary(n, n, g, True)
        s = W.gapped_sigma(n, g)
        dU = g.standard_normal((m, n)) + 1j * g.standard_normal((m, n))
        dV = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
        ds = g.standard_normal(n)
        for lst, x in zip((Us, Ss, Vs, dUs, dSs, dVs, ref), (U, s, V, dU, ds, dV, _complex_formula(U, s, V, dU, ds, dV))):
            lst.append(x)
    tdt = torch.float64 if dt == "z" else torch.float32
    cm = lambda xs, f: torch.from_numpy(np.stack([f(x).T for x in xs]).copy()).to(tdt).cuda()  # noqa: E731
    vec = lambda xs: torch.from_numpy(np.stack(xs)).to(tdt).cuda()  # noqa: E731
    planes = [cm(xs, f) for xs in (Us, Vs, dUs, dVs) for f in (np.real, np.imag)]
    Ur, Ui, Vr, Vi, dUr, dUi, dVr, dVi = planes
    St, dSt = vec(Ss), vec(dSs)
    outr = torch.full((nb, n, m), float("nan"), dtype=tdt, device="cuda")
    outi = torch.full_like(outr, float("nan"))
    L = _lib.lib()
    f = L.vjh_zsvd_bwd if dt == "z" else L.vjh_csvd_bwd
    print(f"{dt}svd_bwd nb={nb} m={m} n={n}", flush=True)
    rc = f(nb, m, n, Ur.data_ptr(), Ui.data_ptr(), St.data_ptr(), Vr.data_ptr(), Vi.data_ptr(), dUr.data_ptr(),
           dUi.data_ptr(), dSt.data_ptr(), dVr.data_ptr(), dVi.data_ptr(), outr.data_ptr(), outi.data_ptr(),
           torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == 0, _lib.last_error()
    got = (outr.double().cpu().numpy() + 1j * outi.double().cpu().numpy()).transpose(0, 2, 1)
    want = np.stack(ref)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"  max rel err {err:.3e}")
    assert err <= (1e-12 if dt == "z" else 1e-4)


def _complex_case(nb, m, n, seed):
    """A = Q1 diag(s) Q2^H (complex Haar factors, gapped s in [1, 10]) and
    the weights of a loss that does not depend on the phases of the
    singular vector pairs."""
    g = W.rng(seed)
    A = np.empty((nb, m, n), dtype=complex)
    for b in range(nb):
        A[b] = (W.haar_unitary(m, n, g, True) * W.gapped_sigma(n, g)) @ W.haar_unitary(n, n, g, True).conj().T
    w = g.standard_normal((nb, n))
    W1 = g.standard_normal((nb, m, n))
    W2 = g.standard_normal((nb, n, n))
    P = g.standard_normal((nb, m, n)) + 1j * g.standard_normal((nb, m, n))
    Q = g.standard_normal((nb, n, n)) + 1j * g.standard_normal((nb, n, n))
    return A, (w, W1, W2, P, Q)


def _complex_loss(U, S, V, wts):
    """sum w s + sum W1 |U|^2 + sum W2 |V|^2 + Re sum_k (P_k^H u_k)(v_k^H Q_k):
    unchanged by (u_k, v_k) -> (u_k e^{i phi}, v_k e^{i phi})."""
    w, W1, W2, P, Q = wts
    a = (P.conj() * U).sum(dim=-2)
    b = (V.conj() * Q).sum(dim=-2)
    return (w * S).sum() + (W1 * U.abs() ** 2).sum() + (W2 * V.abs() ** 2).sum() + (a * b).real.sum()


def test_complex_grad_vs_torch_cpu():
    """jacobi_svd on complex128 input: the gradient of a phase-independent
    loss against torch.linalg.svd's complex autograd on the CPU, relative
    1e-8 of the gradient's max-norm."""
    import torch
    from vecjac_amd.autograd import jacobi_svd
    A, wts = _complex_case(8, 48, 24, 21)
    At = torch.from_numpy(A).cuda().requires_grad_(True)
    print("complex jacobi_svd forward/backward", flush=True)
    U, S, V = jacobi_svd(At)
    _complex_loss(U, S, V, [torch.from_numpy(x).cuda() for x in wts]).backward()
    Ac = torch.from_numpy(A).requires_grad_(True)
    Uc, Sc, Vh = torch.linalg.svd(Ac, full_matrices=False)
    _complex_loss(Uc, Sc, Vh.mH, [torch.from_numpy(x) for x in wts]).backward()
    g_gpu, g_cpu = At.grad.cpu(), Ac.grad
    err = float((g_gpu - g_cpu).abs().max() / g_cpu.abs().max())
    print(f"vs torch.linalg.svd complex autograd: rel err {err:.3e}")
    rec = float((U @ torch.diag_embed(S.to(U.dtype)) @ V.mH - At.detach()).abs().max())
    assert rec <= 1e-12 * 10 and float((S.cpu() - Sc.detach()).abs().max()) <= 1e-12 * 10
    assert err <= 1e-8


def test_complex_grad_vs_mpmath_finite_differences():
    """Central differences (h = 1e-6) of the same phase-independent loss
    evaluated with a 40-digit mpmath complex SVD, along random complex
    directions E: |Re <dA, E> - FD| <= 1e-8 max(1, |Re <dA, E>|)."""
    mpmath = pytest.importorskip("mpmath")
    import torch
    from vecjac_amd.autograd import jacobi_svd
    mpmath.mp.dps = 40
    A, wts = _complex_case(1, 12, 6, 23)
    At = torch.from_numpy(A).cuda().requires_grad_(True)
    U, S, V = jacobi_svd(At)
    _complex_loss(U, S, V, [torch.from_numpy(x).cuda() for x in wts]).backward()
    dA = At.grad.cpu().numpy()[0]
    w, W1, W2, P, Q = (x[0] for x in wts)

    def loss_mp(M):
        Um, Sm, Vh = mpmath.svd_c(mpmath.matrix(M), full_matrices=False)
        m, n = Um.rows, Um.cols
        s = [Sm[i] for i in range(n)]
        order = sorted(range(n), key=lambda i: -s[i])
        tot = mpmath.mpf(0)
        for jj, j in enumerate(order):
            u = [Um[r, j] for r in range(m)]
            v = [mpmath.conj(Vh[j, r]) for r in range(n)]
            tot += w[jj] * s[j]
            tot += sum(W1[r][jj] * abs(u[r]) ** 2 for r in range(m))
            tot += sum(W2[r][jj] * abs(v[r]) ** 2 for r in range(n))
            a = sum(mpmath.conj(complex(P[r][jj])) * u[r] for r in range(m))
            b = sum(mpmath.conj(v[r]) * complex(Q[r][jj]) for r in range(n))
            tot += mpmath.re(a * b)
        return tot

    g = W.rng(24)
    h = mpmath.mpf("1e-6")
    for _ in range(2):
        E = g.standard_normal(A[0].shape) + 1j * g.standard_normal(A[0].shape)
        Am = mpmath.matrix(A[0].tolist())
        Em = mpmath.matrix(E.tolist())
        fd = (loss_mp(Am + h * Em) - loss_mp(Am - h * Em)) / (2 * h)
        an = float(np.sum(dA.real * E.real + dA.imag * E.imag))
        err = abs(an - float(fd))
        print(f"directional derivative {an:.12e} FD {float(fd):.12e} err {err:.2e}")
        assert err <= 1e-8 * max(1.0, abs(an))
