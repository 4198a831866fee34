"""Differentiable batched SVD: the GPU Jacobi SVD forward and the
hand-written HIP adjoint backward, wrapped in a torch.autograd.Function.

    U, S, V = jacobi_svd(A)      # A: (batch, m, n) CUDA tensor, m >= n, n even;
                                 # float64, float32, complex128 or complex64

A member whose SVD fails (a non-finite entry: the driver's -80/-92 codes,
dvjsvd.c:78-82) or does not converge within the plan's sweep limit makes
jacobi_svd raise a RuntimeError that names the members and their codes;
JacobiSVD.apply(A) returns the per-member codes as a fourth output instead,
for callers that mask failed members themselves.

The outputs are the thin SVD A = U diag(S) V^H with S sorted in decreasing
order and each singular-vector pair normalised so that the entry of largest
magnitude of every U column is real and positive (a fixed gauge, so that the
outputs are deterministic).  The backward applies the SVD adjoint formula
with F_ij = 1/(s_j^2 - s_i^2) (vjh_dsvd_bwd; complex data: vjh_zsvd_bwd),
which requires distinct singular values.  For complex data the loss must
not depend on the phases of the singular vector pairs (u_k, v_k) ->
(u_k e^{i phi}, v_k e^{i phi}) (e.g. functions of S, of |U_ij|, of the
products u_k v_k^H), as for torch.linalg.svd; the real gauge is a sign, and
any loss of U and V is differentiable.
"""
from __future__ import annotations

import torch

from . import _lib
from . import batched as BT

_PLANS: dict = {}

# A new plan times 1, 2, 4 and 8 waves per workgroup on its first input and
# keeps the fastest (BatchedSVD.autotune).  The best count varies with the
# shape and not monotonically: for 4096 FP64 matrices of m = 2n it is 1 at
# n = 16 and 24, 4 at n = 32 and 2 at n = 48.  The results do not depend on it.
AUTOTUNE = True


SWEEP_LIMIT = 999


def _plan(B: int, m: int, n: int, kind: str):
    """The cached plan of this shape and whether it was just created."""
    key = (B, m, n, kind, torch.cuda.current_device())
    p = _PLANS.get(key)
    if p is None:
        p = BT.BatchedSVD([m] * B, [n] * B, kind=kind, swp=SWEEP_LIMIT)
        _PLANS[key] = p
        return p, True
    return p, False


def _forward(A: torch.Tensor):
    if A.dim() != 3 or not A.is_cuda:
        raise ValueError("A must be a (batch, m, n) CUDA tensor")
    B, m, n = A.shape
    if m < n or n % 2:
        raise ValueError("need m >= n and n even")
    kind = {torch.float64: "d", torch.float32: "s", torch.complex128: "z", torch.complex64: "c"}.get(A.dtype)
    if kind is None:
        raise TypeError("A must be float64, float32, complex128 or complex64")
    cplx = kind in "zc"
    rdt = A.real.dtype
    Ad = A.detach()
    cols = lambda t: t.transpose(1, 2).contiguous().reshape(-1)  # noqa: E731  (columns as rows, packed)
    G = cols(Ad.real if cplx else Ad)
    Gi = cols(Ad.imag) if cplx else None
    V = torch.empty(B * n * n, dtype=rdt, device=A.device)
    Vi = torch.empty(B * n * n, dtype=rdt, device=A.device) if cplx else None
    eS = torch.empty(B * n, dtype=rdt, device=A.device)
    fS = torch.empty(B * n, dtype=rdt, device=A.device)
    info = torch.empty(B, dtype=torch.int64, device=A.device)
    p, new = _plan(B, m, n, kind)
    if new and AUTOTUNE:
        p.autotune(G, V, eS, fS, info, Gi=Gi, Vi=Vi)  # G is left as it is; the run below rewrites V, eS, fS, info
    p.run(G, V, eS, fS, info, Gi=Gi, Vi=Vi)
    S = torch.ldexp(fS, torch.where(torch.isfinite(eS), eS, torch.full_like(eS, -4096)).to(torch.int32))
    S = S.view(B, n)
    if cplx:
        G = torch.complex(G, Gi)
        V = torch.complex(V, Vi)
    U = G.view(B, n, m).transpose(1, 2)
    Vm = V.view(B, n, n).transpose(1, 2)
    order = torch.argsort(S, dim=1, descending=True)
    S = torch.gather(S, 1, order)
    U = torch.gather(U, 2, order.unsqueeze(1).expand(B, m, n))
    Vm = torch.gather(Vm, 2, order.unsqueeze(1).expand(B, n, n))
    idx = U.abs().argmax(dim=1, keepdim=True)
    piv = torch.gather(U, 1, idx)
    mag = piv.abs()
    # the unit factor that makes each column's largest entry real and positive
    ph = torch.where(mag == 0, torch.ones_like(piv), piv.conj() / torch.where(mag == 0, torch.ones_like(mag), mag))
    return (U * ph).contiguous(), S.contiguous(), (Vm * ph).contiguous(), info


def svd_backward(U, S, V, gU, gS, gV):
    """dA of the thin SVD through the HIP adjoint kernel (vjh_dsvd_bwd, or
    vjh_zsvd_bwd on split planes for complex data)."""
    B, m, n = U.shape
    L = _lib.lib()
    z = lambda t, shape, dt: t if t is not None else torch.zeros(shape, dtype=dt, device=U.device)  # noqa: E731
    gU = z(gU, (B, m, n), U.dtype)
    gS = z(gS, (B, n), S.dtype)
    gV = z(gV, (B, n, n), V.dtype)
    cm = lambda t: t.transpose(1, 2).contiguous()                     # noqa: E731  (column-major per matrix)
    Sc, gSc = S.contiguous(), gS.contiguous()
    stream = torch.cuda.current_stream().cuda_stream
    if U.is_complex():
        f = {torch.complex128: L.vjh_zsvd_bwd, torch.complex64: L.vjh_csvd_bwd}[U.dtype]
        Ur, Ui, Vr, Vi = cm(U.real), cm(U.imag), cm(V.real), cm(V.imag)
        gUr, gUi, gVr, gVi = cm(gU.real), cm(gU.imag), cm(gV.real), cm(gV.imag)
        dAr = torch.empty((B, n, m), dtype=S.dtype, device=U.device)
        dAi = torch.empty_like(dAr)
        rc = f(B, m, n, Ur.data_ptr(), Ui.data_ptr(), Sc.data_ptr(), Vr.data_ptr(), Vi.data_ptr(), gUr.data_ptr(),
               gUi.data_ptr(), gSc.data_ptr(), gVr.data_ptr(), gVi.data_ptr(), dAr.data_ptr(), dAi.data_ptr(), stream)
        dA = torch.complex(dAr, dAi)
    else:
        f = {torch.float64: L.vjh_dsvd_bwd, torch.float32: L.vjh_ssvd_bwd}[U.dtype]
        Uc, Vc, gUc, gVc = cm(U), cm(V), cm(gU), cm(gV)
        dA = torch.empty((B, n, m), dtype=U.dtype, device=U.device)
        rc = f(B, m, n, Uc.data_ptr(), Sc.data_ptr(), Vc.data_ptr(), gUc.data_ptr(), gSc.data_ptr(), gVc.data_ptr(),
               dA.data_ptr(), stream)
    _lib.check(rc, "svd backward")
    if rc < 0:
        raise RuntimeError(f"svd backward failed ({rc}): {_lib.last_error()}")
    return dA.transpose(1, 2)


class JacobiSVD(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A):
        U, S, V, info = _forward(A)
        ctx.save_for_backward(U, S, V)
        ctx.mark_non_differentiable(info)
        return U, S, V, info

    @staticmethod
    def backward(ctx, gU, gS, gV, ginfo):
        U, S, V = ctx.saved_tensors
        return svd_backward(U, S, V, gU, gS, gV)


def check_info(info: torch.Tensor, swp: int = SWEEP_LIMIT) -> None:
    """Raise if any member failed (negative code) or used all `swp` sweeps
    without converging (code == swp, dvjsvd.c:604); one device sync."""
    bad = ((info < 0) | (info >= swp)).nonzero().flatten()
    if bad.numel():
        idx = bad[:8].tolist()
        codes = info[bad[:8]].tolist()
        more = f" and {bad.numel() - len(idx)} more" if bad.numel() > len(idx) else ""
        raise RuntimeError(f"jacobi_svd: members {idx}{more} failed or did not converge "
                           f"(codes {codes}; negative = error, {swp} = sweep limit)")


def jacobi_svd(A: torch.Tensor):
    """(U, S, V) of a batch of matrices, differentiable (see module doc).
    Raises RuntimeError when a member fails or does not converge."""
    U, S, V, info = JacobiSVD.apply(A)
    check_info(info)
    return U, S, V


def _declare():
    L = _lib.lib()
    import ctypes as C
    for name in ("vjh_dsvd_bwd", "vjh_ssvd_bwd"):
        fn = getattr(L, name)
        fn.restype = C.c_int64
        fn.argtypes = [C.c_int64, C.c_int64, C.c_int64] + [C.c_void_p] * 7 + [C.c_void_p]
    for name in ("vjh_zsvd_bwd", "vjh_csvd_bwd"):
        fn = getattr(L, name)
        fn.restype = C.c_int64
        fn.argtypes = [C.c_int64, C.c_int64, C.c_int64] + [C.c_void_p] * 12 + [C.c_void_p]


_declare()
