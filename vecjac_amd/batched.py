"""Batched Jacobi SVD of many uneven matrices (one persistent launch per class
of matrices per run: long columns; short columns on global memory; short
columns whose G stays resident in LDS).

Layout (device memory, one buffer per quantity): matrix b has its m_b x n_b
G (A on entry, U on exit) at element offset offG[b], column-major with
leading dimension m_b; V_b (n_b x n_b, ld n_b) at offV[b]; eS/fS (n_b each)
at offS[b].  Widths must be even; an odd-width matrix is solved as the
matrix with one zero column appended (its extra singular value is 0), which
`pack_shapes` arranges.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

KINDS = {"d": 0, "z": 1, "s": 2, "c": 3}


def _declare(L):
    if getattr(L, "_vj_batched_declared", False):
        return
    i64p = C.POINTER(C.c_int64)
    L.vjh_bplan_create.restype = C.c_void_p
    L.vjh_bplan_create.argtypes = [C.c_int, C.c_int64, i64p, i64p, i64p, i64p, i64p, C.c_uint, C.c_int, i64p]
    L.vjh_bplan_run.restype = C.c_int64
    L.vjh_bplan_run.argtypes = [C.c_void_p] + [C.c_void_p] * 6 + [C.c_void_p, C.c_void_p, C.c_void_p]
    L.vjh_bplan_destroy.restype = None
    L.vjh_bplan_destroy.argtypes = [C.c_void_p]
    L.vjh_bplan_waves.restype = C.c_int
    L.vjh_bplan_waves.argtypes = [C.c_void_p]
    L.vjh_bplan_evals.restype = C.c_int64
    L.vjh_bplan_evals.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.vjh_bplan_traffic.restype = C.c_int64
    L.vjh_bplan_traffic.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.vjh_bplan_order.restype = C.c_int64
    L.vjh_bplan_order.argtypes = [C.c_void_p, C.c_int]
    L.vjh_bplan_queue.restype = C.c_int64
    L.vjh_bplan_queue.argtypes = [C.c_void_p, C.c_int]
    L.vjh_bplan_vpreset.restype = C.c_int64
    L.vjh_bplan_vpreset.argtypes = [C.c_void_p, C.c_int]
    L.vjh_bplan_set.restype = C.c_int64
    L.vjh_bplan_set.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_void_p, C.c_void_p]
    L.vjh_bplan_config.restype = C.c_int64
    L.vjh_bplan_config.argtypes = [C.c_void_p, C.c_void_p]
    L.vjh_bplan_autotune.restype = C.c_int64
    L.vjh_bplan_autotune.argtypes = [C.c_void_p] + [C.c_void_p] * 9 + [C.POINTER(C.c_float), C.c_void_p]
    L._vj_batched_declared = True


def pack_shapes(m, n):
    """Offsets for packed storage; odd widths are padded to even (and a
    square odd matrix gets a zero row as well, so that n <= m still holds).
    Returns (m_padded, n_padded, offG, offV, offS, totals)."""
    m = np.asarray(m, dtype=np.int64)
    n = np.asarray(n, dtype=np.int64)
    npad = n + (n & 1)
    m = np.maximum(m, npad)
    sizeG = m * npad
    sizeV = npad * npad
    offG = np.concatenate([[0], np.cumsum(sizeG)[:-1]]).astype(np.int64)
    offV = np.concatenate([[0], np.cumsum(sizeV)[:-1]]).astype(np.int64)
    offS = np.concatenate([[0], np.cumsum(npad)[:-1]]).astype(np.int64)
    return m, npad, offG, offV, offS, (int(sizeG.sum()), int(sizeV.sum()), int(npad.sum()))


class BatchedSVD:
    """A plan for one batch of shapes; `run` is one kernel launch."""

    def __init__(self, m, n, kind: str = "d", swp: int = 999, waves: int = 0):
        L = _lib.lib()
        _declare(L)
        self.kind = kind
        self.m, self.n, self.offG, self.offV, self.offS, self.totals = pack_shapes(m, n)
        if np.any(self.n != np.asarray(n)):
            raise ValueError("odd widths must be padded first (use pack_shapes)")
        st = C.c_int64(0)
        P = C.POINTER(C.c_int64)
        self._keep = [np.ascontiguousarray(a) for a in (self.m, self.n, self.offG, self.offV, self.offS)]
        ptrs = [a.ctypes.data_as(P) for a in self._keep]
        self.plan = L.vjh_bplan_create(KINDS[kind], len(self.m), *ptrs, swp, waves, C.byref(st))
        if not self.plan:
            raise RuntimeError(f"vjh_bplan_create failed ({st.value}): {_lib.last_error()}")
        self.waves = L.vjh_bplan_waves(self.plan)

    def run(self, G, V, eS, fS, info, transforms=None, Gi=None, Vi=None, stream=None) -> int:
        import torch
        L = _lib.lib()
        s = (stream or torch.cuda.current_stream()).cuda_stream
        rc = L.vjh_bplan_run(self.plan, G.data_ptr(), Gi.data_ptr() if Gi is not None else None, V.data_ptr(),
                             Vi.data_ptr() if Vi is not None else None, eS.data_ptr(), fS.data_ptr(),
                             info.data_ptr(), transforms.data_ptr() if transforms is not None else None, s)
        return _lib.check(rc, "vjh_bplan_run")

    def config(self) -> dict:
        """The plan's choices (vjh_bplan_config): matrices and waves per
        workgroup of each launch, the split row count, the LDS budget and the
        LDS bytes per workgroup of the LDS-resident launch."""
        out = np.zeros(12, dtype=np.int64)
        _lib.check(_lib.lib().vjh_bplan_config(self.plan, out.ctypes.data), "vjh_bplan_config")
        names = ("long", "short_global", "short_lds")
        return {"matrices": dict(zip(names, out[0:3].tolist())), "waves": dict(zip(names, out[3:6].tolist())),
                "lanes_per_pair": dict(zip(names, out[6:9].tolist())), "split_rows": int(out[9]),
                "lds_budget": int(out[10]), "lds_bytes_per_workgroup": int(out[11]), "queue_model": self.queue(),
                "ordering": self.order()}

    def set(self, split_rows: int, lds_budget: int, waves=None, lanes=None) -> None:
        """Fix the split row count, the LDS budget (bytes) and optionally the
        waves per workgroup of the three launches and the lanes per pair of
        the two lane-group launches (vjh_bplan_set)."""
        w = None if waves is None else np.asarray(waves, dtype=np.int32)
        ln = None if lanes is None else np.asarray(lanes, dtype=np.int32)
        rc = _lib.lib().vjh_bplan_set(self.plan, int(split_rows), int(lds_budget),
                                      None if w is None else w.ctypes.data, None if ln is None else ln.ctypes.data)
        _lib.check(rc, "vjh_bplan_set")
        if rc < 0:
            raise ValueError(f"vjh_bplan_set: {_lib.last_error()}")
        self.waves = _lib.lib().vjh_bplan_waves(self.plan)

    def autotune(self, G, V, eS, fS, info, Gi=None, Vi=None, stream=None):
        """Tune the plan on a scratch copy of G (and Gi): the split row count,
        the LDS budget and each launch's waves per workgroup
        (vjh_bplan_autotune); the choices stay in the plan (config()).  G is
        not modified; V, eS, fS and info are.  Returns {waves: ms} of the
        runs at the default split."""
        import torch
        L = _lib.lib()
        s = (stream or torch.cuda.current_stream()).cuda_stream
        Gw = torch.empty_like(G)
        Gwi = torch.empty_like(Gi) if Gi is not None else None
        ms = (C.c_float * 4)()
        rc = L.vjh_bplan_autotune(self.plan, G.data_ptr(), Gi.data_ptr() if Gi is not None else None, Gw.data_ptr(),
                                  Gwi.data_ptr() if Gwi is not None else None, V.data_ptr(),
                                  Vi.data_ptr() if Vi is not None else None, eS.data_ptr(), fS.data_ptr(),
                                  info.data_ptr(), ms, s)
        _lib.check(rc, "vjh_bplan_autotune")
        if rc < 0:
            raise RuntimeError(f"vjh_bplan_autotune failed ({rc}): {_lib.last_error()}")
        self.waves = L.vjh_bplan_waves(self.plan)
        return {w: float(ms[k]) for k, w in enumerate((1, 2, 4, 8)) if ms[k] >= 0}

    def evals(self, stream=None) -> np.ndarray:
        """Pair evaluations of each matrix in the last run (vjh_bplan_evals)."""
        import torch
        out = np.zeros(len(self.m), dtype=np.uint64)
        s = (stream or torch.cuda.current_stream()).cuda_stream
        _lib.check(_lib.lib().vjh_bplan_evals(self.plan, out.ctypes.data, s), "vjh_bplan_evals")
        return out

    def traffic(self, stream=None) -> np.ndarray:
        """Elements of G and V each matrix's solve moved to and from device
        memory in the last run (vjh_bplan_traffic)."""
        import torch
        out = np.zeros(len(self.m), dtype=np.uint64)
        s = (stream or torch.cuda.current_stream()).cuda_stream
        _lib.check(_lib.lib().vjh_bplan_traffic(self.plan, out.ctypes.data, s), "vjh_bplan_traffic")
        return out

    def order(self, id_: int = 0) -> int:
        """Set the ordering of the pairs (pjs id 6, the default, or 4) and
        return it; 0 only returns it (vjh_bplan_order)."""
        rc = _lib.lib().vjh_bplan_order(self.plan, int(id_))
        if rc < 0:
            raise ValueError(f"vjh_bplan_order: {_lib.last_error()}")
        return int(rc)

    def queue(self, model: int = -1) -> int:
        """Set the cost model of the processing order (0, 1 or 2) and
        return it; -1 only returns it (vjh_bplan_queue)."""
        rc = _lib.lib().vjh_bplan_queue(self.plan, int(model))
        if rc < 0:
            raise ValueError(f"vjh_bplan_queue: {_lib.last_error()}")
        return int(rc)

    def vpreset(self, on: int = -1) -> bool:
        """Whether runs take V as the starting matrices (vjh_bplan_vpreset);
        on = 1 / 0 sets it, -1 only returns it."""
        return bool(_lib.lib().vjh_bplan_vpreset(self.plan, int(on)))

    def close(self):
        if getattr(self, "plan", None):
            _lib.lib().vjh_bplan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fill_gaussian(m, n, n_real, offG, seed: int, dtype=np.float64, m_real=None):
    """Packed G (padded shapes m x n) with N(0,1) entries in the leading
    m_real[b] x n_real[b] block of each matrix and zeros in its padding."""
    g = np.random.Generator(np.random.Philox(seed))
    m = np.asarray(m)
    n = np.asarray(n)
    m_real = m if m_real is None else np.asarray(m_real)
    total = int(np.sum(m * n))
    G = g.standard_normal(total)  # float64 draws, so the FP32 batch is the same batch rounded
    if G.dtype != dtype:
        G = G.astype(dtype)
    for b in np.nonzero((np.asarray(n_real) != n) | (m_real != m))[0]:
        mm, nn = int(m[b]), int(n[b])
        blk = G[offG[b]: offG[b] + mm * nn].reshape(nn, mm)
        blk[int(n_real[b]):, :] = 0
        blk[:, int(m_real[b]):] = 0
    return G
