#!/usr/bin/env python3
"""Benchmark of the Jacobi SVD / 2x2 EVD hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line.  A "step" is one pass of the hot path over one batch of
synthetic input that is already resident in HBM when timing starts.  The
dominant kernel's average duration is measured with HIP events on the stream
it is launched on, and turned into a roofline fraction from its algorithmic
bytes.  The CPU baseline (rank 0, N=1) times the reference build
(oracle/_ref, when it loads and the host has AVX-512) or the oracle's C
restatement on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def _cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _host_has_avx512() -> bool:
    try:
        return "avx512f" in Path("/proc/cpuinfo").read_text()
    except Exception:
        return False


def _profiled_valu(name: str):
    """VALU issue utilisation of a kernel on its default workload, from the
    rocprofv3 SQ counters recorded in profiles/<name>.md on a line
    `valu_busy: <fraction>`.  None when absent."""
    f = ROOT / "profiles" / f"{name}.md"
    try:
        for line in f.read_text().splitlines():
            if line.strip().startswith("valu_busy:"):
                return float(line.split(":", 1)[1].strip().split()[0])
    except (OSError, ValueError, IndexError):
        pass
    return None


def _profiled_traffic(name: str):
    """HBM bytes per launch of a kernel on its default workload, as measured
    by two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE, each scaled by the
    same counter on a copy of known size) and recorded in profiles/<name>.md
    on a line `traffic_bytes_per_launch: <bytes>`.  None when absent."""
    f = ROOT / "profiles" / f"{name}.md"
    try:
        for line in f.read_text().splitlines():
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1].strip().split()[0]), f"profiles/{name}.md"
    except (OSError, ValueError, IndexError):
        pass
    return None, None


# --------------------------------------------------------------------------- 2x2 EVD
def bench_bjac2(args, kind: str = "d"):
    import torch
    from vecjac_amd import _lib, workloads as W

    L = _lib.lib()
    n = 1 << int(args.lg)
    cplx = kind in "zc"
    dt_np = np.float64 if kind in "dz" else np.float32
    dt = torch.float64 if kind in "dz" else torch.float32
    w = 8 if kind in "dz" else 4
    vl = 8 if kind in "dz" else 16
    ins_np = [a for a in W.jac2_batch(n, 2026, cplx, dtype=dt_np) if a is not None]
    ins = [torch.from_numpy(a).to("cuda") for a in ins_np]
    outs = [torch.empty(n, dtype=dt, device="cuda") for _ in range(5 if cplx else 4)]
    p = torch.empty(n // vl, dtype=torch.int32, device="cuda")
    f = getattr(L, f"vjh_{kind}bjac2")
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    ptrs = [t.data_ptr() for t in ins] + [t.data_ptr() for t in outs] + [p.data_ptr()]

    def step():
        rc = f(-n, *ptrs, s)
        if rc != 0:
            raise RuntimeError(f"vjh_{kind}bjac2 returned {rc}: {_lib.last_error()}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # the K launches back to back between two events on their stream: an event
    # pair around each sub-millisecond launch adds its own dispatch latency
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        step()
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    k_avg_s = ev[0].elapsed_time(ev[1]) / args.steps / 1e3
    per = (len(ins_np) + len(outs)) * w  # algorithmic bytes per matrix (+ 1 bit of p)
    traffic_alg = n * per + n // 8
    achieved = traffic_alg / k_avg_s / 1e9
    value = n * args.steps / wall
    out = {
        "metric": f"2x2 Hermitian EVDs/s ({kind}bjac2, sine form)",
        "value": value, "unit": "EVDs/s",
        "ms_per_step": wall / args.steps * 1e3,
        "dtype": {"d": "f64", "z": "c128", "s": "f32", "c": "c64"}[kind],
        "config": {"workload": f"{kind}bjac2: 2^{args.lg} 2x2 problems, components +-2^U[-100,100], 1% b=0, 1% a=d",
                   "batch": n},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"k_{'c' if cplx else 'r'}bjac2", "kernel_avg_ms": k_avg_s * 1e3,
                     "alg_bytes_per_launch": traffic_alg},
    }
    if int(args.lg) == 26:
        t, src = _profiled_traffic(f"k_{'c' if cplx else 'r'}bjac2_{kind}_2e26")
        out["roofline"]["traffic"] = t
        if src:
            out["roofline"]["traffic_source"] = src
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_bjac2(kind, ins_np)
    return out


def cpu_bjac2(kind: str, ins_np):
    """Reference build when usable, else the oracle; bounded sample.  The
    aligned inputs and outputs are staged once, outside the timed calls."""
    import ctypes as C
    os.environ.setdefault("OMP_NUM_THREADS", str(_cpu_threads()))
    from oracle import vjo
    use_ref = vjo.ref() is not None and _host_has_avx512()
    dt = np.float64 if kind in "dz" else np.float32
    nsamp = min(ins_np[0].size, 1 << 22)
    ins = []
    for a in ins_np:
        b = vjo.aligned(nsamp, dt)
        b[...] = a[:nsamp]
        ins.append(b)
    outs = [vjo.aligned(nsamp, dt) for _ in range(5 if kind in "zc" else 4)]
    p = vjo.aligned(max(1, nsamp // (8 if dt == np.float64 else 16)), np.uint32)
    f = getattr(vjo.ref(), f"{kind}bjac2_") if use_ref else getattr(vjo.oracle(), f"vjo_{kind}bjac2")
    nn = C.c_int64(-nsamp)
    argv = [C.byref(nn)] + [a.ctypes.data for a in ins + outs] + [p.ctypes.data]
    reps, t_total = 0, 0.0
    while t_total < 10.0 and reps < 5000:
        t0 = time.perf_counter()
        f(*argv)
        t_total += time.perf_counter() - t0
        reps += 1
    return {"value": nsamp * reps / t_total, "unit": "EVDs/s",
            "cores": int(os.environ["OMP_NUM_THREADS"]) if use_ref else 1,
            "kind": "reference" if use_ref else "port",
            "sample": f"{reps} x {nsamp} problems ({'AVX-512 reference build, OpenMP' if use_ref else 'oracle C restatement, 1 thread'}); "
                      "inputs staged once, only the calls timed"}



# --------------------------------------------------------------------------- batched SVD
def _bsvd_alg_bytes(m, n, evals, transforms, w):
    """Per-step column bytes of one batched run, as if every pair moved
    through memory: every pair evaluation reads the pair's two G columns
    (w 2m) and a transformed pair writes them and reads/writes its two V
    columns (w (2m + 4n)).  evals counts the pair evaluations the run made (a
    pair whose columns did not change since its previous evaluation is not
    evaluated again).  (The LDS-resident form moves the G bytes between LDS
    and registers, and the V bytes through the caches.)"""
    m = m.astype(np.float64)
    n = n.astype(np.float64)
    return float(np.sum(evals.astype(np.float64) * w * 2 * m + transforms.astype(np.float64) * w * (2 * m + 4 * n)))


def bench_bsvd(args, kind: str = "d"):
    import torch
    from vecjac_amd import batched as BT
    from vecjac_amd import svd as S
    from vecjac_amd import workloads as W

    nb = args.nb
    m_real, n_real = W.batch_shapes(nb, 4242, 8, 128)
    m, npad, offG, offV, offS, tot = BT.pack_shapes(m_real, n_real)
    G64 = BT.fill_gaussian(m, npad, n_real, offG, 4243, np.float64, m_real=m_real)

    def run_batch(kind, steps, warmup, timed=True):
        dt_np = np.float64 if kind == "d" else np.float32
        dt = torch.float64 if kind == "d" else torch.float32
        plan = BT.BatchedSVD(m, npad, kind=kind, swp=999, waves=args.waves)
        G0 = torch.from_numpy(G64.astype(dt_np)).cuda()
        G = torch.empty_like(G0)
        V = torch.empty(tot[1], dtype=dt, device="cuda")
        eS = torch.empty(tot[2], dtype=dt, device="cuda")
        fS = torch.empty(tot[2], dtype=dt, device="cuda")
        info = torch.empty(nb, dtype=torch.int64, device="cuda")
        tr = torch.empty(nb, dtype=torch.int64, device="cuda")
        stream = torch.cuda.current_stream()
        tune = None
        if args.order:
            plan.order(args.order)
        if args.plan:
            sp, bud, wv, ln = args.plan.split(":")
            plan.set(int(sp), int(bud), [int(x) for x in wv.split(",")], [16] + [int(x) for x in ln.split(",")])
            tune = {"fixed": args.plan}
        elif args.waves == 0 and not args.no_tune:
            # plan-time autotuner (split, LDS budget, waves per workgroup),
            # outside the timed region
            t_tune = time.perf_counter()
            tune = plan.autotune(G0, V, eS, fS, info)
            tune = {"ms_by_waves": tune, "seconds": time.perf_counter() - t_tune}
        for _ in range(warmup):
            G.copy_(G0)
            plan.run(G, V, eS, fS, info, tr)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * steps)]
        t0 = time.perf_counter()
        for i in range(steps):
            G.copy_(G0)
            ev[2 * i].record(stream)
            plan.run(G, V, eS, fS, info, tr)
            ev[2 * i + 1].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kms = [ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(steps)]
        res = dict(plan=plan, wall=wall, kernel_ms=float(np.mean(kms)), info=info.cpu().numpy(),
                   tr=tr.cpu().numpy(), G=G, eS=eS.cpu().numpy(), fS=fS.cpu().numpy(), waves=plan.waves, tune=tune,
                   evals=plan.evals(), xfer=plan.traffic(), plan_config=plan.config(), order=plan.order())
        return res

    w = 8 if kind == "d" else 4
    r = run_batch(kind, args.steps, args.warmup)
    info, tr = r["info"], r["tr"]
    if np.any(info < 0):
        raise RuntimeError(f"batched SVD failed on {int(np.sum(info < 0))} matrices")
    # algorithmic bytes (the roofline's numerator): the per-step column
    # traffic of the one-sided Jacobi method, w m per column read by a pair
    # evaluation and w (2m + 4n) per transformed pair, over the evaluations
    # the run made
    step_bytes = _bsvd_alg_bytes(m, npad, r["evals"], tr, w)
    # the G and V elements this design actually moves to and from memory
    # (every column load and store of the steps or block rounds, the set-up
    # scan, V := I, the scaling and the epilogue; G in and U out once for a
    # matrix whose G is held in LDS), counted by the kernel (vjh_bplan_traffic): below
    # the algorithmic bytes where block rounds and LDS residence reuse columns
    # on chip
    moved = float(np.sum(r["xfer"].astype(np.float64))) * w
    k_s = r["kernel_ms"] / 1e3
    achieved = step_bytes / k_s / 1e9
    value = nb * args.steps / r["wall"]
    out = {
        "metric": "SVDs/s (batched vs plain) at sigma rel err <= 8 n eps",
        "value": value, "unit": "SVDs/s",
        "ms_per_step": r["wall"] / args.steps * 1e3,
        "dtype": "f64" if kind == "d" else "f32",
        "config": {"workload": f"batched {'FP64' if kind == 'd' else 'FP32'} one-sided Jacobi SVD of {nb} matrices, "
                               "n ~ U{8..128} (odd n padded with a zero column), m ~ U{n..4n}, N(0,1) entries, "
                               "one persistent launch per class (long columns; short on global memory; short "
                               "resident in LDS)",
                   "batch": nb, "ordering": f"pjs id {r['order']}", "plan": r["plan_config"], "autotune": r["tune"],
                   "sweeps_mean": float(np.mean(info)), "sweeps_max": int(np.max(info)),
                   "pair_evaluations": int(np.sum(r["evals"])),
                   "pair_slots": int(np.sum((np.minimum(info, 998) + 1) * npad * (npad // 2)))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_bsvd",
                     "kernel_avg_ms": r["kernel_ms"], "alg_bytes_per_launch": step_bytes,
                     "moved_bytes_per_launch": moved,
                     "moved_frac": moved / k_s / 1e9 / HBM_PEAK_GBS},
    }
    if nb == 16384 and args.waves == 0:
        t, src = _profiled_traffic(f"k_bsvd_{'fp64' if kind == 'd' else 'fp32'}_16384")
        out["roofline"]["traffic"] = t
        if src:
            out["roofline"]["traffic_source"] = src
            # the long-column launch is bound by VALU issue, not by HBM
            out["roofline"]["valu_busy"] = _profiled_valu(f"k_bsvd_{'fp64' if kind == 'd' else 'fp32'}_16384")
    # sigma accuracy on a sample, against the oracle on the same matrices
    if not args.no_check:
        out["sigma_check"] = _bsvd_sigma_check(m, npad, offG, offS, G64, r, kind)
    # plain per-matrix launches on a sample of the same batch
    if not args.no_plain:
        out["plain"] = _bsvd_plain_sample(m, npad, offG, G64, kind, args.plain_sample)
        out["plain"]["batched_over_plain"] = value / out["plain"]["value"]
    # FP32 run of the same batch (one timed step)
    if kind == "d" and not args.no_fp32:
        r32 = run_batch("s", 1, 1)
        out["fp32"] = {"value": nb / (r32["kernel_ms"] / 1e3), "unit": "SVDs/s", "kernel_ms": r32["kernel_ms"],
                       "sweeps_mean": float(np.mean(r32["info"]))}
        r32["plan"].close()
    r["plan"].close()
    if not args.no_cpu:
        out["cpu_baseline"] = _bsvd_cpu(m_real, n_real, m, offG, npad, G64, kind)
    return out


def _bsvd_sigma_check(m, npad, offG, offS, G64, r, kind, nsamp=6):
    from oracle import vjo
    from vecjac_amd import svd as S
    eps = np.finfo(np.float64 if kind == "d" else np.float32).eps
    g = np.random.Generator(np.random.Philox(7))
    idx = g.choice(len(m), nsamp, replace=False)
    worst = {0: 0.0, 1: 0.0}
    for b in idx:
        mm, nn = int(m[b]), int(npad[b])
        A = G64[offG[b]: offG[b] + mm * nn].reshape(nn, mm).astype(np.float64 if kind == "d" else np.float32)
        js, stp = S.pjs(6, nn)
        Go = vjo.aligned((nn, mm), A.dtype)
        Go[...] = A
        sg = np.sort(S.sigma(r["eS"][offS[b]: offS[b] + nn], r["fS"][offS[b]: offS[b] + nn]))
        for mode in (1, 0):  # the GPU's norm and dot; the reference's own arithmetic
            Go[...] = A
            with vjo.norm_mode(mode):
                rc, eS, fS, _, _, _ = vjo.vjsvd_core(kind, Go, js=js, stp=stp, swp=999)
            so = np.sort(S.sigma(eS, fS))
            nz = so > 0
            rel = float(np.max(np.abs(sg[nz] - so[nz]) / so[nz]) / (8 * nn * eps))
            worst[mode] = max(worst[mode], rel)
    return {"matrices": int(nsamp), "max_rel_err_over_8n_eps": worst[1], "max_rel_err_over_8n_eps_mode0": worst[0],
            "reference": "oracle, same input and table (mode 1: the GPU's norm and dot; mode 0: the reference's)"}


def _bsvd_plain_sample(m, npad, offG, G64, kind, nsamp):
    import torch
    from vecjac_amd import svd as S
    g = np.random.Generator(np.random.Philox(8))
    idx = g.choice(len(m), nsamp, replace=False)
    dt_np = np.float64 if kind == "d" else np.float32
    dt = torch.float64 if kind == "d" else torch.float32
    mats = []
    for b in idx:
        mm, nn = int(m[b]), int(npad[b])
        A = G64[offG[b]: offG[b] + mm * nn].reshape(nn, mm).astype(dt_np)
        js, stp = S.pjs(6, nn)
        mats.append((torch.from_numpy(A).cuda(), torch.empty((nn, nn), dtype=dt, device="cuda"),
                     torch.empty(nn, dtype=dt, device="cuda"), torch.empty(nn, dtype=dt, device="cuda"),
                     torch.from_numpy(js.reshape(-1).view(np.int32)).cuda(), stp))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for (A, V, eS, fS, js, stp) in mats:
        Gw = A.clone()
        S.device_vjsvd(kind, Gw, V, eS, fS, js, stp)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    return {"value": nsamp / t, "unit": "SVDs/s", "sample": f"{nsamp} matrices of the batch, each through "
            "vjh_dvjsvd (one launch per step)"}


def _bsvd_cpu(m_real, n_real, m, offG, npad, G64, kind):
    """Reference dvjsvd_ (its own sources, AVX-512; long-double norm) over a
    thread pool, one matrix per call, inputs zero-padded to m%8 == 0 and
    n%16 == 0 as the reference requires; the oracle port when the reference
    build cannot run here.  Matrices are staged (padded copies, step tables,
    workspaces) in chunks outside the timed region; only the calls of each
    chunk, run side by side on the pool, are timed."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor
    from oracle import vjo
    from vecjac_amd import svd as S
    threads = _cpu_threads()
    use_ref = vjo.ref() is not None and _host_has_avx512() and kind == "d"
    dt = np.float64 if kind == "d" else np.float32
    g = np.random.Generator(np.random.Philox(9))
    order = g.permutation(len(m_real))
    budget = 15.0
    tables = {}

    def table(nn):
        if nn not in tables:
            js, stp = S.pjs(6, nn)
            jsa = vjo.aligned(js.shape, np.uint32)
            jsa[...] = js
            tables[nn] = (jsa, stp)
        return tables[nn]

    def stage(b):
        """Everything one call needs, allocated and filled before timing."""
        mm, nn_pad, nn, mr = int(m[b]), int(npad[b]), int(n_real[b]), int(m_real[b])
        src = G64[offG[b]: offG[b] + mm * nn_pad].reshape(nn_pad, mm)
        if use_ref:
            m8 = (mr + 7) // 8 * 8
            n16 = max(16, (nn + 15) // 16 * 16)
            m8 = max(m8, n16)
            A = vjo.aligned((n16, m8), np.float64)
            A[:nn, :mr] = src[:nn, :mr]
            rows, cols = m8, n16
        else:
            A = vjo.aligned((nn_pad, mm), dt)
            A[...] = src
            rows, cols = mm, nn_pad
        jsa, stp = table(cols)
        return dict(A=A, rows=rows, cols=cols, js=jsa, stp=stp, V=vjo.aligned((cols, cols), dt),
                    eS=vjo.aligned(cols, dt), fS=vjo.aligned(cols, dt), work=vjo.aligned(8 * cols, dt),
                    iw=vjo.aligned(max(1, cols // 8), np.uint32))

    if use_ref:
        R = vjo.ref()
        gomp = C.CDLL("libgomp.so.1")

        def init():
            gomp.omp_set_num_threads(1)

        def one(x):
            mmc, nnc = C.c_uint64(x["rows"]), C.c_uint64(x["cols"])
            return R.dvjsvd_(C.byref(mmc), C.byref(nnc), x["A"].ctypes.data, C.byref(mmc), x["V"].ctypes.data,
                             C.byref(nnc), x["eS"].ctypes.data, x["fS"].ctypes.data, x["js"].ctypes.data,
                             C.byref(C.c_uint(x["stp"])), C.byref(C.c_uint(999)), x["work"].ctypes.data,
                             x["iw"].ctypes.data)
    else:
        L = vjo.oracle()
        core = getattr(L, f"vjo_{kind}vjsvd_core")

        def init():
            pass

        def one(x):
            mm_, nn_ = vjo._fnat(x["rows"]), vjo._fnat(x["cols"])
            return core(0, C.byref(mm_), C.byref(nn_), x["A"].ctypes.data, C.byref(mm_), None, C.byref(mm_),
                        x["V"].ctypes.data, C.byref(nn_), None, C.byref(nn_), x["eS"].ctypes.data, x["fS"].ctypes.data,
                        x["js"].ctypes.data, C.byref(C.c_uint(x["stp"])), C.byref(C.c_uint(999)), x["work"].ctypes.data,
                        x["iw"].ctypes.data, None)
    done, t = 0, 0.0
    with ThreadPoolExecutor(max_workers=threads, initializer=init) as ex:
        list(ex.map(lambda _: None, range(threads)))  # start the workers
        while t < budget and done < len(order):
            chunk = [stage(b) for b in order[done: done + 16 * threads]]
            t0 = time.perf_counter()
            list(ex.map(one, chunk))
            t += time.perf_counter() - t0
            done += len(chunk)
    return {"value": done / t, "unit": "SVDs/s", "cores": threads, "kind": "reference" if use_ref else "port",
            "sample": f"{done} matrices of the batch, {t:.1f} s of calls ("
                      + ("reference dvjsvd_ built from its sources, zero-padded to m%8, n%16, one matrix per thread"
                         if use_ref else "oracle C restatement, one matrix per thread")
                      + "; staging, padding and step tables outside the timed calls)"}


def bench_plain(args, kind: str):
    """Plain SVD of one large matrix through the device driver (one launch
    per step, the V updates of each sweep applied in one pass after it):
    dvjsvd on a 1024 x 1024 N(0,1) matrix, or zvjsvd on an 8192 x 2048
    complex A = Q1 diag(sigma) Q2^H with sigma geometric from 1 to 1e-12.
    The roofline numerator is the algorithm's bytes (every pair evaluation
    reads its two G columns, every transformation writes them and reads and
    writes its two V columns) over the whole run.""" 
    import torch
    from vecjac_amd import svd as S
    from vecjac_amd import workloads as W
    if kind == "d":
        m, n = args.m or 1024, args.n or 1024
        A = W.gaussian_matrix(m, n, 1024)
        Ar, Ai = A, None
        desc = f"plain dvjsvd, FP64 {m}x{n}, N(0,1), ME table, swp=999, V=I"
    else:
        m, n = args.m or 8192, args.n or 2048
        sig = 10.0 ** (-12.0 * np.arange(n) / (n - 1))
        A = W.graded_svd_matrix(m, n, sig, 8192, True)  # m x n complex
        Ar = np.ascontiguousarray(A.real.T)
        Ai = np.ascontiguousarray(A.imag.T)
        desc = f"plain zvjsvd, complex FP64 {m}x{n}, A = Q1 diag(10^(-12 i/(n-1))) Q2^H, Haar Q1 Q2, ME table"
    print(f"bench plain {kind}: {m}x{n} input ready", file=sys.stderr, flush=True)
    js, stp = S.pjs(2, n)
    jsd = torch.from_numpy(js.reshape(-1).view(np.int32)).cuda()
    G0 = torch.from_numpy(np.ascontiguousarray(Ar)).cuda()
    Gi0 = torch.from_numpy(Ai).cuda() if Ai is not None else None
    G = torch.empty_like(G0)
    Gi = torch.empty_like(Gi0) if Gi0 is not None else None
    V = torch.empty((n, n), dtype=torch.float64, device="cuda")
    Vi = torch.empty((n, n), dtype=torch.float64, device="cuda") if Ai is not None else None
    eS = torch.empty(n, dtype=torch.float64, device="cuda")
    fS = torch.empty(n, dtype=torch.float64, device="cuda")
    walls, sweeps, hist = [], None, None
    for it in range(args.warmup + args.steps):
        G.copy_(G0)
        if Gi is not None:
            Gi.copy_(Gi0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc, st = S.device_vjsvd(kind, G, V, eS, fS, jsd, stp, 999, Gi=Gi, Vi=Vi)
        torch.cuda.synchronize()
        if it >= args.warmup:
            walls.append(time.perf_counter() - t0)
        sweeps, hist = rc, st
        print(f"bench plain {kind}: run {it} {time.perf_counter() - t0:.2f} s, {rc} sweeps", file=sys.stderr, flush=True)
    wall = float(np.mean(walls))
    w = 8 if kind == "d" else 16
    R = float(np.sum(hist[2][: max(0, sweeps) + 1]))
    steps_run = (sweeps + 1) * stp
    alg = steps_run * w * m * n + R * w * (2 * m + 4 * n)
    achieved = alg / wall / 1e9
    out = {"metric": "SVDs/s (plain, one matrix)", "value": 1.0 / wall, "unit": "SVDs/s",
           "ms_per_step": wall * 1e3, "dtype": "f64" if kind == "d" else "c128",
           "config": {"workload": desc, "sweeps": int(sweeps), "launched_sweeps": int(hist[1]),
                      "transformations": R},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "k_step, k_fold and k_vflush (the whole run)",
                        "kernel_avg_ms": wall / max(1, int(hist[1]) * stp) * 1e3,
                        "alg_bytes_per_launch": alg / max(1, steps_run)}}
    sg = np.sort(S.sigma(eS.cpu().numpy(), fS.cpu().numpy()))[::-1]
    if kind == "z":
        # against the nominal sigma of the construction: A is formed in FP64, so
        # sigma_i moves by about eps |A| = eps; compared where that is far below 8 n eps sigma_i
        sel = sig >= 1e-4
        out["sigma_check"] = {"vs": "nominal sigma of A = Q1 diag(sigma) Q2^H, sigma_i >= 1e-4",
                              "max_rel_err_over_8n_eps": float(np.max(np.abs(sg[sel] - sig[sel]) / sig[sel]) /
                                                               (8 * n * np.finfo(np.float64).eps))}
    if kind == "z" and not args.no_cpu:
        out["cpu_baseline"] = _plain_cpu_z(Ar, Ai, m, n, js, steps_run)
    if kind == "d" and not args.no_cpu:
        out["cpu_baseline"], s_ref = _plain_cpu(A, m, n, js, stp)
        if s_ref is not None:
            s_ref = np.sort(s_ref)[::-1]
            out["sigma_check"] = {"vs": out["cpu_baseline"]["kind"] + " dvjsvd_ on the host, same input and table",
                                  "max_rel_err_over_8n_eps": float(np.max(np.abs(sg - s_ref) / s_ref) /
                                                                   (8 * n * np.finfo(np.float64).eps))}
    return out


def _plain_cpu_z(Ar, Ai, m, n, js, steps_gpu, S=16, target_s=10.0):
    """Complex case: the reference's zvjsvd_ needs BLAS and is not built, so
    the oracle's restatement (its norm, dot and apply loops over the pairs of
    a step on an OpenMP team, as the reference's) runs the first steps of the
    first sweep on the same input and table: S steps to size the sample, then
    about target_s seconds of steps, whose rate is extrapolated to the number
    of steps the GPU run took (setup included in the sample, so the CPU
    figure is slightly pessimistic)."""
    os.environ.setdefault("OMP_NUM_THREADS", str(_cpu_threads()))
    from oracle import vjo

    def run(steps):
        G = vjo.aligned((n, m), np.float64)
        G[...] = Ar
        Gi = vjo.aligned((n, m), np.float64)
        Gi[...] = Ai
        jsS = np.ascontiguousarray(js[:steps])
        t0 = time.perf_counter()
        vjo.vjsvd_core("z", G, js=jsS, stp=steps, swp=1, Gi=Gi)
        return time.perf_counter() - t0

    t = run(S)
    S2 = min(int(js.shape[0]), int(S * target_s / max(t, 1e-3)))
    if S2 > S:
        S, t = S2, run(S2)
    return {"value": 1.0 / (t / S * steps_gpu), "unit": "SVDs/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": "port",
            "sample": f"first {S} steps of sweep 1 of the oracle's zvjsvd (OpenMP over the pairs of a step) on the "
                      f"same input and table, "
                      f"{t:.1f} s, extrapolated to the {int(steps_gpu)} steps of the GPU run"}


def _plain_cpu(A, m, n, js, stp):
    """The reference's dvjsvd_ (its own sources, AVX-512, OpenMP over the host
    threads) on the same matrix and table; the oracle port (serial) when the
    reference build cannot run here.  Returns (baseline, sigma)."""
    import ctypes as C
    from oracle import vjo
    from vecjac_amd import svd as S
    threads = _cpu_threads()
    G = vjo.aligned((n, m), np.float64)
    G[...] = A.reshape(n, m)
    jsa = vjo.aligned(js.shape, np.uint32)
    jsa[...] = js
    V = vjo.aligned((n, n), np.float64)
    eS = vjo.aligned(n, np.float64)
    fS = vjo.aligned(n, np.float64)
    if vjo.ref() is not None and _host_has_avx512() and m % 8 == 0 and n % 16 == 0:
        R = vjo.ref()
        C.CDLL("libgomp.so.1").omp_set_num_threads(threads)
        work = vjo.aligned(8 * n, np.float64)
        iw = vjo.aligned(max(1, n // 8), np.uint32)
        mm, nn = C.c_uint64(m), C.c_uint64(n)
        t0 = time.perf_counter()
        rc = R.dvjsvd_(C.byref(mm), C.byref(nn), G.ctypes.data, C.byref(mm), V.ctypes.data, C.byref(nn),
                       eS.ctypes.data, fS.ctypes.data, jsa.ctypes.data, C.byref(C.c_uint(stp)), C.byref(C.c_uint(999)),
                       work.ctypes.data, iw.ctypes.data)
        t = time.perf_counter() - t0
        kind, cores = "reference", threads
    else:
        t0 = time.perf_counter()
        rc, eS, fS, *_ = vjo.vjsvd_core("d", G, js=js, stp=stp, swp=999)
        t = time.perf_counter() - t0
        kind, cores = "port", 1
    base = {"value": 1.0 / t, "unit": "SVDs/s", "cores": cores, "kind": kind,
            "sample": f"the whole {m}x{n} SVD, one call ({t:.1f} s, {int(rc)} sweeps)"}
    return base, (S.sigma(eS, fS) if rc >= 0 else None)


def bench_copy(args):
    """Plain device-to-device copy of 2^lg doubles: the calibration kernel
    for the FETCH_SIZE / WRITE_SIZE counters (its traffic is known)."""
    import torch
    n = 1 << int(args.lg)
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        b.copy_(a)
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    k = ev[0].elapsed_time(ev[1]) / args.steps / 1e3
    bytes_ = 2 * 8 * n
    return {"metric": "device copy bandwidth (calibration)", "value": bytes_ / k / 1e9, "unit": "GB/s",
            "ms_per_step": wall / args.steps * 1e3, "dtype": "f64",
            "config": {"workload": f"copy of 2^{args.lg} doubles", "bytes_per_copy": bytes_},
            "roofline": {"bound": "hbm", "achieved": bytes_ / k / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": bytes_ / k / 1e9 / HBM_PEAK_GBS, "traffic": None, "kernel": "copy",
                         "kernel_avg_ms": k * 1e3, "alg_bytes_per_launch": bytes_}}

# --------------------------------------------------------------------------- differentiable SVD
def bench_grad(args):
    """Forward (batched Jacobi SVD) plus the hand-written adjoint backward of
    4096 FP64 matrices of 64 x 32, A = Q1 diag(s) Q2^T with s ~ U[1, 10] and
    minimum relative gap 1e-3, loss sum w_i s_i + <G_U, U> + <G_V, V>."""
    import torch
    from vecjac_amd import autograd as AG
    from vecjac_amd import workloads as W

    nb = 4096 if args.nb == 16384 else args.nb
    m, n = 64, 32
    A, w, GU, GV = W.grad_batch(nb, m, n, 515)
    At = torch.from_numpy(A).cuda()
    wt, GUt, GVt = (torch.from_numpy(x).cuda() for x in (w, GU, GV))
    stream = torch.cuda.current_stream()

    def step():
        X = At.detach().requires_grad_(True)
        U, S, V = AG.jacobi_svd(X)
        loss = (wt * S).sum() + (GUt * U).sum() + (GVt * V).sum()
        loss.backward()
        return X.grad

    AG._forward(At)  # set-up: creates and autotunes the plan, outside the timed steps even at --warmup 0
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g = step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    # the two kernels alone, with HIP events on their stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(stream)
    U, S, V, info = AG._forward(At)
    ev[1].record(stream)
    torch.cuda.synchronize()
    f_ms = ev[0].elapsed_time(ev[1])
    # the backward kernel alone: its operands already column-major per matrix
    # (the wrapper's transposes are outside), averaged over the launches
    from vecjac_amd import _lib
    cm = lambda t: t.transpose(1, 2).contiguous()  # noqa: E731
    Uc, Vc, gUc, gVc = cm(U), cm(V), cm(GUt), cm(GVt)
    Sc, gSc = S.contiguous(), wt.contiguous()
    dAk = torch.empty((nb, n, m), dtype=torch.float64, device="cuda")
    L = _lib.lib()
    bw = []
    for _ in range(max(3, args.steps)):
        ev[2].record(stream)
        rc = L.vjh_dsvd_bwd(nb, m, n, Uc.data_ptr(), Sc.data_ptr(), Vc.data_ptr(), gUc.data_ptr(), gSc.data_ptr(),
                            gVc.data_ptr(), dAk.data_ptr(), stream.cuda_stream)
        ev[3].record(stream)
        torch.cuda.synchronize()
        assert rc == 0, _lib.last_error()
        bw.append(ev[2].elapsed_time(ev[3]))
    b_ms = float(np.mean(bw[1:]))
    # backward kernel bytes: reads U, dU (m n), V, dV (n^2), s, ds (n); writes dA (m n)
    bwd_bytes = nb * 8 * (3 * m * n + 2 * n * n + 2 * n)
    # the forward kernel alone (the cached plan's run, 90% of the step), with the
    # pair evaluations and transformations its algorithmic bytes are counted from
    plan, _ = AG._plan(nb, m, n, "d")
    G0 = At.transpose(1, 2).contiguous().reshape(-1)
    Gk = torch.empty_like(G0)
    Vk = torch.empty(nb * n * n, dtype=torch.float64, device="cuda")
    eSk, fSk = (torch.empty(nb * n, dtype=torch.float64, device="cuda") for _ in range(2))
    infok, trk = (torch.empty(nb, dtype=torch.int64, device="cuda") for _ in range(2))
    fk = []
    for _ in range(max(1, args.steps)):
        Gk.copy_(G0)
        ev[0].record(stream)
        plan.run(Gk, Vk, eSk, fSk, infok, trk)
        ev[1].record(stream)
        torch.cuda.synchronize()
        fk.append(ev[0].elapsed_time(ev[1]))
    fk_ms = float(np.mean(fk))
    fwd_bytes = _bsvd_alg_bytes(np.full(nb, m), np.full(nb, n), plan.evals(), trk.cpu().numpy(), 8)
    # gradient against torch.linalg.svd FP64 autograd on the CPU (same gauge), on a sample
    ns = 64
    Ac = torch.from_numpy(A[:ns]).requires_grad_(True)
    Uc, Sc, Vh = torch.linalg.svd(Ac, full_matrices=False)
    order = torch.argsort(Sc, dim=-1, descending=True)
    Sc = torch.gather(Sc, -1, order)
    Uc = torch.gather(Uc, -1, order.unsqueeze(-2).expand_as(Uc))
    Vc = torch.gather(Vh.transpose(-2, -1), -1, order.unsqueeze(-2).expand(ns, n, n))
    idx = Uc.abs().argmax(dim=-2, keepdim=True)
    sgn = torch.sign(torch.gather(Uc, -2, idx))
    Uc, Vc = Uc * sgn, Vc * sgn
    lc = (torch.from_numpy(w[:ns]) * Sc).sum() + (torch.from_numpy(GU[:ns]) * Uc).sum() + \
        (torch.from_numpy(GV[:ns]) * Vc).sum()
    lc.backward()
    gref = Ac.grad
    rel = float((g[:ns].cpu() - gref).abs().max() / gref.abs().max())
    out = {
        "metric": "batched SVD forward + adjoint backward ms",
        "value": wall * 1e3, "unit": "ms", "higher_is_better": False,
        "ms_per_step": wall * 1e3, "dtype": "f64",
        "config": {"workload": f"differentiable batched SVD: {nb} FP64 matrices of {m}x{n}, s ~ U[1,10] "
                               "(min rel gap 1e-3), loss sum w s + <G_U,U> + <G_V,V>, torch.autograd",
                   "batch": nb, "forward_kernel_ms": f_ms, "backward_kernel_ms": b_ms,
                   "sweeps_mean": float(info.double().mean()),
                   "plan": plan.config(), "plan_tuning": "autotuned in the first (set-up) call"},
        "roofline": {"bound": "hbm", "achieved": fwd_bytes / (fk_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": fwd_bytes / (fk_ms / 1e3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_bsvd (forward)", "kernel_avg_ms": fk_ms, "alg_bytes_per_launch": fwd_bytes,
                     "note": "algorithmic bytes are the per-step column traffic; in the LDS-resident class the G "
                             "columns move between LDS and registers and the V columns through the caches, and HBM "
                             "sees about 2.7 GB per run (profiles/k_bsvd_fp64_4096x64x32.md): the launch is bound by "
                             "VALU instruction issue",
                     "backward": {"kernel": "k_svd_bwd_stg" if (m * n <= 2048 and n * n <= 2048) else "k_svd_bwd",
                                  "kernel_avg_ms": b_ms, "alg_bytes_per_launch": bwd_bytes,
                                  "achieved": bwd_bytes / (b_ms / 1e3) / 1e9,
                                  "frac": bwd_bytes / (b_ms / 1e3) / 1e9 / HBM_PEAK_GBS}},
        "grad_check": {"vs": "torch.linalg.svd FP64 autograd (CPU), same gauge", "matrices": ns,
                       "max_rel_err": rel},
    }
    if not args.no_cpu:
        # the reference's forward (dvjsvd_) on the host; the reference has no backward
        m_real = np.full(nb, m, dtype=np.int64)
        n_real = np.full(nb, n, dtype=np.int64)
        from vecjac_amd import batched as BT
        mm, npad, offG, _, _, _ = BT.pack_shapes(m_real, n_real)
        G64 = np.ascontiguousarray(np.transpose(A, (0, 2, 1))).reshape(-1)
        out["cpu_baseline"] = _bsvd_cpu(m_real, n_real, mm, offG, npad, G64, "d")
        out["cpu_baseline"]["note"] = "forward only (dvjsvd_ per matrix); the reference has no SVD backward"
    return out


WORKLOADS = {
    "copy": bench_copy,
    "svd1024": lambda a: bench_plain(a, "d"),
    "zsvd": lambda a: bench_plain(a, "z"),
    "bsvd": lambda a: bench_bsvd(a, "d"),
    "bsvd32": lambda a: bench_bsvd(a, "s"),
    "grad": bench_grad,
    "bjac2": lambda a: bench_bjac2(a, "d"),
    "zbjac2": lambda a: bench_bjac2(a, "z"),
    "sbjac2": lambda a: bench_bjac2(a, "s"),
    "cbjac2": lambda a: bench_bjac2(a, "c"),
}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="bsvd", choices=sorted(WORKLOADS))
    ap.add_argument("--nb", type=int, default=16384, help="matrices in the batched workloads")
    ap.add_argument("--waves", type=int, default=0, help="waves per workgroup of the batched kernel (0: auto)")
    ap.add_argument("--plain-sample", type=int, default=16, help="matrices timed through the plain path")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--m", type=int, default=0, help="rows of the plain workloads (0: the config's)")
    ap.add_argument("--n", type=int, default=0, help="columns of the plain workloads (0: the config's)")
    ap.add_argument("--lg", type=int, default=26, help="log2 batch size of the 2x2 workloads")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-tune", action="store_true", help="skip the batched plan's autotuner")
    ap.add_argument("--plan", default="", help="fixed batched plan instead of the autotuner, as "
                    "split_rows:lds_budget:w0,w1,w2:l1,l2 (waves per launch, lanes per pair of the short launches)")
    ap.add_argument("--order", type=int, default=0, help="pjs id of the batched ordering (4 or 6; 0: the plan's)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark measures one GPU only (asked for --gpus {args.gpus}); "
              "the batched SVD would shard by matrix with no communication", file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 3
    res = WORKLOADS[args.workload](args)
    line = {"metric": res.pop("metric"), "value": res.pop("value"), "unit": res.pop("unit"),
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": res.pop("ms_per_step"), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": res.pop("dtype"), "data": "synthetic (seeded Philox)"}
    line.update(res)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
