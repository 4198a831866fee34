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
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[2 * i].record(stream)
        step()
        ev[2 * i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)]
    k_avg_s = float(np.mean(kms)) / 1e3
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
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_bjac2(kind, ins_np)
    return out


def cpu_bjac2(kind: str, ins_np):
    """Reference build when usable, else the oracle; bounded sample."""
    import ctypes as C
    os.environ.setdefault("OMP_NUM_THREADS", str(_cpu_threads()))
    from oracle import vjo
    use_ref = vjo.ref() is not None and _host_has_avx512()
    nsamp = min(ins_np[0].size, 1 << 22)
    sample = [np.ascontiguousarray(a[:nsamp]) for a in ins_np]
    reps, t_total = 0, 0.0
    while t_total < 10.0 and reps < 200:
        t0 = time.perf_counter()
        rc, _, _ = vjo.bjac2(kind, -nsamp, sample, use_ref=use_ref)
        t_total += time.perf_counter() - t0
        reps += 1
    return {"value": nsamp * reps / t_total, "unit": "EVDs/s",
            "cores": int(os.environ["OMP_NUM_THREADS"]) if use_ref else 1,
            "kind": "reference" if use_ref else "port",
            "sample": f"{reps} x {nsamp} problems ({'AVX-512 reference build, OpenMP' if use_ref else 'oracle C restatement, 1 thread'}), incl. staging copies"}


WORKLOADS = {
    "bjac2": lambda a: bench_bjac2(a, "d"),
    "zbjac2": lambda a: bench_bjac2(a, "z"),
    "sbjac2": lambda a: bench_bjac2(a, "s"),
    "cbjac2": lambda a: bench_bjac2(a, "c"),
}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="bjac2", choices=sorted(WORKLOADS))
    ap.add_argument("--lg", type=int, default=26, help="log2 batch size of the 2x2 workloads")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the batched SVD shards by matrix with no communication, see DESIGN.md", file=sys.stderr)
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
