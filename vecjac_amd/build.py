"""Build recipe for libvecjac_hip.so (gfx950) and the CPU oracle.

    python -m vecjac_amd.build          # HIP library + oracle (+ reference build)

The HIP library is compiled with hipcc straight from vecjac_amd/csrc into
the package directory, so the .so travels with the source tree.  Two
libraries are linked from it: libvecjac_hip.so, whose reference interfaces
take 64-bit Fortran integers (the reference's MKL_ILP64 default), and
libvecjac_hip_lp64.so, the C-ABI sources compiled again with -DVJ_LP64
(32-bit integers, the reference's LP64 build) around the same kernels.  Floating
point contraction is off (every fused multiply-add in the kernels is written
out) and packed FP32 instructions are disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libvecjac_hip.so"
LIB_LP64 = PKG / "libvecjac_hip_lp64.so"
# the sources whose entry points take the reference's integer types
ABI_SRCS = ("capi_batched.hip", "capi_bjac2.cpp", "capi_nssvj.hip", "capi_svd.hip", "svd_bwd.hip", "zsplit.hip")
DRIVERS = PKG / "drivers"
BIN = PKG / "bin"
ARCH = os.environ.get("VJ_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]


def _compile(src: Path, lp64: bool = False) -> Path:
    obj = BUILD / (("lp64_" if lp64 else "") + src.name + ".o")
    cmd = [HIPCC, *COMMON, *(["-DVJ_LP64"] if lp64 else []), "-c", str(src), "-o", str(obj)]
    if src.suffix == ".hip":
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _link(objs, lib: Path) -> None:
    tmp = lib.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)


def build_hip(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = [(s, False) for s in srcs] + [(CSRC / n, True) for n in ABI_SRCS]
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    _link(objs[:len(srcs)], LIB)
    lp = {o.name[len("lp64_"):]: o for o in objs[len(srcs):]}
    _link([lp.get(o.name, o) for o in objs[:len(srcs)]], LIB_LP64)
    if verbose:
        print(f"built {LIB} and {LIB_LP64}")
    return LIB


def build_drivers(verbose: bool = True) -> list:
    """The command-line drivers (drivers/*.c, the reference's tdvjsvd /
    tzvjsvd / tdbjac2 interfaces) into bin/, linked against the library."""
    BIN.mkdir(exist_ok=True)
    out = []
    for src in sorted(DRIVERS.glob("*.c")):
        exe = BIN / src.stem
        cmd = ["gcc", "-O2", "-std=gnu11", "-Wall", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
               f"-L{LIB.parent}", "-lvecjac_hip", "-Wl,-rpath,$ORIGIN/..", "-lm"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        out.append(exe)
    if verbose:
        print(f"built {len(out)} drivers in {BIN}")
    return out


def build_oracle(verbose: bool = True) -> None:
    r = subprocess.run(["make", "-C", str(ROOT / "oracle"), "-j8"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built oracle")


def main() -> int:
    build_hip()
    build_drivers()
    build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
