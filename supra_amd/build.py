"""Build the HIP beamformer library in-tree (supra_amd/libsupra_bf.so).

Host C++ sources are compiled with g++ and the .hip sources with hipcc for
gfx950, then linked with hipcc.  Floating-point contraction is disabled
everywhere: the kernels reproduce the host restatement of the reference's
single-precision arithmetic bit for bit, which an implicit FMA would break.
Packed FP32 instructions are disabled for the device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libsupra_bf.so"
SHIM_TEST = PKG / "shim_test"
TORCH_LIB = PKG / "libsupra_torch.so"

ARCH = os.environ.get("SUPRA_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["geometry.cpp", "bf_host.cpp", "status.cpp", "mock_io.cpp"]
HIP_SOURCES = ["bf_context.hip", "das_kernels.hip", "das_adjoint.hip", "mv_kernels.hip", "synth.hip", "batch.hip",
               "envelope.hip", "das_variants.hip", "scan_converter.hip"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP beamformer cannot be built")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    """Compile every source and link libsupra_bf.so; returns its path."""
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        jobs.append(["g++", *COMMON, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        jobs.append([hipcc, *COMMON, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, LIB)
    # the C++ engine shim's test program (include/supra_bf.hpp), next to the library
    _run([hipcc, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "shim_test.cpp"),
          f"-L{PKG}", "-lsupra_bf", "-Wl,-rpath,$ORIGIN", "-o", str(SHIM_TEST)], verbose)
    build_torch_op(hipcc, verbose)
    return LIB


def build_torch_op(hipcc: str, verbose: bool) -> None:
    """libsupra_torch.so: the libtorch autograd operator supra_bf::das_c
    (include/supra_bf_torch.hpp, csrc/torch_op.cpp), host code against the
    installed torch's headers and libraries."""
    import torch
    tdir = Path(torch.__file__).resolve().parent
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    tmp = TORCH_LIB.with_suffix(".so.tmp")
    _run([hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
          f"-D_GLIBCXX_USE_CXX11_ABI={abi}", f"-I{tdir / 'include'}", f"-I{tdir / 'include' / 'torch' / 'csrc' / 'api' / 'include'}",
          f"-I{ROOT / 'include'}", str(CSRC / "torch_op.cpp"), f"-L{tdir / 'lib'}", "-ltorch", "-ltorch_cpu", "-lc10",
          "-lc10_hip", "-ltorch_hip", f"-L{PKG}", "-lsupra_bf", f"-Wl,-rpath,{tdir / 'lib'}", "-Wl,-rpath,$ORIGIN",
          "-o", str(tmp)], verbose)
    os.replace(tmp, TORCH_LIB)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
