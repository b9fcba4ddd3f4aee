"""Builds sseqs_amd/libsseqs_amd.so from csrc/*.hip (hipcc, gfx950) and the
host-only csrc/*.cpp (the host compiler; torch_ext.cpp is the separate
PyTorch extension).

Run as ``python -m sseqs_amd.build`` (or through ``__graft_entry__.build()``).
Objects are compiled in parallel into ``sseqs_amd/_build/`` and linked into
one shared library next to this file; the library exports exactly the
functions declared in ``include/sseqs_amd.h``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libsseqs_amd.so"
BUILD = PKG / "_build"
ARCH = os.environ.get("SSEQS_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
    # packed FP32 VALU ops stay disabled in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


# host-only sources of the library (everything but the PyTorch extension)
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__=1",
              "-I/opt/rocm/include", f"-I{ROOT / 'include'}"]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    if src.suffix == ".cpp":
        cmd = [os.environ.get("CXX", "g++"), *HOST_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    return obj


EXT_SRC = CSRC / "torch_ext.cpp"
EXT_OUT = PKG / "_torch_ext.so"


def build_torch_ext(verbose: bool = True) -> Path:
    """The PyTorch extension module (xbit, conv, sw over the C ABI), built
    in-tree with the host compiler against torch's headers and libraries;
    it finds libsseqs_amd.so next to itself through its run path."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-shared", "-fPIC", "-std=c++17",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_EXTENSION_NAME=_torch_ext",
           "-DTORCH_API_INCLUDE_EXTENSION_H", "-DUSE_ROCM=1", "-D__HIP_PLATFORM_AMD__=1",
           *[f"-I{p}" for p in ce.include_paths()], f"-I{sysconfig.get_paths()['include']}",
           "-I/opt/rocm/include", f"-I{ROOT / 'include'}", str(EXT_SRC),
           *[f"-L{p}" for p in ce.library_paths()], f"-L{PKG}", "-lsseqs_amd",
           "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python",
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN",
           *[f"-Wl,-rpath,{p}" for p in ce.library_paths()],
           "-o", str(EXT_OUT.with_suffix(".so.tmp"))]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"torch extension build failed:\n{r.stderr[-4000:]}")
    os.replace(EXT_OUT.with_suffix(".so.tmp"), EXT_OUT)
    if verbose:
        print(f"built {EXT_OUT.relative_to(ROOT)}")
    return EXT_OUT


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    if not srcs:
        raise RuntimeError("no HIP sources found")
    srcs += sorted(p for p in CSRC.glob("*.cpp") if p != EXT_SRC)
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT.relative_to(ROOT)} from {len(srcs)} sources")
    build_torch_ext(verbose)
    return OUT


if __name__ == "__main__":
    build()
