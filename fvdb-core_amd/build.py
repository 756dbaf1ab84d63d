"""Builds the two native libraries of the package in-tree, next to the Python code.

* ``fvdb/libfvdbk.so``  - the gfx950 HIP kernels behind the C ABI in
  ``include/fvdbk.h`` (hipcc, no torch dependency).
* ``fvdb/_fvdbk_ext*.so`` - the PyTorch-ROCm extension (g++) that calls the C ABI.

Usage: ``python fvdb-core_amd/build.py`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
PKG = os.path.join(HERE, "fvdb")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("FVDBK_ARCH", "gfx950")

HIP_SOURCES = ["fvdbk_index.hip", "fvdbk_conv.hip", "fvdbk_conv_tile_bf16.hip", "fvdbk_conv_tile_f16.hip",
               "fvdbk_conv_tile_f32.hip", "fvdbk_pool.hip", "fvdbk_norm.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-mcode-object-version=5",
    # packed fp32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops' is not a recognized" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {cmd[0]}")


def ext_filename() -> str:
    return "_fvdbk_ext" + sysconfig.get_config_var("EXT_SUFFIX")


def build_hip() -> str:
    """Compiles each HIP source to an object (in parallel), then links the library."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    out = os.path.join(PKG, "libfvdbk.so")
    tmp = out + ".tmp"
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    with tempfile.TemporaryDirectory(prefix="fvdbk_build_") as td:
        objs = [os.path.join(td, os.path.splitext(src)[0] + ".o") for src in HIP_SOURCES]
        cmds = [[hipcc, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", os.path.join(CSRC, src), "-o", obj]
                for src, obj in zip(HIP_SOURCES, objs)]
        jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with ThreadPoolExecutor(jobs) as pool:
            for f in [pool.submit(_run, c) for c in cmds]:
                f.result()
        _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_ext() -> str:
    import torch
    from torch.utils import cpp_extension

    out = os.path.join(PKG, ext_filename())
    tmp = out + ".tmp"
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-w",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_API_INCLUDE_EXTENSION_H", "-DTORCH_EXTENSION_NAME=_fvdbk_ext",
           "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
           f"-I{INCLUDE}", f"-I{os.path.join(ROCM, 'include')}", f"-I{sysconfig.get_paths()['include']}"]
    cmd += [f"-I{p}" for p in cpp_extension.include_paths()]
    cmd += [os.path.join(CSRC, "ext.cpp"), "-o", tmp]
    libdirs = cpp_extension.library_paths()
    cmd += [f"-L{p}" for p in libdirs] + [f"-L{PKG}"]
    cmd += ["-lfvdbk", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python"]
    cmd += ["-Wl,-rpath,$ORIGIN"] + [f"-Wl,-rpath,{p}" for p in libdirs]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_hip()
    build_ext()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"hip": build_hip, "ext": build_ext, "all": build_all}[what]()
