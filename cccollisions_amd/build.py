"""In-tree build of the native parts of cccollisions_amd.

* ``libccc_amd.so``   -- HIP kernels + C ABI, compiled by hipcc for gfx950.
* ``_cccollisions*.so`` -- the PyTorch extension (pybind11 + ATen), compiled by g++ and
  linked against ``libccc_amd.so`` (rpath $ORIGIN).

Run ``python cccollisions_amd/build.py`` (or ``__graft_entry__.build()``); the script is
loaded by path so that it works before the package can be imported.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(ROOT, "_build", "obj")
LIB = os.path.join(HERE, "libccc_amd.so")
EXT_NAME = "_cccollisions"

HIP_SOURCES = ["ccc_context.hip", "ccc_bvh.hip", "ccc_query.hip", "ccc_sim.hip"]
ARCH = os.environ.get("CCC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # every expression is evaluated as written; fma() appears only where intended
    "-ffp-contract=off",
    # hardware fp64 atomic add on coarse-grained device memory
    "-munsafe-fp-atomics",
    # keep the compiler from forming packed fp32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    return r


def ext_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, EXT_NAME + suffix)


def build_hip(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_torch_ext(verbose: bool = False) -> str:
    import torch
    from torch.utils import cpp_extension as ce

    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    incs = ce.include_paths(device_type="cuda")
    out = ext_path()
    tmp = out + ".tmp"
    cmd = [
        os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared",
        os.path.join(CSRC, "ccc_torch.cpp"), "-o", tmp,
        f"-DTORCH_EXTENSION_NAME={EXT_NAME}", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
        f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}",
        "-I/opt/rocm/include", f"-I{INCLUDE}", f"-I{sysconfig.get_paths()['include']}",
        *[f"-I{i}" for i in incs],
        f"-L{HERE}", "-lccc_amd",
        f"-L{torch_lib}", "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python", "-lc10_hip", "-ltorch_hip",
        "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{torch_lib}",
    ]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build(verbose: bool = False) -> None:
    build_hip(verbose)
    build_torch_ext(verbose)


def clean() -> None:
    for p in (LIB, ext_path()):
        if os.path.exists(p):
            os.remove(p)
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, ext_path())
