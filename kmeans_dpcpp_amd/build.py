"""In-tree build of the gfx950 library and its pybind11 module.

Outputs (next to this file):
  libkmeans_dpcpp_amd.so          HIP kernels + host driver + C ABI
  _kmeans_dpcpp_amd<EXT_SUFFIX>   pybind11 boundary (plain C++, links the above)

Usage: python kmeans_dpcpp_amd/build.py   (or __graft_entry__.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIP_SOURCES = ["util.hip", "sort.hip", "lloyd.hip", "lloyd_fast.hip", "lloyd_f64.hip",
               "lloyd_mixed.hip", "lloyd_tiled.hip", "batched.hip", "seeding.hip", "noalloc.hip",
               "driver.hip", "capi.hip"]
LIB_NAME = "libkmeans_dpcpp_amd.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-munsafe-fp-atomics",
    # no packed FP32 (v_pk_*_f32) instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    path = os.path.join(ROCM, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): "
                           f"{' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


def _compile_hip(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([_hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj],
         verbose)
    return obj


def pybind_module_path() -> str:
    return os.path.join(HERE, "_kmeans_dpcpp_amd" +
                        sysconfig.get_config_var("EXT_SUFFIX"))


def library_path() -> str:
    return os.path.join(HERE, LIB_NAME)


def build(verbose: bool = False, jobs: int | None = None) -> None:
    """Compile every HIP source for gfx950, link the C-ABI library, then
    compile and link the pybind11 module against it."""
    import pybind11

    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(HIP_SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile_hip(s, verbose), HIP_SOURCES))
    tmp_lib = os.path.join(BUILD, LIB_NAME)
    _run([_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC",
          f"-Wl,-soname,{LIB_NAME}", *objs, "-o", tmp_lib], verbose)
    shutil.copyfile(tmp_lib, library_path() + ".tmp")
    os.replace(library_path() + ".tmp", library_path())

    py_inc = sysconfig.get_paths()["include"]
    tmp_mod = os.path.join(BUILD, os.path.basename(pybind_module_path()))
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
          f"-I{py_inc}", f"-I{pybind11.get_include()}", f"-I{INCLUDE}",
          os.path.join(CSRC, "pybind_module.cpp"),
          f"-L{HERE}", "-lkmeans_dpcpp_amd", "-Wl,-rpath,$ORIGIN",
          "-o", tmp_mod], verbose)
    shutil.copyfile(tmp_mod, pybind_module_path() + ".tmp")
    os.replace(pybind_module_path() + ".tmp", pybind_module_path())


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", library_path(), pybind_module_path())
