"""Builds the native parts in-tree:

  egg_sync/libeggsync.so   HIP kernels + C ABI (include/egg_sync.h), hipcc for gfx950
  egg_sync/_egg_sync_C.so  PyTorch extension exposing the reference's binding names
                           (src/maps/map.cpp:35-44), linked against libeggsync.so

Flags: -ffp-contract=off keeps every FP32 product and sum separately rounded
(the membership tests must equal the CPU oracle bit for bit); packed FP32
instructions are disabled.
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                      # egg-sync_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "libeggsync.so")
EXT = os.path.join(PKG, "_egg_sync_C.so")
HIP_SOURCES = ["egg_scan_sort.hip", "egg_grid.hip", "egg_update.hip", "egg_labels.hip",
               "egg_solver.hip", "egg_mlp.hip", "egg_brute.hip"]
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
               "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({' '.join(cmd[:3])} ...):\n{out}")
    return out


def build_lib(verbose: bool = False) -> str:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objs = []
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        jobs.append([hipcc, *HIPCC_FLAGS, "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the sources compile independently: one hipcc per source, side by side
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 8))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(_run, cmd) for cmd in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc, *HIPCC_FLAGS, "-shared", "-Wl,-soname,libeggsync.so", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_ext(verbose: bool = False) -> str:
    import ctypes

    import torch.utils.cpp_extension as ce
    ctypes.CDLL(LIB)  # load() imports the module; its libeggsync.so dependency resolves by soname
    bdir = os.path.join(PKG, "_build", "ext")
    os.makedirs(bdir, exist_ok=True)
    ce.load(name="_egg_sync_C", sources=[os.path.join(CSRC, "torch_ext.cpp")],
            extra_include_paths=[INCLUDE], extra_cflags=["-O2", "-std=c++17"],
            extra_ldflags=[f"-L{PKG}", "-leggsync", "-Wl,-rpath,$ORIGIN"],
            build_directory=bdir, verbose=verbose, is_python_module=True)
    built = glob.glob(os.path.join(bdir, "_egg_sync_C*.so"))
    if not built:
        raise RuntimeError("torch extension build produced no library")
    tmp = EXT + ".tmp"
    shutil.copyfile(built[0], tmp)
    os.replace(tmp, EXT)
    if verbose:
        print("built", EXT)
    return EXT


def build(verbose: bool = False) -> None:
    build_lib(verbose)
    build_ext(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
