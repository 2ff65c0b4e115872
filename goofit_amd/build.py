"""Build recipe of the native libraries, in-tree (outputs under goofit_amd/lib/).

* libgoofit_hip.so -- the gfx950 kernels and the C ABI of include/goofit_hip.h
  (hipcc, one object per source, compiled in parallel).
* libgoofit.so      -- the GooFit C++ API (GooPdf, AddPdf, Amp3Body, FitManager, ...)
  on top of libgoofit_hip.so (g++).
* _goofit*.so       -- the Python binding of that API (pybind11, g++).

Usage: ``python -m goofit_amd.build`` or ``goofit_amd.build.build_all()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GOOFIT_AMD_ARCH", "gfx950")
HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-parameter"]
HIP_SOURCES = ["gf_kernels.hip", "gf_mlp.hip", "gf_capi.hip"]

CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    objs, cmds = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    so = os.path.join(LIB, "libgoofit_hip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", so + ".tmp", *objs])
    os.replace(so + ".tmp", so)
    return so


def host_sources():
    if not os.path.isdir(HOST):
        return []
    return sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))


def build_host(verbose=False):
    srcs = host_sources()
    if not srcs:
        return None
    os.makedirs(OBJ, exist_ok=True)
    objs, cmds = [], []
    for src in srcs:
        obj = os.path.join(OBJ, "host_" + os.path.basename(src).replace(".cpp", ".o"))
        objs.append(obj)
        cmds.append([CXX, *CXX_FLAGS, "-I", INCLUDE, "-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    so = os.path.join(LIB, "libgoofit.so")
    _run([CXX, "-shared", "-fPIC", "-o", so + ".tmp", *objs, f"-L{LIB}", "-lgoofit_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(so + ".tmp", so)
    return so


def build_python(verbose=False):
    src = os.path.join(HERE, "bindings", "goofit_py.cpp")
    if not os.path.exists(src):
        return None
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    so = os.path.join(HERE, "_goofit" + ext)
    py_inc = sysconfig.get_paths()["include"]
    _run([CXX, *CXX_FLAGS, "-shared", "-I", INCLUDE, "-I", pybind11.get_include(), "-I", py_inc, src,
          "-o", so + ".tmp", f"-L{LIB}", "-lgoofit", "-lgoofit_hip", "-Wl,-rpath,$ORIGIN/lib"])
    os.replace(so + ".tmp", so)
    return so


def build_all(verbose=False):
    out = [build_hip(verbose)]
    out.append(build_host(verbose))
    out.append(build_python(verbose))
    return [o for o in out if o]


if __name__ == "__main__":
    for p in build_all(verbose="-v" in sys.argv):
        print("built", p)
