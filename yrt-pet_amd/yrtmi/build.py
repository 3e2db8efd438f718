"""Build the yrtmi native library in-tree with hipcc for gfx950.

    python -m yrtmi.build          (from yrt-pet_amd/)

Produces ``yrtmi/libyrtmi.so`` (HIP kernels + C ABI) and the C++ host layer
with its pybind11 module ``_pyyrtmi`` (``yrt-pet_amd/_pyyrtmi*.so``, imported
through ``pyyrtmi.py``; host code
only, linked against libyrtmi.so).  Everything is rebuilt when this runs;
loaders never build on their own.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)            # yrt-pet_amd/
REPO = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-munsafe-fp-atomics",
    # packed FP32 instructions are disabled on this pool
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

LIB_SOURCES = ["yrtmi_kernels.hip", "yrtmi_tiles.hip", "yrtmi_f64.hip", "yrtmi_order.hip",
               "yrtmi_image.hip", "yrtmi_capi.hip"]


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def _run(cmd):
    r = subprocess.run(cmd, cwd=CSRC, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({' '.join(cmd[:3])} ...):\n{out}")
    return out


def build_library(verbose=False):
    out = os.path.join(HERE, "libyrtmi.so")
    tmp = out + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-shared",
           "-o", tmp, *LIB_SOURCES]
    log = _run(cmd)
    os.replace(tmp, out)
    if verbose and log.strip():
        print(log)
    return out


HOST_SOURCES = ["host/yrtmi_host.cpp", "host/yrtmi_io.cpp", "host/pyyrtmi.cpp",
                "host/pyyrtmi_io.cpp"]


def build_host_module(verbose=False):
    """The C++ host layer + pybind11 bindings (g++, no device code)."""
    import sysconfig
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG_ROOT, "_pyyrtmi" + suffix)
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-shared", "-fPIC",
           "-fvisibility=hidden", "-Wall", "-Wno-unused-function",
           f"-I{INCLUDE}", f"-I{os.path.join(CSRC, 'host')}",
           f"-I{pybind11.get_include()}", f"-I{py_inc}",
           *HOST_SOURCES, f"-L{HERE}", "-lyrtmi", "-lz", "-Wl,-rpath,$ORIGIN/yrtmi",
           "-o", tmp]
    log = _run(cmd)
    os.replace(tmp, out)
    if verbose and log.strip():
        print(log)
    return out


def main():
    print(build_library(verbose=True))
    print(build_host_module(verbose=True))


if __name__ == "__main__":
    sys.exit(main())
