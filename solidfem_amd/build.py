"""Build recipe for the native parts of the package (in-tree, gfx950 only).

    python -m solidfem_amd.build          # libsfem_hip.so (+ pysolidfem module)
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["sfem_kernels.hip", "sfem_gnn.hip", "sfem_mg.hip", "sfem_noalloc.hip", "sfem_capi.cpp", "sfem_gnn_host.cpp",
               "sfem_mg.cpp", "sfem_host.cpp"]
LIB = os.path.join(HERE, "libsfem_hip.so")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib():
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", f"-I{CSRC}", *srcs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_pysolidfem():
    """pybind11 module `pysolidfem` (host C++ only, links libsfem_hip.so)."""
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "pysolidfem" + suffix)
    host = os.path.join(CSRC, "host")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
           f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}", f"-I{INCLUDE}", f"-I{host}",
           os.path.join(host, "FemPhysicsMatrixFree.cpp"), os.path.join(host, "FemIO.cpp"), os.path.join(host, "FemBody.cpp"),
           os.path.join(host, "pysolidfem.cpp"),
           f"-L{HERE}", "-lsfem_hip", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_lib()
    build_pysolidfem()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
