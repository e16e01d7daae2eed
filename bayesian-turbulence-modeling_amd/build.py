"""Builds the native parts of the package in place.

  librsg.so                      HIP engine for gfx950 (csrc/rsg_capi.hip, csrc/rsg_team.hip), C-ABI of include/rsg.h
  rans_sst_py*.so                pybind11 host module (host/*.cpp) linked against librsg.so

Run as a script or call build_all().  Both artefacts are written next to this
file; nothing is installed.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# gfx950 only; packed FP32 instructions disabled; no FMA contraction so the
# FP64 assembly rounds exactly like the reference solver's
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-function"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_engine():
    out = os.path.join(HERE, "librsg.so")
    srcs = [os.path.join(HERE, "csrc", f) for f in ("rsg_capi.hip", "rsg_team.hip")]
    _run([HIPCC] + HIP_FLAGS + ["-I", INCLUDE, "-I", os.path.join(HERE, "csrc")] + srcs +
         ["-o", out + ".tmp", "-lpthread"])
    os.replace(out + ".tmp", out)
    return out


def build_module():
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "rans_sst_py" + suffix)
    host = os.path.join(HERE, "host")
    srcs = [os.path.join(host, f) for f in ("bindings.cpp", "forward_model.cpp", "sensitivity.cpp", "model.cpp",
                                            "mesh.cpp")]
    py_inc = sysconfig.get_paths()["include"]
    cmd = ["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC",
           "-I", pybind11.get_include(), "-I", py_inc, "-I", INCLUDE] + srcs + \
          ["-L", HERE, "-lrsg", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_all():
    # the module links against librsg.so, so the engine goes first
    build_engine()
    build_module()


if __name__ == "__main__":
    build_all()
    print("built", os.listdir(HERE))
    sys.exit(0)
