"""Builds the native parts of the package in place.

  librsg.so                      HIP engine for gfx950 (csrc/*.hip), C-ABI of include/rsg.h
  rans_sst_py*.so                pybind11 host module (host/*.cpp) linked against librsg.so

Run as a script or call build_all().  Both artefacts are written next to this
file (objects under _build/); nothing is installed.
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


BUILD = os.path.join(HERE, "_build")
# (source, object, extra flags): rsg_chunk.hip once per workgroup size and precision
HIP_UNITS = [("rsg_capi.hip", "rsg_capi", []), ("rsg_team.hip", "rsg_team", []), ("rsg_big.hip", "rsg_big", [])] + \
    [("rsg_chunk.hip", "rsg_chunk_%d_%d" % (bs, prec), ["-DRSG_CHUNK_BS=%d" % bs, "-DRSG_CHUNK_PREC=%d" % prec])
     for bs in (512, 256) for prec in (0, 1)]
HOST_SOURCES = ("bindings.cpp", "forward_model.cpp", "sensitivity.cpp", "model.cpp", "mesh.cpp", "compressible.cpp")


def _compile_hip(unit):
    src, name, extra = unit
    obj = os.path.join(BUILD, name + ".o")
    _run([HIPCC] + [f for f in HIP_FLAGS if f != "-shared"] + extra +
         ["-I", INCLUDE, "-I", os.path.join(HERE, "csrc"), "-c", os.path.join(HERE, "csrc", src), "-o", obj])
    return obj


def _compile_host(src):
    import pybind11
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-fPIC", "-I", pybind11.get_include(),
          "-I", sysconfig.get_paths()["include"], "-I", INCLUDE, "-c", src, "-o", obj])
    return obj


def build_all():
    """Every translation unit compiled in parallel (the member kernels of each
    workgroup size are a unit of their own), then librsg.so linked, then the
    module linked against it."""
    os.makedirs(BUILD, exist_ok=True)
    host = [os.path.join(HERE, "host", f) for f in HOST_SOURCES]
    with ThreadPoolExecutor(max_workers=8) as ex:
        hip_f = [ex.submit(_compile_hip, u) for u in HIP_UNITS]
        host_f = [ex.submit(_compile_host, f) for f in host]
        hip_o = [f.result() for f in hip_f]
        host_o = [f.result() for f in host_f]
    eng = os.path.join(HERE, "librsg.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + hip_o + ["-o", eng + ".tmp", "-lpthread"])
    os.replace(eng + ".tmp", eng)
    mod = os.path.join(HERE, "rans_sst_py" + sysconfig.get_config_var("EXT_SUFFIX"))
    _run(["g++", "-shared", "-fPIC"] + host_o + ["-L", HERE, "-lrsg", "-Wl,-rpath,$ORIGIN", "-o", mod + ".tmp"])
    os.replace(mod + ".tmp", mod)
    return eng, mod


if __name__ == "__main__":
    build_all()
    print("built", os.listdir(HERE))
    sys.exit(0)
