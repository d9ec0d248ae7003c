"""Build recipe for the native parts of fcpw_amd (run by __graft_entry__.build()).

  lib/libfcpw_hip.so          HIP kernels + the C ABI (include/fcpw_hip.h), gfx950
  _fcpw_amd<ext-suffix>.so    pybind11 host module (Scene, GPUScene), linked to it

Everything is compiled in-tree so the built objects travel with the source.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("FCPW_HIP_ARCH", "gfx950")

HIP_SOURCES = [os.path.join(HERE, "csrc", "hip", "fcpw_hip.hip")]
HIP_LIB = os.path.join(HERE, "lib", "libfcpw_hip.so")
EXT = os.path.join(HERE, "_fcpw_amd" + sysconfig.get_config_var("EXT_SUFFIX"))

# -ffp-contract=off: every product and sum rounded separately, so the kernels
# reproduce the reference arithmetic step for step (and match the oracle bit
# for bit).  -packed-fp32-ops: no packed FP32 instructions are generated.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             f"--offload-arch={ARCH}"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    os.makedirs(os.path.dirname(HIP_LIB), exist_ok=True)
    tmp = HIP_LIB + ".tmp"
    _run(["hipcc", *HIP_FLAGS, "-o", tmp, *HIP_SOURCES])
    os.replace(tmp, HIP_LIB)


def build_ext():
    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    tmp = EXT + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden", "-ffp-contract=off",
          f"-I{py_inc}", f"-I{pybind11.get_include()}",
          os.path.join(HERE, "csrc", "py", "bindings.cpp"),
          f"-L{os.path.dirname(HIP_LIB)}", "-lfcpw_hip", "-Wl,-rpath,$ORIGIN/lib",
          "-o", tmp])
    os.replace(tmp, EXT)


def build_all():
    build_hip()
    build_ext()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
