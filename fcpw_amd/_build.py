"""Build recipe for the native parts of fcpw_amd (run by __graft_entry__.build()).

  lib/libfcpw_hip.so          HIP kernels + the C ABI (include/fcpw_hip.h), gfx950
  _fcpw_amd<ext-suffix>.so    pybind11 host module (Scene, GPUScene), linked to it

Everything is compiled in-tree so the built objects travel with the source.

``python fcpw_amd/_build.py --sanitize`` builds the host code with
AddressSanitizer and UndefinedBehaviorSanitizer instead, for the CPU test
suite: the pybind11 module (the host BVH, silhouette and instance builders and
the GPUScene front end) and the oracle, staged with copies of the Python
sources under build/sanitize/ so the in-tree build is left alone.  The HIP
library is linked unsanitized (its device code cannot be, and the CPU tests
only reach its argument checks).  It prints the command that runs the CPU
tests against the staged copy.
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
# -Wno-pass-failed: the eight-waves occupancy hint cannot be met by the kernel
# instances whose LDS stack window is too large for eight waves per SIMD; the
# compiler then keeps its own register budget, which is the intended outcome.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-pass-failed",
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


SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined"]


def build_sanitized():
    """ASan + UBSan builds of the host module and the oracle, staged under build/sanitize/."""
    import pybind11
    import shutil
    stage = os.path.join(ROOT, "build", "sanitize")
    if os.path.isdir(stage):
        shutil.rmtree(stage)
    ignore = shutil.ignore_patterns("__pycache__", "*.so", "*.o", "_san", "obj")
    shutil.copytree(HERE, os.path.join(stage, "fcpw_amd"), ignore=ignore)
    shutil.copytree(os.path.join(ROOT, "oracle"), os.path.join(stage, "oracle"), ignore=ignore)
    shutil.copytree(os.path.join(ROOT, "tests"), os.path.join(stage, "tests"), ignore=ignore)
    shutil.copytree(os.path.join(ROOT, "include"), os.path.join(stage, "include"))
    for f in ("pytest.ini", "__graft_entry__.py"):
        shutil.copy(os.path.join(ROOT, f), stage)
    if not os.path.exists(HIP_LIB):
        build_hip()
    os.makedirs(os.path.join(stage, "fcpw_amd", "lib"), exist_ok=True)
    shutil.copy(HIP_LIB, os.path.join(stage, "fcpw_amd", "lib"))
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", *SAN_FLAGS, "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden", "-ffp-contract=off",
          f"-I{py_inc}", f"-I{pybind11.get_include()}", os.path.join(HERE, "csrc", "py", "bindings.cpp"),
          f"-L{os.path.dirname(HIP_LIB)}", "-lfcpw_hip", "-Wl,-rpath,$ORIGIN/lib",
          "-o", os.path.join(stage, "fcpw_amd", os.path.basename(EXT))])
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "SANITIZE=1"])
    shutil.copy(os.path.join(ROOT, "oracle", "_san", "liboracle.so"), os.path.join(stage, "oracle", "liboracle.so"))
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    cmd = (f"cd {stage} && LD_PRELOAD={asan}:{ubsan} ASAN_OPTIONS=detect_leaks=0:detect_odr_violation=0 "
           f"UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 python -m pytest tests -m 'not gpu' -x -q "
           f"-p no:cacheprovider")
    print(cmd)
    return cmd


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        build_sanitized()
    else:
        build_all()
    sys.exit(0)
