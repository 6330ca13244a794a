"""Builds heffte_amd/libheffte_mi355.so (HIP kernels + C ABI) for gfx950 with hipcc.

Every source is compiled on each call (the build check recompiles even when a
library is present); objects go to heffte_amd/_build and the library is linked
in-tree next to this file.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libheffte_mi355.so")
ARCH = os.environ.get("HMI_OFFLOAD_ARCH", "gfx950")

SOURCES = ["hmi_kernels.hip", "hmi_mfma.hip", "hmi_scratch.hip", "hmi_plan.cpp", "hmi_batch.cpp", "hmi_capi.cpp"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def common_flags():
    return [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        # packed FP32 (v_pk_*_f32) instructions stay disabled for all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-command-line-argument",
        "-I" + os.path.join(HERE, "..", "include"),
    ]


# AddressSanitizer and UndefinedBehaviorSanitizer for the host code only (planner, C ABI, batch
# host logic): every -fsanitize= goes to the host compilation (-Xarch_host); device code is
# built as usual.
SAN_HOST_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
SAN_LIB = os.path.join(HERE, "libheffte_mi355_san.so")


def _compile(src, sanitize=False):
    bdir = os.path.join(BUILD, "san") if sanitize else BUILD
    obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + common_flags() + (SAN_HOST_FLAGS if sanitize else []) + \
        ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile of {src} failed:\n{r.stderr}")
    # the host pass warns that it does not know the packed-fp32 feature; drop that noise
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32" not in l and l.strip())
    return obj, warn


def build(verbose=False, sanitize=False):
    """Compiles and links the library; sanitize=True builds libheffte_mi355_san.so with the host
    code under ASan/UBSan instead (CPU tests only: load it with the sanitizer runtime preloaded)."""
    os.makedirs(os.path.join(BUILD, "san") if sanitize else BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(lambda src: _compile(src, sanitize), SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, w in results:
            if w:
                print(w, file=sys.stderr)
    target = SAN_LIB if sanitize else LIB
    tmp = target + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + \
        (["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libasan"]
         if sanitize else []) + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, target)
    return target


FRONTEND_TEST_SRC = os.path.join(HERE, "..", "tests", "cpp", "test_frontend.cpp")
FRONTEND_TEST_BIN = os.path.join(HERE, "..", "tests", "cpp", "test_frontend.bin")


def build_frontend_test():
    """Compiles the C++ front-end test program (include/heffte_mi355.hpp) against the library."""
    cmd = [hipcc(), "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-DHEFFTE_MI355_NO_MPI",
           "-I" + os.path.join(HERE, "..", "include"), FRONTEND_TEST_SRC, "-o", FRONTEND_TEST_BIN,
           "-L" + HERE, "-lheffte_mi355", "-Wl,-rpath,$ORIGIN/../../heffte_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"front-end test build failed:\n{r.stderr}")
    return FRONTEND_TEST_BIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_frontend_test())
