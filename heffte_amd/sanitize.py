"""Runs the CPU tests of the host code under AddressSanitizer + UndefinedBehaviorSanitizer.

    python heffte_amd/sanitize.py

1. builds heffte_amd/libheffte_mi355_san.so (host code of the planner, the C ABI and the batch
   plan instrumented; device code as usual) and oracle/liboracle_san.so;
2. runs tests/test_capi.py (plan creation, geometry, argument and aliasing checks through the C
   ABI, no GPU needed) against the sanitized library, with clang's ASan runtime preloaded;
3. runs tests/test_oracle.py against the sanitized oracle, with gcc's ASan/UBSan runtimes
   preloaded.
Any sanitizer report ends its run with a non-zero status.  Leak checking is off: the Python
interpreter itself keeps allocations alive until exit.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _clang_asan_runtime():
    hits = glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not hits:
        raise RuntimeError("clang's ASan runtime (libclang_rt.asan-x86_64.so) not found under /opt/rocm/lib/llvm")
    return hits[0]


def _gcc_runtime(name):
    return subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True,
                          check=True).stdout.strip()


def main():
    sys.path.insert(0, HERE)
    import build as hb
    lib = hb.build(sanitize=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    common = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
                  UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    runs = [
        ("C ABI", dict(common, LD_PRELOAD=_clang_asan_runtime(), HEFFTE_MI355_LIBRARY=lib),
         ["tests/test_capi.py"]),
        ("oracle", dict(common, LD_PRELOAD=_gcc_runtime("libasan.so") + ":" + _gcc_runtime("libubsan.so"),
                        HEFFTE_ORACLE_LIBRARY=os.path.join(ROOT, "oracle", "liboracle_san.so")),
         ["tests/test_oracle.py"]),
    ]
    status = 0
    for name, env, tests in runs:
        print(f"== {name} under ASan/UBSan: {' '.join(tests)}", flush=True)
        r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider"] + tests,
                           cwd=ROOT, env=env)
        status = status or r.returncode
    return status


if __name__ == "__main__":
    sys.exit(main())
