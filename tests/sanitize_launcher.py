"""Host-side AddressSanitizer + UndefinedBehaviorSanitizer run of the C-ABI
launcher (the library's host code: argument validation, dispatch, error
reporting).  Test tooling, not part of the library.

    python tests/sanitize_launcher.py          # build (2-3 min) + run

Builds libbslstm_tfhe.so with the host pass instrumented (`-Xarch_host
-fsanitize=address,undefined`; the gfx950 device code is compiled as usual)
into blind-spiking-lstm_amd/_build/sanitize/, then runs tests/test_abi.py
against that build in a fresh process with clang's ASan runtime preloaded.
No GPU is needed: every call those tests make returns from host validation
before any HIP call.  Exit status 0 = every test passed with no sanitizer
report.
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "blind-spiking-lstm_amd")
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build", "sanitize")
LIB = os.path.join(OUT, "libbslstm_tfhe_sanitized.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=all"]
FLAGS = ["--offload-arch=gfx950", "-O1", "-g", "-fPIC", "-std=c++20",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include")]


def asan_runtime() -> str:
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not found:
        raise RuntimeError("clang ASan runtime not found under /opt/rocm/llvm")
    return found[-1]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def up_to_date() -> bool:
    if not os.path.exists(LIB):
        return False
    t = os.path.getmtime(LIB)
    deps = sources() + glob.glob(os.path.join(CSRC, "*.h")) + \
        glob.glob(os.path.join(ROOT, "include", "*.h"))
    return all(os.path.getmtime(d) <= t for d in deps)


def build() -> str:
    if up_to_date():
        return LIB
    os.makedirs(OUT, exist_ok=True)

    def one(src):
        obj = os.path.join(OUT, os.path.basename(src)[:-4] + ".o")
        r = subprocess.run([HIPCC, *FLAGS, *SAN, "-c", src, "-o", obj], capture_output=True,
                           text=True)
        if r.returncode:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=6) as ex:
        objs = list(ex.map(one, sources()))
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *SAN[:4], "-o", LIB,
                        *objs], capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    return LIB


def run_abi_tests(lib: str) -> subprocess.CompletedProcess:
    """tests/test_abi.py against `lib`, in a fresh process under ASan/UBSan."""
    code = ("import sys; sys.path.insert(0, {root!r}); sys.path.insert(0, {pkg!r})\n"
            "from cuTFHE import _core; _core.LIB_PATH = {lib!r}\n"
            "import pytest\n"
            "sys.exit(pytest.main(['-q', '-m', 'not gpu', '-p', 'no:cacheprovider', "
            "{test!r}]))\n").format(root=ROOT, pkg=PKG, lib=lib,
                                    test=os.path.join(ROOT, "tests", "test_abi.py"))
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:detect_odr_violation=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True,
                          text=True, timeout=600)


def main() -> int:
    lib = build()
    r = run_abi_tests(lib)
    out = r.stdout + r.stderr
    print(out[-3000:])
    clean = "ERROR: AddressSanitizer" not in out and "runtime error" not in out
    print("launcher under ASan/UBSan:", "clean" if (r.returncode == 0 and clean) else "FAILED")
    return 0 if (r.returncode == 0 and clean) else 1


if __name__ == "__main__":
    sys.exit(main())
