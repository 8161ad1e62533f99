"""Build the native library ``mentpy_amd/lib/libmbqc_hip.so`` for gfx950.

One ``hipcc`` invocation compiles the HIP kernels and the host plan compiler
into a shared library with a plain C ABI (declared in ``include/mbqc_hip.h``).
Packed FP32 instructions are disabled for all device code.

``--sanitize`` builds ``mentpy_amd/lib/libmbqc_hip_sanitize.so`` instead: the
same sources with AddressSanitizer and UndefinedBehaviorSanitizer on the host
code only (plan compiler, causal flow and the entry points' argument
checks; device code is built as usual), and then runs the CPU test suite
(``pytest -m "not gpu"``, which drives the plan compiler, the causal flow and
the entry points' validation through the C ABI) against it, with the
sanitizer runtime preloaded and any report fatal.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
ROOT = PKG.parent
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libmbqc_hip.so"
SANITIZE_LIB = LIB_DIR / "libmbqc_hip_sanitize.so"
SOURCES = [HERE / "kernels.hip", HERE / "plan.cpp", HERE / "flow.cpp"]
ARCH = os.environ.get("MBQC_OFFLOAD_ARCH", "gfx950")
SANITIZE_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host",
                  "-fno-sanitize-recover=undefined", "-shared-libasan"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    return "hipcc"


def _compile(target: Path, extra, opt: str, verbose: bool) -> Path:
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    tmp = target.with_suffix(".so.tmp")
    cmd = [
        hipcc(), f"--offload-arch={ARCH}", opt, "-std=c++17", "-shared", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument", *extra,
        "-I", str(ROOT / "include"),
        *[str(s) for s in SOURCES], "-o", str(tmp),
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")
    os.replace(tmp, target)
    return target


def build(verbose: bool = False) -> Path:
    return _compile(LIB, [], "-O3", verbose)


def asan_runtime() -> str:
    res = subprocess.run([hipcc(), "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True)
    path = res.stdout.strip()
    if res.returncode != 0 or not Path(path).exists():
        raise RuntimeError("the compiler's AddressSanitizer runtime was not found")
    return path


def build_sanitized(verbose: bool = False, run_tests: bool = True) -> int:
    """Build the host-sanitized library and run the CPU tests on it; returns
    the test run's exit status."""
    _compile(SANITIZE_LIB, SANITIZE_FLAGS, "-O1", verbose)
    if not run_tests:
        return 0
    env = dict(os.environ)
    env.update({
        "MENTPY_AMD_LIB": str(SANITIZE_LIB),
        "LD_PRELOAD": asan_runtime(),
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
    })
    cmd = [sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
           str(ROOT / "tests")]
    if verbose:
        print("sanitized run: " + " ".join(cmd), flush=True)
    return subprocess.run(cmd, env=env, cwd=str(ROOT)).returncode


if __name__ == "__main__":
    if "--sanitize" in sys.argv[1:]:
        sys.exit(build_sanitized(verbose=True))
    print(build(verbose=True))
