"""Runs the CPU tests of the library's host code and of the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (build container only; no GPU is used).

    python tests/run_sanitized.py

Builds oracle/_build/liboracle_san.so (make -C oracle sanitize) and
dfpsr_amd/_build/libdfpsr_hip_san.so (python -m dfpsr_amd.build --sanitize: host code
instrumented, device code not), then runs test_abi.py, test_dmf1.py and test_oracle.py with
clang's sanitizer runtime preloaded into the interpreter and both libraries named through
DFPSR_HIP_LIBRARY and DFPSR_ORACLE_LIBRARY.  Any sanitizer report fails the run (UBSan does not
recover; ASan aborts).  Leak checking is off: the interpreter itself keeps memory to the end.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = ["tests/test_abi.py", "tests/test_dmf1.py", "tests/test_oracle.py"]


def runtime() -> str:
    found = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not found:
        raise SystemExit("clang's AddressSanitizer runtime is not installed")
    return found[-1]


def main() -> int:
    import tempfile
    logs = tempfile.mkdtemp(prefix="dfpsr_san_")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    subprocess.run([sys.executable, "-m", "dfpsr_amd.build", "--sanitize"], cwd=ROOT, check=True, stdout=subprocess.DEVNULL)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": runtime(),
        "ASAN_OPTIONS": f"detect_leaks=0:abort_on_error=1:halt_on_error=1:log_path={logs}/asan",
        "UBSAN_OPTIONS": f"halt_on_error=1:print_stacktrace=1:log_path={logs}/ubsan",
        "DFPSR_HIP_LIBRARY": os.path.join(ROOT, "dfpsr_amd", "_build", "libdfpsr_hip_san.so"),
        "DFPSR_ORACLE_LIBRARY": os.path.join(ROOT, "oracle", "_build", "liboracle_san.so"),
    })
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider"] + TESTS,
                         cwd=ROOT, env=env, capture_output=True, text=True)
    out = res.stdout + res.stderr
    print(out[-2000:])
    reports = []
    for name in sorted(os.listdir(logs)):
        with open(os.path.join(logs, name)) as f:
            text = f.read()
        print(f"--- {name}\n{text[:3000]}")
        reports += [l for l in text.splitlines() if "ERROR: AddressSanitizer" in l or "runtime error:" in l]
    if res.returncode != 0 or reports:
        print(f"sanitized run failed: exit {res.returncode}, {len(reports)} sanitizer lines")
        return 1
    print("sanitized run clean")
    return 0


if __name__ == "__main__":
    sys.exit(main())
