"""Builds libdfpsr_hip.so in-tree with hipcc for gfx950.

    python -m dfpsr_amd.build

Float behaviour is pinned to match the reference: no FMA contraction, no fast-math,
correctly rounded fp32 division and square root.  Packed-FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdfpsr_hip.so")
SOURCES = ["api.hip", "frame.hip", "kernels.hip", "depth.hip"]
HEADERS = ["geometry.h", "occlusion.h", "shading.h", "pipeline.h", "frame.h", "dmf1.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++14", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


# Host-only sources of the library, plain C++14 built with g++ and linked into the library.
HOST_SOURCES = ["dmf1.cpp"]
HOST_FLAGS = ["-std=c++14", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra"]
OBJ_DIR = os.path.join(HERE, "_build")


# Host-side sanitizers (AddressSanitizer, UndefinedBehaviorSanitizer) for the library's CPU
# code: argument checks, model and texture bookkeeping, the DMF1 parser, the occlusion grid's
# host part.  Device code is not instrumented.  The plain host sources are built with clang here
# (not g++), so that the process needs only clang's sanitizer runtime.
SANITIZE_FLAGS = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
SANITIZED_LIB = os.path.join(HERE, "_build", "libdfpsr_hip_san.so")
CLANGXX = "/opt/rocm/lib/llvm/bin/clang++"


def build_host_objects(verbose: bool = False, sanitize: bool = False) -> list:
    os.makedirs(OBJ_DIR, exist_ok=True)
    objects = []
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ("_san.o" if sanitize else ".o"))
        cmd = ([CLANGXX] + HOST_FLAGS + SANITIZE_FLAGS if sanitize else ["g++"] + HOST_FLAGS) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"g++ failed to build {src}")
        objects.append(obj)
    return objects


def build_sanitized(verbose: bool = False) -> str:
    """The library with its host code sanitized, into dfpsr_amd/_build/ (tests/run_sanitized.py)."""
    objects = build_host_objects(verbose, sanitize=True)
    host_only = []
    for flag in SANITIZE_FLAGS:
        host_only += ["-Xarch_host", flag]
    cmd = [hipcc()] + HIPCC_FLAGS + host_only + ["-I" + os.path.join(ROOT, "include")] + \
        [os.path.join(CSRC, s) for s in SOURCES] + ["-x", "none"] + objects + ["-o", SANITIZED_LIB]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed to build the sanitized library")
    return SANITIZED_LIB


def build(verbose: bool = False) -> str:
    objects = build_host_objects(verbose)
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I" + os.path.join(ROOT, "include")] + \
        [os.path.join(CSRC, s) for s in SOURCES] + ["-x", "none"] + objects + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed to build libdfpsr_hip.so")
    os.replace(LIB + ".tmp", LIB)
    build_facade_program(verbose)
    return LIB


FACADE_SOURCE = os.path.join(ROOT, "tests", "host", "facade_render.cpp")
FACADE_PROGRAM = os.path.join(ROOT, "tests", "host", "_build", "facade_render")


def build_facade_program(verbose: bool = False) -> str:
    """The C++14 facade's test program: plain g++ at the reference's language level, linked
    against the in-tree library."""
    os.makedirs(os.path.dirname(FACADE_PROGRAM), exist_ok=True)
    cmd = ["g++", "-std=c++14", "-O2", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"), FACADE_SOURCE,
           "-L" + HERE, "-ldfpsr_hip", "-Wl,-rpath," + os.path.relpath(HERE, os.path.dirname(FACADE_PROGRAM)).join(["$ORIGIN/", ""]),
           "-o", FACADE_PROGRAM]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed to build the facade test program")
    return FACADE_PROGRAM


if __name__ == "__main__":
    if "--sanitize" in sys.argv[1:]:
        print(build_sanitized(verbose=True))
    else:
        print(build(verbose=True))
