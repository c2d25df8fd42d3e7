"""Compiles the gfx950 library liteqg_amd/liblqg.so with hipcc.

    python -m liteqg_amd.build              # liteqg_amd/liblqg.so
    python -m liteqg_amd.build --sanitize   # liteqg_amd/_san/liblqg.so: host code
                                            # with AddressSanitizer + UBSan

The sanitizer build instruments host code only (GPU code is built as usual):
load it with LQG_LIB=liteqg_amd/_san/liblqg.so and the clang ASan runtime
preloaded (see oracle/Makefile, target check-san).

All device code is built for gfx950 only, with packed-FP32 instruction
selection disabled and FP contraction off (every fused multiply-add in the
kernels is written explicitly, so rounding follows the source).
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblqg.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _compile(src: str, obj: str, extra, verbose: bool) -> None:
    flags = [f for f in FLAGS if f != "-shared"]
    cmd = [HIPCC, *flags, *extra, "-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}")


SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g", "-shared-libsan"]


def build(verbose: bool = False, extra=(), sanitize: bool = False) -> str:
    """Compile every source into liblqg.so (always recompiles; one hipcc
    process per source, in parallel, then one link).  sanitize: host code
    with AddressSanitizer and UBSan, into liteqg_amd/_san/liblqg.so."""
    from concurrent.futures import ThreadPoolExecutor

    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    lib_path = LIB
    if sanitize:
        extra = (*extra, *SAN_FLAGS)
        lib_path = os.path.join(HERE, "_san", "liblqg.so")
        os.makedirs(os.path.dirname(lib_path), exist_ok=True)
    objdir = os.path.join(HERE, "_san_obj" if sanitize else "_obj")
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_compile, s, o, extra, verbose) for s, o in zip(srcs, objs)]:
            f.result()
    tmp = lib_path + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
           *(SAN_FLAGS if sanitize else ())]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, lib_path)
    return lib_path


if __name__ == "__main__":
    print(build(verbose=True, sanitize="--sanitize" in sys.argv[1:]))
