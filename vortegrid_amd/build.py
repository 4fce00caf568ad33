"""Build libvortegrid.so (HIP for gfx950) in-tree with hipcc.

    python -m vortegrid_amd.build          # compile + link
    python -m vortegrid_amd.build --clean

Every .hip under csrc/ is compiled to an object under _build/ and linked
into vortegrid_amd/libvortegrid.so.  Packed FP32 instructions are disabled
for all device code (-packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvortegrid.so")
ARCH = os.environ.get("VG_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CFLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unknown-pragmas",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


HASH_PATH = LIB + ".srchash"
HEADER = os.path.join(os.path.dirname(HERE), "include", "vortegrid.h")


def source_hash() -> str:
    """SHA-256 over every file the library is built from (csrc/ and the
    public header), so a stale library can be told from a current one
    without trusting file times."""
    h = hashlib.sha256()
    files = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)) + [HEADER]
    for path in files:
        if os.path.isfile(path):
            h.update(os.path.basename(path).encode())
            with open(path, "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def built_hash():
    """The source hash recorded when the library was last built, or None."""
    if not (os.path.exists(LIB) and os.path.exists(HASH_PATH)):
        return None
    with open(HASH_PATH) as fh:
        return fh.read().strip()


def is_current() -> bool:
    return built_hash() == source_hash()


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    digest = source_hash()
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    with open(HASH_PATH, "w") as fh:
        fh.write(digest + "\n")
    if os.path.exists(FACADE_SRC):
        build_facade_driver()
    if verbose:
        print(f"built {LIB}")
    return LIB


FACADE_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "facade_frame.cpp")
FACADE_EXE = os.path.join(os.path.dirname(HERE), "tests", "cpp", "facade_frame")


def build_facade_driver() -> str:
    """Compile the C++ facade frame driver (tests/cpp) against libvortegrid."""
    root = os.path.dirname(HERE)
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", f"-I{root}", FACADE_SRC,
           f"-L{HERE}", "-lvortegrid", "-Wl,-rpath,$ORIGIN/../../vortegrid_amd", "-o", FACADE_EXE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade driver build failed:\n{r.stdout}\n{r.stderr}")
    return FACADE_EXE


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    if a.clean:
        clean()
        sys.exit(0)
    build(verbose=True)
