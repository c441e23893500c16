"""Builds the HIP extension libtpic_mi.so for gfx950 in-tree (tpic_amd/lib/).

Each .hip translation unit is compiled with hipcc and the objects are linked
into one shared library exposing the C-ABI of include/tpic_mi.h.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib")
LIB = os.path.join(OUT, "libtpic_mi.so")
ARCH = os.environ.get("TPIC_MI_ARCH", "gfx950")

SOURCES = ["context.hip", "sort.hip", "mover.hip", "deposit.hip", "maxwell.hip", "gmres.hip",
           "diagnostics.hip", "loader.hip", "adjoint.hip", "checkpoint.hip"]
# Contraction-free translation units: their results are compared bit for bit.
NO_CONTRACT = {"mover.hip", "loader.hip"}
# Per-file code generation: the deposition's MFMA accumulators stay in the
# architectural VGPRs (no copies to and from the accumulation registers
# around the evaluation code).
EXTRA_FLAGS = {"deposit.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}

BASE_FLAGS = ["-O3", "-std=c++20", "-fPIC", "--offload-arch=" + ARCH,
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OUT, "obj", os.path.splitext(src)[0] + ".o")
    flags = list(BASE_FLAGS)
    if src in NO_CONTRACT:
        flags.append("-ffp-contract=off")
    flags += EXTRA_FLAGS.get(src, [])
    cmd = [hipcc()] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(os.path.join(OUT, "obj"), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as pool:
        results = list(pool.map(_compile, SOURCES))
    if verbose:
        for obj, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results] + [_compile_host(s) for s in DECK_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    build_driver()
    return LIB


# The deck front end (parser, admission, startup over the C-ABI), also linked
# into the library for tpic_mi_load_deck.
DECK_SOURCES = ["driver/deck.cpp", "driver/deck_start.cpp", "driver/deck_abi.cpp"]


def host_cxx():
    return os.environ.get("CXX", "g++")


def _compile_host(src):
    obj = os.path.join(OUT, "obj", os.path.basename(os.path.splitext(src)[0]) + ".o")
    cmd = [host_cxx(), "-std=c++20", "-O2", "-fPIC", "-Wall", "-Wextra", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host compile failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


# The deck-driven command-line driver: host C++ over the C-ABI only.
DRIVER_SOURCES = ["driver/deck.cpp", "driver/deck_start.cpp", "driver/output.cpp", "driver/tpic_run.cpp"]
DRIVER = os.path.join(OUT, "tpic_mi_run")


# The driver under AddressSanitizer + UndefinedBehaviorSanitizer (host code:
# the deck parser, admission, output writers), for the CPU tests.
SANITIZED_DRIVER = os.path.join(OUT, "sanitize", "tpic_mi_run")


def build_driver(sanitize=False):
    """Builds tpic_mi_run; sanitize=True builds the sanitized copy instead."""
    target = SANITIZED_DRIVER if sanitize else DRIVER
    os.makedirs(os.path.dirname(target), exist_ok=True)
    tmp = target + ".tmp"
    flags = ["-std=c++20", "-Wall", "-Wextra"]
    if sanitize:
        flags += ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                  "-fno-sanitize-recover=all"]
        rpath = "-Wl,-rpath,$ORIGIN/.."
    else:
        flags += ["-O2"]
        rpath = "-Wl,-rpath,$ORIGIN"
    cmd = ([host_cxx()] + flags + [os.path.join(CSRC, s) for s in DRIVER_SOURCES]
           + ["-L", OUT, "-ltpic_mi", rpath, "-o", tmp])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, target)
    return target


if __name__ == "__main__":
    print(build(verbose=True))
