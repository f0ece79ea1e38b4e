"""Builds libqumasun_hip.so (the C-ABI HIP library) in-tree for gfx950.

    python qumasun-workshop-template_amd/build.py [--jobs N] [--force]

Each csrc/*.hip (and host-only csrc/*.cpp) file is compiled to an object under _build/ (rebuilt when it or
a header is newer), then linked with rocBLAS / rocSOLVER / rocFFT into
qumasun-workshop-template_amd/libqumasun_hip.so.  The C++ consumer test
(tests/consumer/opea_consumer.cpp, a QUMASUN-style caller of the C ABI) is
built and linked against it afterwards.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libqumasun_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

CXXFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay disabled for every kernel in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    f"-I{ROOT}/include",
]
LDFLAGS = ["--offload-arch=gfx950", "-shared", f"-L{ROCM}/lib", "-lrocsolver", "-lrocblas", "-lrocfft",
           f"-Wl,-rpath,{ROCM}/lib"]
CONSUMER_SRC = os.path.join(ROOT, "tests", "consumer", "opea_consumer.cpp")
CONSUMER_BIN = os.path.join(ROOT, "tests", "consumer", "opea_consumer")


def _headers():
    return glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(ROOT, "include", "*.h"))


def _compile(src, force):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    newest = max([os.path.getmtime(src)] + [os.path.getmtime(h) for h in _headers()])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return obj, None
    cmd = [HIPCC, *CXXFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"$ {' '.join(cmd)}\n{r.stdout}{r.stderr}"
    return obj, None


def build(jobs: int = 8, force: bool = False, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        results = list(ex.map(lambda s: _compile(s, force), srcs))
    errors = [e for _, e in results if e]
    if errors:
        raise RuntimeError("HIP compilation failed:\n" + "\n".join(errors))
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *LDFLAGS, "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n$ {' '.join(cmd)}\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_consumer(verbose)
    return LIB


def build_consumer(verbose: bool = True) -> str:
    """The C++ consumer test program, linked against the library by a relative rpath."""
    rel = os.path.relpath(PKG, os.path.dirname(CONSUMER_BIN))
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", f"-I{ROOT}/include", CONSUMER_SRC, f"-L{PKG}", "-lqumasun_hip",
           f"-Wl,-rpath,$ORIGIN/{rel}", "-o", CONSUMER_BIN + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"consumer build failed:\n$ {' '.join(cmd)}\n{r.stdout}{r.stderr}")
    os.replace(CONSUMER_BIN + ".tmp", CONSUMER_BIN)
    if verbose:
        print(f"built {CONSUMER_BIN}")
    return CONSUMER_BIN


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    try:
        build(a.jobs, a.force)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
