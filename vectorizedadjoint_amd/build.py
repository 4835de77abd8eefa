"""Build libvectorizedadjoint_amd.so (HIP, gfx950) in-tree with hipcc.

    python -m vectorizedadjoint_amd.build [--incremental] [-j N]

Every source under csrc/ is compiled to an object in parallel, then linked
into vectorizedadjoint_amd/libvectorizedadjoint_amd.so next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvectorizedadjoint_amd.so")
EXAMPLES_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "va_examples.cpp")
EXAMPLES_BIN = os.path.join(os.path.dirname(HERE), "tests", "cpp", "va_examples")
USER_SYSTEM_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "va_user_system.hip")
USER_SYSTEM_LIB = os.path.join(os.path.dirname(HERE), "tests", "cpp", "libva_user_system.so")
USER_HOST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "brusselator_host.cpp")
USER_HOST_LIB = os.path.join(os.path.dirname(HERE), "tests", "cpp", "libva_brusselator_host.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("VA_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: only explicit fma() calls are fused, so the IEEE operation
# sequence is the one written in the sources.  The packed-FP32 feature is
# switched off for all device code.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I", INCLUDE,
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _headers_mtime() -> float:
    hs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(INCLUDE, "*.h"))
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _compile(src: str, obj: str, verbose: bool) -> tuple[str, float]:
    cmd = [hipcc()] + HIPCC_FLAGS + ["-x", "hip", "-c", src, "-o", obj]
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed for {os.path.basename(src)}:\n{msg}")
    if verbose:
        print(f"  compiled {os.path.basename(src)} in {time.time() - t:.1f}s", flush=True)
    return obj, time.time() - t


def build(incremental: bool = False, jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    hdr_t = _headers_mtime()
    todo, objs = [], []
    for s in srcs:
        o = os.path.join(OBJDIR, os.path.basename(s) + ".o")
        objs.append(o)
        if (not incremental or not os.path.exists(o) or os.path.getmtime(o) < max(
                os.path.getmtime(s), hdr_t)):
            todo.append((s, o))
    jobs = jobs or min(len(todo) or 1, os.cpu_count() or 4, 16)
    t0 = time.time()
    if todo:
        with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
            futs = [ex.submit(_compile, s, o, verbose) for s, o in todo]
            for f in cf.as_completed(futs):
                f.result()
    if todo or not os.path.exists(LIB):
        tmp = LIB + ".tmp"
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr}")
        os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} ({len(todo)} objects recompiled, {time.time() - t0:.1f}s)", flush=True)
    return LIB


def build_user_system(verbose: bool = True) -> str:
    """Compile the test suite's own device system (tests/cpp/va_user_system.hip, the
    Brusselator) the way a caller compiles one: hipcc, the library's kernel
    templates included through vectorizedadjoint/device_system.hpp.  The same
    functor's host build (tests/cpp/brusselator_host.cpp, host compiler only:
    a library without device code) gives the oracle its callbacks."""
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", CSRC, "-x", "hip", "-shared", USER_SYSTEM_SRC,
                                     "-o", USER_SYSTEM_LIB]
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"user system build failed:\n{msg}")
    cxx = os.environ.get("CXX", shutil.which("g++") or "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", USER_HOST_SRC,
           "-o", USER_HOST_LIB]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"user system host build failed:\n{r.stderr}")
    if verbose:
        print(f"built {USER_SYSTEM_LIB} and {USER_HOST_LIB} in {time.time() - t:.1f}s",
              flush=True)
    return USER_SYSTEM_LIB


def build_examples(verbose: bool = True) -> str:
    """Compile the drop-in C++ example drivers (tests/cpp/va_examples.cpp) with the
    host compiler against the C++ headers and the built library: the headers are
    plain C++17 over the C ABI and the HIP runtime API, no device code.  The
    'user' scenario's device system comes from tests/cpp/libva_user_system.so
    (build_user_system first)."""
    cxx = os.environ.get("CXX", shutil.which("g++") or "g++")
    libdir = os.path.relpath(HERE, os.path.dirname(EXAMPLES_BIN))
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", INCLUDE,
           "-I", os.path.join(ROCM, "include"), EXAMPLES_SRC, "-o", EXAMPLES_BIN,
           "-L", HERE, "-lvectorizedadjoint_amd", "-L", os.path.dirname(EXAMPLES_BIN),
           "-lva_user_system", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
           f"-Wl,-rpath,$ORIGIN/{libdir}", "-Wl,-rpath,$ORIGIN",
           f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"example build failed:\n{r.stderr}")
    if verbose:
        print(f"built {EXAMPLES_BIN} in {time.time() - t:.1f}s", flush=True)
    return EXAMPLES_BIN


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    build(incremental=a.incremental, jobs=a.jobs)
    build_user_system()
    build_examples()
    return 0


if __name__ == "__main__":
    sys.exit(main())
