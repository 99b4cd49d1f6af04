"""Build the gfx950 HIP library ``grim_amd/libgrim_hip.so`` with hipcc.

Every ``grim_amd/csrc/*.hip`` file is compiled for gfx950 only and linked
into one shared library exporting the C ABI of ``include/grim_hip.h``.

Flags that matter:
  -ffp-contract=off   no fused multiply-add contraction, so FP64 results are
                      rounded operation by operation as written (every file
                      except *_fast.hip, which use
                      -ffp-contract=fast-honor-pragmas: a*b+c may fuse,
                      except inside a `#pragma clang fp contract(off)` block,
                      which the shared fold arithmetic uses to keep its bits
                      in both kinds of translation unit);
  -fno-hip-fp32-correctly-rounded-divide-sqrt
                      *_fast.hip only: FP32 division and square root by the
                      hardware reciprocal sequences (the FP32 Jacobian and LU
                      of the mixed path need no correctly rounded quotients;
                      FP64 arithmetic is unaffected);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keeps the compiler from emitting packed FP32 math.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libgrim_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 8) -> str:
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]
    procs = []
    objs = []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        flags = list(HIPCC_FLAGS)
        if src.endswith("_fast.hip"):
            # fused translation units: a*b+c may become one FMA
            flags[flags.index("-ffp-contract=off")] = "-ffp-contract=fast-honor-pragmas"
            flags.append("-fno-hip-fp32-correctly-rounded-divide-sqrt")
        cmd = [hipcc(), *flags, *inc, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        # the host pass does not know the packed-FP32 feature flag: harmless
        text = "".join(line for line in out.decode(errors="replace").splitlines(True)
                       if "packed-fp32-ops' is not a recognized feature" not in line)
        if text:
            sys.stdout.write(text)
        if p.returncode != 0:
            failed.append(src)
    if failed:
        raise RuntimeError("hipcc failed on: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_io(verbose)
    build_driver(verbose)
    return LIB


IO_LIB = os.path.join(HERE, "libgrim_io.so")


def build_io(verbose: bool = False) -> str:
    """grim_amd/libgrim_io.so: grim's HDF5 dumps and restart files
    (include/grim_io.h), host code only; the HDF5 C library is loaded at run
    time."""
    tmp = IO_LIB + ".tmp"
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-shared", "-fPIC",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "host", "grim_io.cpp"), "-ldl", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, IO_LIB)
    return IO_LIB


DRIVER = os.path.join(HERE, "grim_step")
FACADE_CHECK = os.path.join(HERE, "facade_check")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def host_program(src: str, out: str, libs, verbose: bool = False) -> str:
    """A host program over the C ABI and the C++ interfaces of grim_amd/host
    (g++; the HIP runtime's host API for device memory)."""
    tmp = out + ".tmp"
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "host"), "-isystem",
           os.path.join(ROCM, "include"), os.path.join(HERE, "host", src), "-L", HERE, *libs,
           "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib"),
           "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_driver(verbose: bool = False) -> str:
    """The C++ driver grim_amd/grim_step and the interface check
    grim_amd/facade_check (host code only: g++ against the C ABI and the C++
    interfaces of grim_amd/host)."""
    host_program("grim_main.cpp", DRIVER, ["-lgrim_hip", "-lgrim_io"], verbose)
    host_program("facade_check.cpp", FACADE_CHECK, ["-lgrim_hip"], verbose)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
