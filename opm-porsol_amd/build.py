"""Build recipe for the HIP library libmhs.so (gfx950) and the C++ drop-in examples.

Objects go to opm-porsol_amd/_build/ (not tracked); the shared library is
written next to this file so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmhs.so")
ARCH = os.environ.get("MHS_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 (v_pk_*_f32) code generation is disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-value",
          "-Wno-unused-result", f"--offload-arch={ARCH}",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]

SOURCES = ["mhs_kernels.hip", "mhs_transport.hip", "mhs_scratch.hip", "mhs_api.cpp", "mhs_structure.cpp",
           "mhs_amg.cpp", "mhs_ebe.cpp", "mhs_deck.cpp"]


def _compile(src: str) -> tuple[subprocess.Popen, str]:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".hip"):
        cmd += DEVICE_FLAGS + ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True), obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = [(_compile(s), s) for s in SOURCES]
    objs = []
    for (proc, obj), src in jobs:
        out, _ = proc.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if proc.returncode != 0:
            raise RuntimeError(f"compiling {src} failed:\n{out}")
        if verbose and out:
            print(out, file=sys.stderr)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"linking libmhs.so failed:\n{res.stdout}")
    os.replace(tmp, LIB)
    return LIB




EXAMPLES = os.path.join(HERE, "examples")
HOST_INCLUDE = os.path.join(HERE, "host")


def build_examples(verbose: bool = False) -> list[str]:
    """Compile the C++ drop-in drivers (examples/*.cpp) against libmhs.so."""
    out = []
    if not os.path.isdir(EXAMPLES):
        return out
    for name in sorted(os.listdir(EXAMPLES)):
        if not name.endswith(".cpp"):
            continue
        exe = os.path.join(BUILD, os.path.splitext(name)[0])
        os.makedirs(BUILD, exist_ok=True)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", HOST_INCLUDE, "-I", os.path.join(ROOT, "include"),
               os.path.join(EXAMPLES, name), "-o", exe, "-L", HERE, "-lmhs", f"-Wl,-rpath,{HERE}"]
        res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"compiling {name} failed:\n{res.stdout}")
        if verbose and res.stdout:
            print(res.stdout, file=sys.stderr)
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_examples(verbose=True))
