"""Builds the native libraries of pixelforge_amd in-tree.

    python -m pixelforge_amd.build             # everything
    python -m pixelforge_amd.build --check     # exit 1 if a library is missing
    python -m pixelforge_amd.build --sanitize  # host layer under ASan/UBSan (_san/)

Outputs (next to this file):
    libpixelforge_amd.so  drop-in PixelForge API: host C state machine
                          (csrc/pf_host.c) + HIP kernels for gfx950
                          (csrc/pf_device.hip)
    libpf_scenes.so       seeded synthetic scenes issued through any library
                          exporting the PixelForge API (csrc/pf_scenes.c)

hipcc cross-compiles for gfx950 without a GPU.  Device code is built with
-ffp-contract=off (no FMA contraction: parity with the CPU arithmetic) and
with packed FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")

LIB = os.path.join(HERE, "libpixelforge_amd.so")
SCENES = os.path.join(HERE, "libpf_scenes.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
C_FLAGS = ["-std=gnu99", "-O3", "-fPIC", "-fno-semantic-interposition", "-ffp-contract=off", "-fno-fast-math",
           "-Wall", "-Wextra", "-Wno-unused-parameter", "-fvisibility=hidden"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: install ROCm or put hipcc on PATH")
    return path


SAN_DIR = os.path.join(HERE, "_san")
SAN_LIB = os.path.join(SAN_DIR, "libpixelforge_amd.so")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined", "-g"]


def build_sanitized():
    """The host layer (pf_host.c) built with AddressSanitizer and
    UndefinedBehaviorSanitizer into _san/libpixelforge_amd.so, linked with
    the same device object.  Run the CPU tests against it with

        LD_PRELOAD=$(gcc -print-file-name=libasan.so) ASAN_OPTIONS=detect_leaks=0 \
        PIXELFORGE_AMD_LIB=pixelforge_amd/_san/libpixelforge_amd.so \
        python -m pytest tests -m "not gpu" -p no:xdist

    (the sanitizer runtime must be loaded first, hence LD_PRELOAD; leaks of
    the Python interpreter itself are not reported)."""
    os.makedirs(SAN_DIR, exist_ok=True)
    dev_o = os.path.join(BUILD, "pf_device.o")
    if not os.path.exists(dev_o):
        build()
    host_o = os.path.join(SAN_DIR, "pf_host.o")
    _run(["gcc"] + C_FLAGS + SAN_FLAGS + ["-I", INCLUDE, "-c", os.path.join(CSRC, "pf_host.c"), "-o", host_o])
    tmp = SAN_LIB + ".tmp"
    _run(["g++", "-shared", "-fPIC", "-o", tmp, dev_o, host_o, "-L/opt/rocm/lib", "-lamdhip64", "-lubsan",
          "-Wl,-Bsymbolic", "-lm"])
    os.replace(tmp, SAN_LIB)
    return SAN_LIB


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    dev_o = os.path.join(BUILD, "pf_device.o")
    host_o = os.path.join(BUILD, "pf_host.o")
    _run([hipcc()] + HIP_FLAGS + ["-I", INCLUDE, "-c", os.path.join(CSRC, "pf_device.hip"), "-o", dev_o])
    _run(["gcc"] + C_FLAGS + ["-I", INCLUDE, "-c", os.path.join(CSRC, "pf_host.c"), "-o", host_o])
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, dev_o, host_o,
          "-Wl,-Bsymbolic", "-lm"])
    os.replace(tmp, LIB)
    tmp = SCENES + ".tmp"
    _run(["gcc"] + [f for f in C_FLAGS if f != "-fvisibility=hidden"] +
         ["-shared", "-I", INCLUDE, "-o", tmp, os.path.join(CSRC, "pf_scenes.c"), "-ldl", "-lm"])
    os.replace(tmp, SCENES)
    return [LIB, SCENES]


def built():
    return all(os.path.exists(p) for p in (LIB, SCENES))


if __name__ == "__main__":
    if "--check" in sys.argv:
        sys.exit(0 if built() else 1)
    if "--sanitize" in sys.argv:
        build_sanitized()
    else:
        build()
