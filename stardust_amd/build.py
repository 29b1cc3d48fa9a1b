"""Builds libstardust_hip.so (hipcc, gfx950) in-tree.

Usage: python -m stardust_amd.build   (or stardust_amd.build.build())

Every HIP translation unit is compiled with
  -ffp-contract=off                       no a*b+c contraction: the contact law relies on
                                          exact sign symmetry of each pair force
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division / sqrt
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libstardust_hip.so")
ARCH = os.environ.get("SD_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include")]
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-fno-gpu-rdc",
          "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST = ["-D__HIP_PLATFORM_AMD__"]


def sources():
    out = []
    for root, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith((".hip", ".cpp")):
                out.append(os.path.join(root, f))
    return sorted(out)


def _compile(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    obj = os.path.join(BUILD, rel + ".o")
    flags = COMMON + (DEVICE if src.endswith(".hip") else HOST)
    cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, "from", len(srcs), "sources")
    build_engine(verbose)
    return LIB


ENGINE_DIR = os.path.join(HERE, "engine")
ENGINE_LIB = os.path.join(HERE, "libstardust_engine.so")
CXX = shutil.which("g++") or "g++"


def build_engine(verbose=True):
    """Host-only C++ engine API (STARDUST::DEMEngine) over the C ABI, linked to libstardust_hip.so."""
    srcs = sorted(os.path.join(ENGINE_DIR, f) for f in os.listdir(ENGINE_DIR) if f.endswith(".cpp"))
    tmp = ENGINE_LIB + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-ffp-contract=off",
           "-I" + os.path.join(os.path.dirname(HERE), "include"), "-I" + ENGINE_DIR] + srcs + \
          ["-L" + HERE, "-lstardust_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("engine build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, ENGINE_LIB)
    if verbose:
        print("built", ENGINE_LIB, "from", len(srcs), "sources")
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    # the driver sits in engine/bin/, the libraries two levels up
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-I" + ENGINE_DIR, os.path.join(ENGINE_DIR, "examples", "dem_driver.cpp"),
           "-L" + HERE, "-lstardust_engine", "-Wl,-rpath,$ORIGIN/../..", "-o", DRIVER + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(DRIVER + ".tmp", DRIVER)
    return ENGINE_LIB


DRIVER = os.path.join(ENGINE_DIR, "bin", "sd_dem_driver")


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
