"""Builds the HIP library (libkp_hip.so) in-tree for gfx950.

Usage: python -m kazen_proto_amd.build   (or via __graft_entry__.build()).
Every call recompiles; objects go to kazen_proto_amd/_build/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkp_hip.so")

SOURCES = ["capi.cpp", "bvh_build.cpp", "host_trace.cpp", "host_api.cpp", "trace.hip", "render.hip",
           "guiding.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-parameter",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-ffp-contract=fast-honor-pragmas",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # packed FP32 instructions stay off (see the build notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _compile(src):
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", path, "-o", out]
    else:
        # host-only translation units: no device code, x86-64-v3 host code
        cmd = [HIPCC, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-march=x86-64-v3",
               "-ffp-contract=off"] + COMMON + ["-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build_library(verbose=False):
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose=False):
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built oracle/liboracle.so")


def build_adapter_tests(verbose=False):
    """Compiles the C++ programs under tests/cpp/ against the adapter headers
    (kazen_proto_amd/cpp/) and libkp_hip.so, into tests/cpp/bin/."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "bin")
    srcs = sorted(f for f in os.listdir(src_dir) if f.endswith(".cpp")) if os.path.isdir(src_dir) else []
    if not srcs:
        return []
    os.makedirs(out_dir, exist_ok=True)
    outs = []
    for f in srcs:
        exe = os.path.join(out_dir, os.path.splitext(f)[0])
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
               "-I" + os.path.join(HERE, "cpp"), os.path.join(src_dir, f), "-o", exe,
               "-L" + HERE, "-lkp_hip", "-Wl,-rpath,$ORIGIN/../../../kazen_proto_amd", "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("adapter test build failed: %s\n%s" % (" ".join(cmd), r.stderr))
        outs.append(exe)
    if verbose:
        print("built", len(outs), "adapter test program(s)")
    return outs


def main():
    build_library(verbose=True)
    build_oracle(verbose=True)
    build_adapter_tests(verbose=True)


if __name__ == "__main__":
    sys.exit(main())
