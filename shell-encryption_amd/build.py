"""Builds libshellgpu.so (the HIP kernels + C-ABI) for gfx950 in-tree.

    python shell-encryption_amd/build.py            # library
    python shell-encryption_amd/build.py --tests    # also the C++ API tests

hipcc cross-compiles without a GPU.  The packed-FP32 feature is disabled for
all device code (the kernels use FP64 and integer arithmetic only).
"""
import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
CPP = os.path.join(HERE, "cpp")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libshellgpu.so")

SOURCES = ["shg_api.hip", "shg_ntt.hip", "shg_ops.hip", "shg_keyswitch.hip", "shg_batch.hip",
           "shg_pir.hip", "shg_serving.hip", "shg_wire.hip", "shg_bfv.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SHG_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + INCLUDE, "-I" + CPP, "-I" + CSRC, "-Wno-unused-result"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib(jobs=8):
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise SystemExit("compile failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise SystemExit("compile failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_cpp_tests():
    """The C++ API test binary; it links the CPU oracle as its checker."""
    src = os.path.join(CPP, "tests", "rns_api_test.cc")
    out = os.path.join(CPP, "tests", "rns_api_test")
    oracle_dir = os.path.join(ROOT, "oracle")
    oracle_obj = os.path.join(CPP, "tests", "shell_oracle_test.o")
    _run(["gcc", "-O2", "-fPIC", "-std=gnu11", "-c", os.path.join(oracle_dir, "shell_oracle.c"),
          "-o", oracle_obj])
    # Host-only C++ over the C-ABI: the system compiler is enough.
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-I" + CPP, "-I" + oracle_dir, src,
          oracle_obj, "-o", out, "-L" + HERE, "-lshellgpu", "-Wl,-rpath," + HERE])
    os.remove(oracle_obj)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--tests", action="store_true")
    args = ap.parse_args(argv)
    build_lib()
    if args.tests:
        build_cpp_tests()


if __name__ == "__main__":
    sys.exit(main())
