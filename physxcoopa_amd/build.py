"""Builds libpxg.so (the HIP contact step for gfx950 behind include/pxg.h) in-tree.

    python -m physxcoopa_amd.build

Every translation unit is compiled with -ffp-contract=off so no multiply-add pair is fused
(the FP32 results must match the reference's unfused arithmetic), with HIP's default correctly
rounded fp32 division and square root, and with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpxg.so")
SOURCES = ["pxg_kernels.hip", "pxg_world.hip", "pxg_batch.hip", "pxg_mlp.hip", "pxg_events.hip", "pxg_query.hip", "pxg_tiles.hip",
           "pxg_soft.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PXG_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def build(verbose=False, jobs=4):
    objs = []
    procs = []
    tmp = tempfile.mkdtemp(prefix="pxg_build_")
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(tmp, src + ".o")
        cmd = [HIPCC] + FLAGS + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    out_tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", out_tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode())
    os.replace(out_tmp, LIB)
    build_facade_test(verbose)
    return LIB


FACADE_TEST = os.path.join(HERE, "pxg_world_test")


def build_facade_test(verbose=False):
    """The reference's PhysicsWorld tests over the C++ facade (tests/cpp), linked to libpxg.so."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "physics_world_test.cpp")
    if not os.path.exists(src):
        return None
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-Wextra", "-pthread", "-I", os.path.join(root, "include"),
           src, "-L", HERE, "-lpxg", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", FACADE_TEST]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("facade test build failed:\n" + r.stdout.decode())
    return FACADE_TEST


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = out.decode()
    if p.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, text))
    errs = [l for l in text.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if errs:
        sys.stderr.write("\n".join(errs[:20]) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
