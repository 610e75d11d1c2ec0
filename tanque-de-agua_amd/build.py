"""Builds libtanque.so (HIP kernels for gfx950 + the C ABI) and the Node-API addon in-tree.

Usage: python tanque-de-agua_amd/build.py [--force]
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libtanque.so")
NODE_ADDON = os.path.join(LIB_DIR, "tanque_node.node")

HIP_SOURCES = ["k_particles.hip", "k_grid.hip", "k_air.hip", "k_solver.hip", "tq_api.hip", "tq_batch.hip",
               "tq_adjoint.hip", "k_scene.hip", "k_render.hip", "tq_tune.hip"]
CPP_SOURCES = ["scene.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed-FP32 instructions are disabled for all device code (known-bad on these boxes).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-result"]


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))] + \
        [os.path.join(ROOT, "include", "tanque.h")]


def _compile(src, force):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    if not force and not _newer(out, [src] + _headers()):
        return out
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", out]
    else:
        cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-c", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build_lib(force=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in HIP_SOURCES + CPP_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    if force or _newer(LIB, objs):
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
        os.replace(tmp, LIB)
    return LIB


SYNC_BENCH = os.path.join(LIB_DIR, "bench_sync")


def build_sync_bench(force=False):
    """The kernel-boundary / grid-barrier cost probe (csrc/bench_sync.hip; bench.py --sync-probe)."""
    src = os.path.join(CSRC, "bench_sync.hip")
    if not force and not _newer(SYNC_BENCH, [src]):
        return SYNC_BENCH
    tmp = SYNC_BENCH + ".tmp"
    r = subprocess.run([HIPCC] + [f for f in HIP_FLAGS if f != "-fPIC"] + [src, "-o", tmp], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("bench_sync build failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, SYNC_BENCH)
    return SYNC_BENCH


STREAM_BENCH = os.path.join(LIB_DIR, "bench_stream")


def build_stream_bench(force=False):
    """The HBM streaming-bandwidth probe (csrc/bench_stream.hip; bench.py --stream-probe)."""
    src = os.path.join(CSRC, "bench_stream.hip")
    if not force and not _newer(STREAM_BENCH, [src]):
        return STREAM_BENCH
    tmp = STREAM_BENCH + ".tmp"
    r = subprocess.run([HIPCC] + [f for f in HIP_FLAGS if f != "-fPIC"] + [src, "-o", tmp], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("bench_stream build failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, STREAM_BENCH)
    return STREAM_BENCH


def build_node_addon(force=False):
    """The thin Node-API addon over the C ABI (needs node's headers; skipped if absent)."""
    src = os.path.join(CSRC, "napi_addon.cpp")
    hdr = "/usr/include/node/node_api.h"
    if not os.path.exists(src) or not os.path.exists(hdr):
        return None
    if not force and not _newer(NODE_ADDON, [src, LIB, os.path.join(ROOT, "include", "tanque.h")]):
        return NODE_ADDON
    tmp = NODE_ADDON + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-I/usr/include/node", src, "-o", tmp,
           "-L" + LIB_DIR, "-ltanque", "-Wl,-rpath,$ORIGIN", "-Wl,--unresolved-symbols=ignore-all"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("addon build failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, NODE_ADDON)
    return NODE_ADDON


def build(force=False):
    lib = build_lib(force)
    addon = build_node_addon(force)
    build_sync_bench(force)
    build_stream_bench(force)
    return lib, addon


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
