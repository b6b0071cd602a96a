"""Builds libfastem.so (HIP kernels + C ABI + C++ host) and the em2 driver
in-tree for gfx950 with hipcc.  Usage: python -m fastem_amd.build"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libfastem.so")

HIP_SOURCES = ["fem_kernels.hip", "fem_cov64.hip", "fem_cov64_narrow.hip", "fem_kernels32.hip", "fem_kernels_bf16.hip", "fem_kmeans.hip",
               "fem_batch.hip", "fem_setnet.hip", "fem_abi.hip", "fem_noalloc.hip", "fem_sort.hip"]
HOST_SOURCES = ["fem_tools.cpp", "host/linalg.cpp", "host/rand.cpp", "host/block_manager.cpp",
                "host/normal_density.cpp", "host/gaussian_mixture.cpp", "host/dataset.cpp",
                "host/em_algorithm.cpp", "host/em_initializer.cpp", "host/setnet_init.cpp"]
TOOLS = {"em2": "host/em2_main.cpp", "eminit": "host/eminit_main.cpp"}
# MFMA accumulators in VGPRs (no accumulator-read moves when results are
# squared, reduced or flushed) for every file but fem_cov64.hip, whose FP64
# covariance kernel for d > 48 measured faster with the default AGPR form.
VGPR_FORM = ["-mllvm", "-amdgpu-mfma-vgpr-form"]
FILE_FLAGS = {"fem_cov64_narrow.hip": VGPR_FORM, "fem_kernels.hip": VGPR_FORM, "fem_kernels32.hip": VGPR_FORM, "fem_kernels_bf16.hip": VGPR_FORM,
              "fem_kmeans.hip": VGPR_FORM, "fem_batch.hip": VGPR_FORM, "fem_setnet.hip": VGPR_FORM}

# Packed FP32 VALU ops are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def _jobs():
    try:
        n = int(os.environ.get("MAX_JOBS", "0"))
    except ValueError:
        n = 0
    return max(1, min(n or (os.cpu_count() or 1), 16))


def build(verbose=True):
    objdir = os.path.join(PKG, "_objs")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for s in HIP_SOURCES + HOST_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, os.path.splitext(s.replace("/", "_"))[0] + ".o")
        flags = list(HIP_FLAGS) + FILE_FLAGS.get(s, [])
        if s.endswith(".cpp"):
            flags += ["-fopenmp"]
        cmds.append([hipcc()] + flags + ["-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj])
        objs.append(obj)
    # the largest translation units first, so the pool drains evenly
    order = sorted(range(len(cmds)), key=lambda i: -os.path.getsize(cmds[i][-3]))
    with ThreadPoolExecutor(_jobs()) as pool:
        for f in [pool.submit(_run, cmds[i]) for i in order]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the em2 and eminit drivers, linked against the library next to them
    for name, rel in TOOLS.items():
        src = os.path.join(CSRC, rel)
        obj = os.path.join(objdir, name + "_main.o")
        exe = os.path.join(PKG, name)
        _run([hipcc(), "-O3", "-std=c++17", "-fopenmp", "-I", os.path.join(ROOT, "include"), "-c",
              src, "-o", obj])
        _run([hipcc(), "-fopenmp", "-o", exe + ".tmp", obj, "-L", PKG, "-lfastem",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
