"""Build libjayci_gpu.so in-tree with hipcc for gfx950.

    python jayci_amd/build.py          (run by path: importing the package loads the library)

Compiles the HIP kernels and the C++ host runtime into jayci_amd/libjayci_gpu.so.  The
build always recompiles; callers that only need the library load it if present.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libjayci_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["kernels.hip", "sigma_fast.hip", "sigma_gemm.hip", "batch_kernels.hip",
               "noalloc.hip"]
CXX_SOURCES = ["space.cpp", "jci.cpp", "capi.cpp", "batch.cpp", "symeig.cpp", "sifs.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-fvisibility=hidden", "-Wno-unused-result", "-Wno-unused-value", "-I" + CSRC,
          "-I" + os.path.join(ROOT, "include")]
# gfx950 only; packed-FP32 instruction selection is disabled for all device code
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False, jobs=None):
    """Compile every source (up to `jobs` hipcc processes at once, default the CPU count
    capped at 4) and link the library."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in HIP_SOURCES + CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip" if src.endswith(".hip") else "c++",
                                           "-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + COMMON + DEVICE + ["-D__HIP_PLATFORM_AMD__", "-c", path, "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    jobs = jobs or max(1, min(4, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for out in outs:
            if out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs +
         ["-lamdhip64"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
