"""Build libtmrx.so (host C++ + HIP kernels for gfx950) in-tree.

Usage: python -m tmr_amd.build  (or tmr_amd.build.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtmrx.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions in generated device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    # matrix-core results in ordinary VGPRs (no accumulator-register copies)
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-I" + os.path.join(ROCM, "include"), "-D__HIP_PLATFORM_AMD__"]


def _sources():
    srcs = sorted(os.listdir(CSRC))
    cpp = [os.path.join(CSRC, s) for s in srcs if s.endswith(".cpp")]
    hip = [os.path.join(CSRC, s) for s in srcs if s.endswith(".hip")]
    return cpp, hip


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    cpp, hip = _sources()
    jobs = []
    for s in cpp:
        o = os.path.join(OUT_DIR, os.path.basename(s) + ".o")
        jobs.append((["g++"] + CXX_FLAGS + ["-c", s, "-o", o], o))
    for s in hip:
        o = os.path.join(OUT_DIR, os.path.basename(s) + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o], o))
    nproc = min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=nproc) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    objs = [o for _, o in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
