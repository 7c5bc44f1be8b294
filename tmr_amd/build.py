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
         ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-lrocblas",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
             "-fno-omit-frame-pointer", "-g"]


def build_sanitized(out_dir=None, verbose=False):
    """libtmrx_asan.so in out_dir (default: a directory under the system
    temporary directory, outside the package): the host C++ (forest,
    problem set-up, C-ABI) compiled by g++ with AddressSanitizer and
    UndefinedBehaviorSanitizer, linked with the device objects of the last
    build().  For CPU tests of the host code, run with the sanitizer runtimes
    preloaded; no GPU code is instrumented."""
    import tempfile
    cpp, hip = _sources()
    dev_objs = [os.path.join(OUT_DIR, os.path.basename(s) + ".o") for s in hip]
    if not all(os.path.exists(o) for o in dev_objs):
        raise RuntimeError("build_sanitized: run build() first (device objects missing)")
    out_dir = out_dir or os.path.join(tempfile.gettempdir(), "tmrx_asan_%d" % os.getuid())
    os.makedirs(out_dir, exist_ok=True)
    jobs = []
    for s in cpp:
        o = os.path.join(out_dir, os.path.basename(s) + ".o")
        jobs.append((["g++"] + CXX_FLAGS + SAN_FLAGS + ["-c", s, "-o", o], o))
    nproc = min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=nproc) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    lib = os.path.join(out_dir, "libtmrx_asan.so")
    _run(["g++", "-shared", "-fPIC"] + SAN_FLAGS + ["-o", lib + ".tmp"] +
         [o for _, o in jobs] + dev_objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-lrocblas",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(lib + ".tmp", lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        build_sanitized(verbose=True)
    else:
        build(verbose=True)
    sys.exit(0)
