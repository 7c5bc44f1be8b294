"""Build the oracle's C restatement (oracle/liboracle.so) and, when the
reference sources are present, oracle/_ref/libtmr_octant_ref.so.

Test infrastructure only: nothing in the product path loads these.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_SRC = "/root/reference/src"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))


def build_oracle():
    srcs = [os.path.join(HERE, s) for s in sorted(os.listdir(HERE)) if s.endswith(".c")]
    tmp = LIB + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-o", tmp] + srcs + ["-lm"])
    os.replace(tmp, LIB)
    return LIB


def build_ref():
    """Compile the reference's octant algebra (TMROctant.cpp, TMRBase.cpp)
    from where it lies, with the MPI headers of this image; returns the path
    or None when the reference is absent."""
    if not os.path.isdir(REF_SRC):
        return None
    mpi_inc, mpi_lib = "/opt/conda/include", "/opt/conda/lib"
    if not os.path.exists(os.path.join(mpi_inc, "mpi.h")):
        return None
    os.makedirs(REF_DIR, exist_ok=True)
    out = os.path.join(REF_DIR, "libtmr_octant_ref.so")
    srcs = [os.path.join(REF_SRC, "TMROctant.cpp"), os.path.join(REF_SRC, "TMRBase.cpp"),
            os.path.join(HERE, "ref_octant_driver.cpp")]
    _run(["g++", "-O2", "-std=gnu++11", "-fPIC", "-shared", "-I" + REF_SRC, "-I" + mpi_inc,
          "-o", out] + srcs + ["-L" + mpi_lib, "-Wl,-rpath," + mpi_lib, "-lmpi"])
    return out


def build():
    build_oracle()
    try:
        build_ref()
    except RuntimeError as e:  # the reference octant build is optional
        print("oracle/_ref not built:", str(e).splitlines()[0])


if __name__ == "__main__":
    build()
    print("built", LIB)
