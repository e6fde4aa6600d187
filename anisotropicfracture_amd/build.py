"""Build the gfx950 shared library libela.so in-tree with hipcc.

Run as ``python -m anisotropicfracture_amd.build`` (or via the repository's
``__graft_entry__.build()``).  Every call recompiles.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libela.so")
SOURCES = ["ela_kernels.hip", "ela_solver.hip", "ela_batch.hip", "ela_amr.hip", "ela_composite.hip", "ela_plotfile.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         # no packed FP32 instructions (v_pk_{add,mul,fma}_f32) anywhere in device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        text = "\n".join(l for l in out.decode().splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            sys.stderr.write(text + "\n")
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
        if verbose and text.strip():
            sys.stderr.write(text + "\n")
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_cpp_tests()
    return LIB


def build_cpp_tests():
    """The C++ interface test (tests/cpp) against include/alamo_elastic.hpp and libela.so."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "test_operator_elastic.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(root, "tests", "cpp", "test_operator_elastic")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(root, "include"), src, "-o", out,
                           "-L", HERE, "-lela", "-Wl,-rpath," + HERE])
    return out


if __name__ == "__main__":
    print(build(verbose=True))
