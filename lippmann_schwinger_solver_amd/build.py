"""Build the in-tree shared library liblsg.so (HIP kernels for gfx950 + host C++).

Host numerics (tables, tree, NCA) are compiled with g++ and OpenMP; the HIP
kernels and the C ABI with hipcc. The library is written next to this file so
that it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblsg.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HOST_SRC = ["tables.cpp", "tree.cpp", "nca.cpp", "precomp.cpp"]
HIP_SRC = ["device.hip", "block.hip", "ragged.hip", "learn.hip", "nca_device.hip", "field.hip", "capi.cpp"]
# device code that restates the host numerics keeps their rounding (no contraction)
HIP_NO_CONTRACT = {"nca_device.hip"}
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def _jobs(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)) for c in cmds]
    err = None
    for c, p in procs:
        out, e = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + e)
            err = err or ("build step failed: " + " ".join(c))
    if err:
        raise RuntimeError(err)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs, cmds = [], []
    inc = ["-I", os.path.join(HERE, "..", "include")]
    for s in HOST_SRC:
        o = os.path.join(BUILD, s + ".o")
        cmds.append(["g++"] + HOST_FLAGS + inc + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRC:
        o = os.path.join(BUILD, s + ".o")
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        # device kernels may fuse multiply-add; host code in the C ABI keeps
        # the host numerics' rounding (no contraction)
        contract = ["-ffp-contract=fast"] if s.endswith(".hip") and s not in HIP_NO_CONTRACT else ["-ffp-contract=off"]
        cmds.append([HIPCC] + HIP_FLAGS + contract + inc + lang + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    if verbose:
        for c in cmds:
            print(" ".join(c))
    _jobs(cmds)
    tmp = LIB + ".tmp"
    # Link with g++ so that complex multiply/divide helpers (__muldc3,
    # __divdc3) come from libgcc, the same ones the host numerics and the
    # oracle use, instead of compiler-rt's variants.
    libdir = os.path.join(ROCM, "lib")
    _run(["g++", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-fopenmp", "-L" + libdir, "-lamdhip64", "-Wl,-rpath," + libdir])
    os.replace(tmp, LIB)
    # reference-style driver (12 positional inputs, result/ files)
    exe = os.path.join(HERE, "lsg_solve")
    _run(["g++", "-O2", "-std=c++17", "-o", exe + ".tmp", os.path.join(HERE, "cpp", "lsg_main.cpp"),
          "-L" + HERE, "-llsg", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
