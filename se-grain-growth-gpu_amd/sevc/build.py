"""Build of the native library libsevc.so (HIP kernels for gfx950 + host
C++), in-tree under se-grain-growth-gpu_amd/build/, and of the oracle used
by the tests (oracle/build/liboracle.so)."""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                      # se-grain-growth-gpu_amd
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
OUT = os.path.join(ROOT, "build")
LIB = os.path.join(OUT, "libsevc.so")
CLI = os.path.join(OUT, "sevc")

ARCH = os.environ.get("SEVC_OFFLOAD_ARCH", "gfx950")
# -ffp-contract=off: no fused multiply-add anywhere, so device FP64 results
# equal the host oracle's bit for bit.  The packed-fp32 feature is disabled
# for all device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
             "-Wno-unused-function", "-Wno-unused-variable"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]

HIP_SOURCES = ["sevc_kernels.hip", "sevc_api.hip", "sevc_adjoint.hip", "sevc_check.hip"]
CXX_SOURCES = ["sevc_host.cpp", "sevc_voronoi.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build(verbose=False, jobs=8):
    os.makedirs(OUT, exist_ok=True)
    cmds, objs = [], []
    for src in HIP_SOURCES:
        p = os.path.join(CSRC, src)
        if not os.path.exists(p):
            continue
        o = os.path.join(OUT, src + ".o")
        cmds.append(["hipcc"] + HIP_FLAGS + ["-c", p, "-o", o])
        objs.append(o)
    for src in CXX_SOURCES:
        p = os.path.join(CSRC, src)
        if not os.path.exists(p):
            continue
        o = os.path.join(OUT, src + ".o")
        cmds.append(["g++"] + CXX_FLAGS + ["-fopenmp", "-c", p, "-o", o])
        objs.append(o)
    procs = []
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        procs.append((c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((c, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        for c, out in failed:
            sys.stderr.write(" ".join(c) + "\n" + out + "\n")
        raise RuntimeError("compilation failed")
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-fopenmp", "-o", LIB] + objs, verbose)
    main = os.path.join(CSRC, "sevc_main.cpp")
    if os.path.exists(main):
        _run(["g++"] + CXX_FLAGS + ["-o", CLI, main, "-L" + OUT, "-lsevc", "-Wl,-rpath,$ORIGIN"], verbose)
    return LIB


def build_oracle(verbose=False):
    d = os.path.join(REPO, "oracle")
    _run(["make", "-C", d], verbose)
    return os.path.join(d, "build", "liboracle.so")


if __name__ == "__main__":
    build(verbose=True)
    build_oracle(verbose=True)
