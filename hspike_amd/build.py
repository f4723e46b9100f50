"""Build libhspike.so in-tree (hspike_amd/lib/) for gfx950.

C host code is compiled with gcc, the HIP kernels and their orchestration with hipcc
(--offload-arch=gfx950), and everything is linked into one shared library whose C ABI is
declared in include/*.h.  Run:  python -m hspike_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OUT_DIR = os.path.join(HERE, "lib")
OUT = os.path.join(OUT_DIR, "libhspike.so")

C_SOURCES = ["spike_matrix.c", "spike_interfaces.c", "spike_backend.c", "spike_gen.c", "spike_rcm.c", "spike_memory.c"]
HIP_SOURCES = ["spike_core.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HSPIKE_ARCH", "gfx950")
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC, "-I" + CSRC,
    # packed FP32 instructions are excluded from generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# the C sources call the HIP runtime API only (pinned host memory), hence the platform define
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-fopenmp", "-Wall", "-Wno-unused-function", "-I" + INC,
           "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + p.stdout)
        raise RuntimeError("build step failed: %s" % os.path.basename(cmd[-1]))
    return p.stdout


def build(verbose=False, jobs=None):
    """Compile every source and link libhspike.so; returns its path."""
    os.makedirs(OUT_DIR, exist_ok=True)
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    cmds = []
    for src in C_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        cmds.append(["gcc"] + C_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((c, out))
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        for c, out in failed:
            sys.stderr.write(" ".join(c) + "\n" + out + "\n")
        raise RuntimeError("compilation failed: %s" % ", ".join(os.path.basename(c[-1]) for c, _ in failed))
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-fopenmp", "-lm"])
    os.replace(tmp, OUT)
    shutil.rmtree(bdir, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
