"""Build of the HIP library (libupf.so, gfx950) and of the CPU oracle.

    python -m unifiedparticleframeworkcuda_amd.build

The library is compiled in-tree with hipcc: every .hip / .cpp file under
csrc/ is one translation unit, linked into unifiedparticleframeworkcuda_amd/libupf.so.
Floating point: -ffp-contract=off and correctly rounded division / sqrt keep
the reference's operation-by-operation arithmetic; packed FP32 instructions
are disabled.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libupf.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("UPF_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC",
    "--offload-arch=" + ARCH,
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, cwd=None):
    sys.stdout.write("+ " + " ".join(cmd) + "\n")
    sys.stdout.flush()
    subprocess.check_call(cmd, cwd=cwd)


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def build_library(jobs=8):
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + HIP_FLAGS + inc + lang + ["-c", src, "-o", obj]
        sys.stdout.write("+ " + " ".join(cmd) + "\n")
        procs.append((subprocess.Popen(cmd), src))
        if len(procs) >= jobs:
            p, s = procs.pop(0)
            if p.wait() != 0:
                raise RuntimeError("compile failed: " + s)
    for p, s in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed: " + s)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_all():
    build_library()
    build_oracle()


if __name__ == "__main__":
    build_all()
