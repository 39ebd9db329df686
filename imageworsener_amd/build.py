"""Build libimageworsener_amd.so (host C++ + gfx950 HIP kernels) in-tree.

The library is a plain C-ABI shared object (include/imagew.h, include/iwamd.h);
Python only drives the compiler here.  Flags that matter for byte parity:
  -ffp-contract=off                 no fused multiply-add anywhere
  -fno-gpu-flush-denormals-to-zero  FP32 intermediate keeps subnormals
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                    no packed FP32 instructions
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libimageworsener_amd.so")
BUILD_DIR = os.path.join(ROOT, "_build", "lib")

SOURCES = ["iw_api.cpp", "iw_prepare.cpp", "iw_finish.cpp", "iw_tables.cpp", "iw_pnm.cpp", "iw_miff.cpp",
           "iw_gpu.hip"]

ARCH = os.environ.get("IWAMD_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
          "-I" + INCLUDE, "-I" + CSRC, "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value"]
HIPFLAGS = ["--offload-arch=" + ARCH, "-fno-gpu-flush-denormals-to-zero",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            raise RuntimeError("missing source: " + path)
        obj = os.path.join(BUILD_DIR, src + ".o")
        cmd = [cc] + COMMON + ["-c", path, "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
            cmd += HIPFLAGS
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([cc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-Wl,-Bsymbolic", "-lm", "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
