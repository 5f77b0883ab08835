"""Build the HIP shared library in-tree (hipcc, gfx950).

The product library ``libmpm_auxetics.so`` holds every kernel and the C ABI
declared in ``include/mpm_auxetics.h``.  Built in place next to this file so
that it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmpm_auxetics.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPM_OFFLOAD_ARCH", "gfx950")

# packed FP32 instructions disabled: ordinary FP32 code only.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
]


def sources():
    return [os.path.join(CSRC, "mpm_engine.hip"), os.path.join(CSRC, "mpm_diff.hip")]


def build(verbose=False):
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include")] + sources() + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout or "") + (r.stderr or "")
    # the host pass does not know the GPU feature flag: harmless
    out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
