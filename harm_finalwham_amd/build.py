"""Build libharmgpu.so (HIP for gfx950) in-tree.

    python -m harm_finalwham_amd.build        (or build() from Python)

Flags: FP64 code must not be contracted into FMA (-ffp-contract=off) so that
the kernels round exactly like the reference's x86-64 build; packed FP32
instructions are disabled for the whole library.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libharmgpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HG_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(REPO, "include"), "-I", os.path.join(HERE, "csrc"),
]


def sources():
    c = os.path.join(HERE, "csrc")
    return sorted(glob.glob(os.path.join(c, "*.hip")) + glob.glob(os.path.join(c, "*.cpp")))


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + sources() + ["-o", tmp]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libharmgpu.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
