"""Build libpowerfit_hip.so (gfx950) and the host-side FFT check library in-tree.

    python -m powerfit_amd.build

Every HIP translation unit is compiled with packed-FP32 instructions
disabled (``-target-feature -packed-fp32-ops``).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpowerfit_hip.so")
HOST_LIB = os.path.join(HERE, "libpf_fft_host.so")
ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["pf_kernels.hip", "pf_fused.hip", "pf_planes.hip", "pf_api.hip", "pf_shapes.hip",
               "pf_peaks.hip", "pf_split.hip"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return proc


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", *COMMON, *NO_PACKED_FP32,
           *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB)
    # host build of the FFT stage code (CPU checks of the device butterflies)
    tmp = HOST_LIB + ".tmp"
    _run([HIPCC, "-shared", *COMMON, os.path.join(CSRC, "pf_fft_host.cpp"), "-o", tmp])
    os.replace(tmp, HOST_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
