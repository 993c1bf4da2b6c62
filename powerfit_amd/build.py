"""Build libpowerfit_hip.so (gfx950) and the host-side FFT check library in-tree
(one compiler process per HIP source, in parallel, then one link).

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

HIP_SOURCES = ["pf_kernels.hip", "pf_fused.hip", "pf_planes.hip", "pf_planes_odd.hip",
               "pf_zct.hip", "pf_api.hip", "pf_shapes.hip", "pf_peaks.hip", "pf_split.hip"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return proc


def build(verbose=False):
    """Compile each HIP source to an object in parallel, then link the library."""
    objdir = os.path.join(HERE, "_objs")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-c", *COMMON, *NO_PACKED_FP32,
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                 stderr=subprocess.PIPE, text=True)))
    objs = []
    failed = None
    for cmd, obj, proc in procs:
        out, err = proc.communicate()
        if proc.returncode != 0 and failed is None:
            sys.stderr.write(out + err)
            failed = cmd
        objs.append(obj)
    if failed is not None:
        raise RuntimeError("build step failed: " + " ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    # host build of the FFT stage code (CPU checks of the device butterflies)
    tmp = HOST_LIB + ".tmp"
    _run([HIPCC, "-shared", *COMMON, os.path.join(CSRC, "pf_fft_host.cpp"), "-o", tmp])
    os.replace(tmp, HOST_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
