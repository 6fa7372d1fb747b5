"""Builds libclarabel_hip.so in-tree for gfx950 (hipcc for the kernels, the
host C++ compiled by the same toolchain, one shared library)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libclarabel_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CLH_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # packed FP32 instructions are not generated for this library
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


# The interior-point kernel is tuned for four resident workgroups of four
# wavefronts per CU (128 VGPRs); a change that makes the compiler give it more
# registers (for instance an engine function emitted out of line for a second
# kernel of the same file) quarters the resident slots.  Fail the build then.
KERNEL_MAX_VGPRS = {"ipm_kernel": 128, "resident_kernel": 128, "grid_ipm_kernel": 128}


def _check_registers(remarks):
    name = None
    for line in remarks.splitlines():
        if "Function Name:" in line:
            name = line.split("Function Name:")[1].split()[0]
        elif "VGPRs:" in line and "AGPRs" not in line and name:
            v = int(line.split("VGPRs:")[1].split()[0])
            for k, cap in KERNEL_MAX_VGPRS.items():
                if k in name and v > cap:
                    raise RuntimeError(f"{k} uses {v} VGPRs (budget {cap}): its occupancy would drop")


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in ["clh_symbolic.cpp", "clh_runtime.cpp"]:
        o = os.path.join(objdir, src.replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in ["clh_backend_hip.hip", "clh_lanes_hip.hip", "clh_cones_hip.hip", "clh_grid_hip.hip"]:
        o = os.path.join(objdir, src.replace(".hip", ".o"))
        out = _run([HIPCC, *HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-c", os.path.join(CSRC, src),
                    "-o", o])
        _check_registers(out)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
