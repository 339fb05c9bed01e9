"""Build libsqc_hip.so for gfx950 in-tree (hipcc, no JIT cache).

``python -m sqcalc_amd.build`` or ``build_library()``.  The library lands
next to this file so that it travels with the source tree.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsqc_hip.so")
ARCH = os.environ.get("SQC_HIP_ARCH", "gfx950")

# Packed-FP32 instructions are kept out of every kernel (the compiler may
# otherwise form v_pk_* from scalar float code).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library(verbose=False, jobs=2):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    common = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(HERE, "..", "include")]
    units = [
        ("sqc_kernels.hip", ["-x", "hip", "--offload-arch=" + ARCH] + DEVICE_FLAGS),
        ("sqc_host.cpp", ["-x", "hip", "--offload-arch=" + ARCH, "-ffp-contract=off"]
         + DEVICE_FLAGS),
        # the matrix-core accumulators stay in VGPRs (no AGPR copies per MFMA)
        ("sqc_dyn.hip", ["-x", "hip", "--offload-arch=" + ARCH, "-mllvm", "-amdgpu-mfma-vgpr-form"]
         + DEVICE_FLAGS),
        # pair displacements are evaluated in the reference's operation order
        ("sqc_debye.hip", ["-x", "hip", "--offload-arch=" + ARCH, "-ffp-contract=off"]
         + DEVICE_FLAGS),
        ("sqc_dump.cpp", ["-x", "c++", "-ffp-contract=off", "-pthread"]),
        ("sqc_elements.cpp", ["-x", "c++"]),
    ]
    procs = []
    for src, extra in units:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc] + common + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    for pr in procs:
        if pr.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "-pthread", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build_library(verbose="-v" in sys.argv)
