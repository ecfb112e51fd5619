"""Build recipes for the native libraries (run on the build host; artefacts stay in-tree).

* ``rabbitqcplus_amd/_lib/libcare_hip.so`` - HIP kernels for gfx950 plus the C-ABI
  declared in ``include/care_hip.h``.
* ``rabbitqcplus_amd/_lib/libcaresim.so``  - the seeded read simulator
  (``include/care_sim.h``).

The oracle libraries under ``oracle/`` are built by ``oracle/Makefile``; this module
only calls that Makefile (``build_oracle``), it does not link them.
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rabbitqcplus_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OFFLOAD_ARCH = "gfx950"

# Device code flags:
#  -ffp-contract=off : no a*b+c -> fma contraction, so float results match the
#                      reference's x86 build operation for operation.
#  -packed-fp32-ops  : disabled (packed FP32 VALU forms are avoided on this target).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={OFFLOAD_ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{INCLUDE}",
]

HIP_SOURCES = ["care_hip.hip", "mlp.hip"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_sim():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libcaresim.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", f"-I{INCLUDE}",
          os.path.join(CSRC, "readsim.cpp"), "-o", out])
    return out


def build_hip():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libcare_hip.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([HIPCC] + HIP_FLAGS + srcs + ["-o", out])
    return out


def build_oracle(with_reference: bool | None = None):
    """Builds oracle/_build/libcarecpu.so, and oracle/_ref/libcareref.so when the
    reference sources are present (they are only needed to build it)."""
    odir = os.path.join(ROOT, "oracle")
    jobs = str(min(8, os.cpu_count() or 1))
    _run(["make", "-j", jobs, "-C", odir, "carecpu"])
    if with_reference is None:
        with_reference = os.path.isdir("/root/reference/src")
    if with_reference:
        _run(["make", "-j", jobs, "-C", odir, "ref"])
        # the reference-side drop-in, linked to libcare_hip.so (built before this step)
        if os.path.exists(os.path.join(LIBDIR, "libcare_hip.so")):
            _run(["make", "-j", jobs, "-C", odir, "dropin"])


def build_all():
    build_sim()
    build_hip()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"sim": build_sim, "hip": build_hip, "oracle": build_oracle, "all": build_all}[w]()
