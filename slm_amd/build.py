"""Build the native libraries in-tree.

  slm_amd/libslm_hip.so   hipcc, gfx950 kernels + C ABI (include/slm_hip.h)
  slm_amd/libslm_host.so  gcc, host helpers (sink filling for synthetic DTMs, subsegment colouring)

Run ``python -m slm_amd.build``.  Every build recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SLM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-shared", "-std=c++17",
    # no implicit contraction: fused multiply-adds only where the code asks
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

HIP_SOURCES = ["slm_hip.hip", "kde.hip", "mapping.hip", "prep.hip", "surrogate.hip", "hsl.hip"]


def _run(cmd):
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(out=None):
    out = out or os.path.join(HERE, "libslm_hip.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + srcs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_host(out=None):
    out = out or os.path.join(HERE, "libslm_host.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11", os.path.join(CSRC, "dtm_fill.c"),
          os.path.join(CSRC, "host_graph.c"), "-o", tmp, "-lm"])
    os.replace(tmp, out)
    return out


def build_oracle():
    """The CPU oracle (test infrastructure) under oracle/."""
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_host()
    build_hip()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host, "oracle": build_oracle}[w]()
