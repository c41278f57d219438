"""In-tree build of the package's native libraries.

  libfastpart_host.so  g++    csrc/fp_host.cpp         (host helpers)
  libfastpart.so       hipcc  csrc/fp_{context,tv,cn,vol,conn,metrics,batch,serve,bfs}.hip  (HIP engine, gfx950)
  libfastpart_legacy_<m>.so  g++  csrc/fp_legacy.cpp  (the reference's partition_gpu_wrapper
                       for objective <m>, over libfastpart.so)

`python -m fastpartitioner_amd.build` builds all of them.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("FP_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# -ffp-contract=off: the LDG score must use exactly the reference's IEEE double
# operations; packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-fPIC", "-shared", "-std=c++17", "-fopenmp", "-ffp-contract=off",
             f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


ENGINE_SOURCES = ["fp_context.hip", "fp_tv.hip", "fp_cn.hip", "fp_vol.hip", "fp_conn.hip", "fp_metrics.hip",
                  "fp_batch.hip", "fp_serve.hip", "fp_bfs.hip"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:2])} ... rc={p.returncode}")


def build_host():
    out = os.path.join(HERE, "libfastpart_host.so")
    _run([CXX, "-O3", "-fPIC", "-shared", "-fopenmp", "-std=c++17", "-Wall",
          "-o", out, os.path.join(CSRC, "fp_host.cpp")])
    return out


def build_engine():
    out = os.path.join(HERE, "libfastpart.so")
    srcs = [os.path.join(CSRC, f) for f in ENGINE_SOURCES]
    _run([HIPCC, *HIP_FLAGS, "-o", out, *srcs])
    return out


LEGACY_OBJECTIVES = ["tv", "cn", "msv", "msrv", "tm", "msm"]


def build_legacy(objective):
    out = os.path.join(HERE, f"libfastpart_legacy_{objective}.so")
    _run([CXX, "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c++17", "-Wall", f"-DFP_LEGACY_{objective.upper()}",
          "-o", out, os.path.join(CSRC, "fp_legacy.cpp"), f"-L{HERE}", "-lfastpart", "-lfastpart_host",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_all():
    return [build_host(), build_engine()] + [build_legacy(m) for m in LEGACY_OBJECTIVES]


if __name__ == "__main__":
    for p in build_all():
        print(p)
