"""Build recipe of libqss.so (HIP, gfx950) — in-tree, no JIT cache.

``python -m tertia_amd.build`` compiles every source under ``csrc/`` with
hipcc for gfx950 and links ``tertia_amd/libqss.so``.  The particle, beam and
spectral kernels are compiled with ``-ffp-contract=off`` so that their
arithmetic is evaluated in the reference's operation order (no FMA fusion);
the transform GEMMs keep contraction (MFMA accumulation has its own order)
and keep their MFMA accumulators in VGPRs.
All HIP code is compiled without packed-FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqss.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QSS_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-command-line-argument"]
NO_CONTRACT = {"particles.hip", "push_tile.hip", "beam.hip", "spectral.hip", "adjoint.hip",
               "driver.hip"}
# MFMA accumulators in VGPRs: with AGPR accumulators the compiler copies
# every loop-carried accumulator AGPR->VGPR->AGPR at the K-loop back-edge,
# draining the MFMA pipeline once per iteration
MFMA_VGPR = {"gemm.hip", "fact_fused.hip"}
SOURCES = ["gemm.hip", "spectral.hip", "beam.hip", "particles.hip", "push_tile.hip",
           "adjoint.hip", "driver.hip", "noalloc.hip", "fact.hip", "fact_fused.hip", "qss_api.cpp",
           "h5file.cpp", "save.cpp"]


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, src.rsplit(".", 1)[0] + ".o")
    flags = list(COMMON)
    if src in NO_CONTRACT:
        flags.append("-ffp-contract=off")
    if src in MFMA_VGPR:
        flags += ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]
    if src.endswith(".cpp"):
        flags = ["-x", "hip"] + flags
    cmd = [HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
