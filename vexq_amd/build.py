"""Builds libvexqgpu.so (gfx950) in-tree with hipcc.

    python -m vexq_amd.build

Compiles every HIP source under vexq_amd/csrc into one shared library with a
plain C ABI (include/vexq_gpu.h).  Flags that matter for results:
  -ffp-contract=off   every float64 op rounds on its own, as Go's materialised
                      vector ops do (no fused multiply-add);
  -packed-fp32-ops off  no packed FP32 instructions are generated.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvexqgpu.so")
SOURCES = ["vq_api.hip", "k_scan.hip", "k_join.hip", "vq_jit.hip", "vq_ingest.hip"]
JIT_TEMPLATE = "jit_template.hip"
JIT_HEADER = "jit_template_src.h"  # generated: the template as a string literal

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VQ_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def embed_template() -> None:
    """Writes csrc/jit_template_src.h: fixchunks.h and jit_template.hip as C++
    raw strings (the run-time source is the prologue, then both)."""
    with open(os.path.join(CSRC, JIT_TEMPLATE)) as f:
        text = f.read()
    with open(os.path.join(CSRC, "fixchunks.h")) as f:
        chunks = f.read()
    assert ")VQJIT\"" not in text and ")VQJIT\"" not in chunks
    out = ("// generated by vexq_amd/build.py from fixchunks.h and jit_template.hip; do not edit\n"
           "#pragma once\n"
           f"static const char kFixChunksSrc[] = R\"VQJIT({chunks})VQJIT\";\n"
           f"static const char kJitTemplateSrc[] = R\"VQJIT({text})VQJIT\";\n")
    path = os.path.join(CSRC, JIT_HEADER)
    if not os.path.exists(path) or open(path).read() != out:
        with open(path, "w") as f:
            f.write(out)


def check_template(verbose: bool = False) -> None:
    """Compiles the run-time template once with its built-in default shape, so
    a syntax error surfaces at build time rather than at the first query."""
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O1", "-std=c++17", "-c", "--cuda-device-only",
           "-ffp-contract=off", "-o", os.devnull, os.path.join(CSRC, JIT_TEMPLATE)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("run-time kernel template does not compile")


def build(verbose: bool = False) -> str:
    embed_template()
    check_template(verbose)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs, "-lhiprtc"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
