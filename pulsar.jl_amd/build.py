"""Builds libpulsar_rocm.so in-tree for gfx950 with hipcc.

    python pulsar.jl_amd/build.py [--debug]

The library is written next to the Python binding
(pulsar.jl_amd/pulsar_rocm/libpulsar_rocm.so) so that it travels with the
source tree; nothing is installed.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "pulsar_rocm", "libpulsar_rocm.so")
SOURCES = ["pulsar_rocm.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(debug: bool = False, verbose: bool = False, out: str = OUT, defines=(), sources=None) -> str:
    flags = [
        "--offload-arch=gfx950",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-O1" if debug else "-O3",
        # packed FP32 VALU instructions stay disabled for all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # MFMA accumulators in VGPRs (no v_accvgpr_read copies on gfx950)
        "-mllvm", "-amdgpu-mfma-vgpr-form",
        "-Wall",
        "-Wno-unused-function",
        *[f"-D{d}" for d in defines],
    ]
    tmp = out + ".tmp"
    cmd = [HIPCC, *flags, "-o", tmp, *[os.path.join(CSRC, s) for s in (sources or SOURCES)]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines() if noise not in ln]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    args = sys.argv[1:]
    out = args[args.index("--out") + 1] if "--out" in args else OUT
    defs = [a[2:] for a in args if a.startswith("-D")]
    print(build(debug="--debug" in args, verbose=True, out=out, defines=defs))
