"""Build recipe of the HIP library ``libaldvc_hip.so`` (gfx950), in place next to this file.

    python -m pyaldvc_amd.build

All device code is compiled with packed-FP32 instruction selection disabled
(``-target-feature -packed-fp32-ops``): ordinary float code is left to the scalar FP32 path.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libaldvc_hip.so")
SOURCES = [
    "aldvc_icgn.hip",
    "aldvc_ncc.hip",
    "aldvc_split.hip",
    "aldvc_volume.hip",
    "aldvc_global.hip",
    "aldvc_warp.hip",
    "aldvc_admm.hip",
    "aldvc_uncertainty.hip",
]
ARCH = os.environ.get("ALDVC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def command(out: str = LIB) -> list[str]:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    return [
        hipcc(),
        "-O3",
        f"--offload-arch={ARCH}",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-o",
        out,
        *srcs,
    ]


def build(verbose: bool = False) -> str:
    """Compile the library (always) and return its path."""
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
