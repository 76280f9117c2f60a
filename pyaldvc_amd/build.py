"""Build recipe of the HIP library ``libaldvc_hip.so`` (gfx950), in place next to this file.

    python -m pyaldvc_amd.build

Every source is compiled to its own object in parallel (``_build/``), then linked. All device
code is compiled with packed-FP32 instruction selection disabled (``-target-feature
-packed-fp32-ops``): ordinary float code is left to the scalar FP32 path.

The library carries a hash of the sources it was built from (``aldvc_source_hash()``); the loader
compares it with the sources on disk, so a library built from other sources is refused.
"""

from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaldvc_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "aldvc_hip.h")
SOURCES = [
    "aldvc_icgn.hip",
    "aldvc_ncc.hip",
    "aldvc_split.hip",
    "aldvc_volume.hip",
    "aldvc_global.hip",
    "aldvc_warp.hip",
    "aldvc_admm.hip",
    "aldvc_uncertainty.hip",
    "aldvc_admm_loop.hip",
]
# compiled with the source hash defined; holds nothing else
INFO_SOURCE = "aldvc_build_info.hip"
ARCH = os.environ.get("ALDVC_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3",
    f"--offload-arch={ARCH}",
    "-std=c++17",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def source_hash(csrc: str | None = None) -> str:
    """SHA-256 (first 16 hex digits) over the names and bytes of the sources under ``csrc`` and of
    the public header the sources include."""
    csrc = csrc or CSRC
    h = hashlib.sha256()
    paths = [os.path.join(csrc, name) for name in sorted(os.listdir(csrc))] + [HEADER]
    for path in paths:
        name = os.path.basename(path)
        if not os.path.isfile(path) or not name.endswith((".hip", ".h")):
            continue
        h.update(name.encode())
        h.update(b"\0")
        with open(path, "rb") as fh:
            h.update(fh.read())
        h.update(b"\0")
    return h.hexdigest()[:16]


def _compile(src: str, obj: str, defines: list[str]) -> tuple[str, subprocess.CompletedProcess]:
    cmd = [hipcc(), *FLAGS, *defines, "-c", "-o", obj, os.path.join(CSRC, src)]
    return " ".join(cmd), subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source (always) and link the library; return its path."""
    os.makedirs(OBJ, exist_ok=True)
    digest = source_hash()
    units = [(s, []) for s in SOURCES] + [(INFO_SOURCE, [f'-DALDVC_SOURCE_HASH="{digest}"'])]
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s, _ in units]
    jobs = jobs or min(len(units), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(lambda u: _compile(u[0], u[1], u[2]), [(s, o, d) for (s, d), o in zip(units, objs)]))
    failed = False
    for cmd, res in results:
        if verbose:
            print(cmd, flush=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    link = [hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", "-o", tmp, *objs]
    if verbose:
        print(" ".join(link), flush=True)
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"link failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
