"""Launch-shape autotuner of the fused IC-GN kernel (12-DOF solves and translation-only solves).

The fused kernel has launch choices that do not change what it computes, only how fast: the LDS
budget of a workgroup's brick (three workgroups per CU with 7 planes, or two with 11), the
voxel-to-lane mapping of the brick path (8 x 8 patches per wave, plane pairs of 32-64-voxel rows,
or sampling order), the node order (spatial tiles or index order), the XCD-aware workgroup ->
node mapping, and bricks versus memory gathers. Which is fastest depends
on the subset shape, the interpolation, the mask/split selection and the deformation (how many
chunks fit a brick). ``choose_icgn12`` / ``choose_icgn3`` time every candidate on a contiguous run of the call's
valid nodes (neighbouring nodes, so that subset overlap and L2 reuse are as in the full call;
enough of them to fill the chip several times), best of three, and keeps the fastest unless the
default is within 3 % of it; the choice is cached per (solve kind, device, subset shape, stride,
interpolation, gradient storage, selection).
The cache lives in memory and in a JSON file (``$PYALDVC_AMD_TUNE_CACHE``, default
``~/.cache/pyaldvc_amd/tune.json``), so a later process starts from the tuned shape.

Candidates can round float32 lane sums differently (a different voxel-to-lane assignment), so a
tuned call agrees with the default to the kernels' own parity bar, not bit for bit; a cached
choice is stable from call to call.
"""

from __future__ import annotations

import json
import os
import threading
import time
from dataclasses import asdict, dataclass

import numpy as np
import torch

TUNE_NO_PAIRS, TUNE_INDEX_ORDER, TUNE_NO_XCD, TUNE_NO_PATCHES = 1, 2, 4, 8
SAMPLE_NODES = 16384  # nodes timed per candidate: 256 CUs x 3 workgroups, 20 times over
MARGIN = 0.03  # a candidate replaces the default only when it is this much faster
REPEATS = 3


@dataclass(frozen=True)
class Tuning:
    variant: int  # 1: fused with LDS bricks where a chunk fits, 2: fused with memory gathers only
    lds_budget: int = 0  # LDS bytes per workgroup for the brick (0: three workgroups per CU)
    flags: int = 0  # TUNE_* bits

    def label(self) -> str:
        parts = ["bricks" if self.variant == 1 else "gathers"]
        if self.lds_budget:
            parts.append(f"lds{self.lds_budget // 1024}k")
        for bit, name in ((TUNE_NO_PATCHES, "no-patches"), (TUNE_NO_PAIRS, "no-pairs"), (TUNE_INDEX_ORDER, "index-order"),
                          (TUNE_NO_XCD, "no-xcd")):  # fmt: skip
            if self.flags & bit:
                parts.append(name)
        return "+".join(parts)


DEFAULT = Tuning(1)
CANDIDATES = (
    DEFAULT,  # bricks, 8 x 8 lane patches
    Tuning(1, 0, TUNE_NO_PATCHES),  # plane pairs for 32-64-voxel rows
    Tuning(1, 0, TUNE_NO_PATCHES | TUNE_NO_PAIRS),  # voxels in sampling order
    Tuning(1, 80 * 1024),  # two workgroups per CU, 11 planes per brick
    Tuning(1, 0, TUNE_NO_XCD),
    Tuning(1, 0, TUNE_INDEX_ORDER),
    Tuning(2),
)

_lock = threading.Lock()
_memo: dict = {}
_loaded = False
last_report: dict = {}  # the most recent tuning: key, per-candidate seconds, choice


def cache_path() -> str:
    return os.environ.get("PYALDVC_AMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "pyaldvc_amd", "tune.json"))


def _load() -> None:
    global _loaded
    if _loaded:
        return
    _loaded = True
    try:
        with open(cache_path()) as fh:
            for k, v in json.load(fh).items():
                _memo[k] = Tuning(**v)
    except (OSError, ValueError, TypeError):
        pass


def _save() -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as fh:
            json.dump({k: asdict(v) for k, v in _memo.items()}, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: the choice stays in memory


def clear(memory_only: bool = False) -> None:
    """Forget the tuned shapes (and delete the cache file unless ``memory_only``)."""
    global _loaded
    with _lock:
        _memo.clear()
        _loaded = memory_only
        if not memory_only:
            try:
                os.remove(cache_path())
            except OSError:
                pass


def _key(kind, hx, hy, hz, stride, mode, stored, selected) -> str:
    name = torch.cuda.get_device_name(torch.cuda.current_device()).replace(" ", "_")
    return f"{kind}|{name}|{hx}x{hy}x{hz}|s{stride}|m{mode}|{'stored' if stored else 'otf'}|{'sel' if selected else 'all'}"


def _sample(valid, n: int) -> torch.Tensor:
    """Indices of ``n`` consecutive valid nodes from the middle of the node order (all of them
    when fewer): neighbours in the lattice, whose subsets overlap as in the full call."""
    idx = torch.nonzero(valid.to(torch.bool)).flatten()
    if idx.numel() <= n:
        return idx
    start = (idx.numel() - n) // 2
    return idx[start : start + n]


def _tune(key: str, nodes: int, run) -> Tuning:
    """Times ``run(candidate)`` for every candidate (best of REPEATS) and caches the choice."""
    times = {}
    for cand in CANDIDATES:
        best = float("inf")
        for _ in range(REPEATS):  # the first run also loads the code object
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run(cand)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        times[cand.label()] = best
    choice = min(CANDIDATES, key=lambda c: times[c.label()])
    if times[choice.label()] > (1.0 - MARGIN) * times[DEFAULT.label()]:
        choice = DEFAULT
    with _lock:
        _memo[key] = choice
        _save()
    last_report.clear()
    last_report.update(key=key, nodes=nodes, seconds=times, choice=choice.label())
    return choice


def _cached(key: str):
    with _lock:
        _load()
        return _memo.get(key)


def choose_icgn12(args, kw) -> Tuning:
    """The tuned launch shape for ``kernels.icgn_12dof_parallel(*args, **kw)`` (see the module
    docstring); times the candidates on first use of a key."""
    from . import kernels as K

    coords, P0, hx, hy, hz, f, gx, gy, gz, mask, g, mode, L_all, meanf, bottomf, valid = args[:16]
    rest = args[16:]
    fd = K._Args.vol(f)
    stored = gx is not None and tuple(gx.shape) == tuple(fd.shape)
    selected = (mask is not None and tuple(mask.shape) == tuple(fd.shape)) or kw.get("split_index") is not None
    key = _key("icgn12", int(hx), int(hy), int(hz), int(kw.get("stride", 1)), int(mode), stored, selected)
    hit = _cached(key)
    if hit is not None:
        return hit

    A = K._Args(fd)
    vd = A.arr(valid, np.uint8)
    sel = _sample(vd, SAMPLE_NODES)
    if sel.numel() == 0:
        return DEFAULT
    take = lambda x, dt: None if x is None else A.arr(x, dt)[sel].contiguous()  # noqa: E731
    sub_args = (take(coords, np.int64), take(P0, np.float64), hx, hy, hz, fd, gx, gy, gz, mask, g, mode,
                take(L_all, np.float64), take(meanf, np.float64), take(bottomf, np.float64), vd[sel].contiguous(), *rest)  # fmt: skip
    sub_kw = dict(kw)
    if kw.get("H_all") is not None:
        sub_kw["H_all"] = take(kw["H_all"], np.float64)
    if kw.get("split_index") is not None:
        sub_kw["split_index"] = take(kw["split_index"], np.int64)
    return _tune(key, int(sel.numel()), lambda cand: K.icgn_12dof_parallel(*sub_args, **sub_kw, tuning=cand))


def choose_icgn3(args, kw) -> Tuning:
    """The tuned launch shape for ``kernels.icgn_3dof_parallel(*args, **kw)``: the same candidates
    as the 12-DOF solve, timed on the translation-only kernel (its few iterations per node weigh
    brick staging against gathers differently), cached under its own key."""
    from . import kernels as K

    (coords, U_old, F_fixed, vdual, hx, hy, hz, f, gx, gy, gz, mask, g, mode, H_all, meanf, bottomf,
     valid) = args[:18]  # fmt: skip
    rest = args[18:]
    fd = K._Args.vol(f)
    stored = gx is not None and tuple(gx.shape) == tuple(fd.shape)
    selected = (mask is not None and tuple(mask.shape) == tuple(fd.shape)) or kw.get("split_index") is not None
    key = _key("icgn3", int(hx), int(hy), int(hz), int(kw.get("stride", 1)), int(mode), stored, selected)
    hit = _cached(key)
    if hit is not None:
        return hit

    A = K._Args(fd)
    vd = A.arr(valid, np.uint8)
    sel = _sample(vd, SAMPLE_NODES)
    if sel.numel() == 0:
        return DEFAULT
    take = lambda x, dt: None if x is None else A.arr(x, dt)[sel].contiguous()  # noqa: E731
    sub_args = (take(coords, np.int64), take(U_old, np.float64), take(F_fixed, np.float64), take(vdual, np.float64),
                hx, hy, hz, fd, gx, gy, gz, mask, g, mode, take(H_all, np.float64), take(meanf, np.float64),
                take(bottomf, np.float64), vd[sel].contiguous(), *rest)  # fmt: skip
    sub_kw = dict(kw)
    if kw.get("split_index") is not None:
        sub_kw["split_index"] = take(kw["split_index"], np.int64)
    return _tune(key, int(sel.numel()), lambda cand: K.icgn_3dof_parallel(*sub_args, **sub_kw, tuning=cand))
