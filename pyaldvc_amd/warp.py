"""Differentiable warp residual: the ZNSSD of every subset under a displacement field on a coarse
hex8 node grid, with its backward pass, as a ``torch.autograd.Function``.

The forward is the reference's ZNCC of a warp (solver/numba_kernels.py:1354-1387,
``evaluate_zncc_parallel``) with each subset's affine warp replaced by the trilinear FE
interpolation of node displacements (mesh/hex8.py:58-61); per subset it returns
``C = sum (f~ - g~)^2 = 2 (1 - ZNCC)``. The backward (the reference has none) gives ``dL/du``
for the node displacements and ``dL/dg`` for the deformed volume, from the HIP kernels of
``csrc/aldvc_warp.hip`` with interpolation weights recomputed rather than stored; its sums run in
a fixed order, so the gradients are bit-identical from run to run.
"""

from __future__ import annotations

import numpy as np
import torch

from . import kernels as K
from ._lib import check, lib

NSTAT = 8  # n, m_f, b_f, m_g, b_g, ZNCC, valid, C


def _geometry(node_u: torch.Tensor, origin, spacing):
    gnz, gny, gnx, three = node_u.shape
    if three != 3:
        raise ValueError("node_u must be (gnz, gny, gnx, 3)")
    ox, oy, oz = (float(v) for v in origin)
    sx, sy, sz = (float(v) for v in spacing)
    return gnx, gny, gnz, ox, oy, oz, sx, sy, sz


def _launch_forward(f, g, centers, half, node_u, origin, spacing, mode):
    nz, ny, nx = (int(v) for v in f.shape)
    S = int(centers.shape[0])
    C = torch.empty(S, dtype=torch.float64, device=f.device)
    stats = torch.empty((S, NSTAT), dtype=torch.float64, device=f.device)
    check(
        lib().aldvc_warp_znssd(K._ptr(f), K._ptr(g), nz, ny, nx, int(mode), K._ptr(centers), S, *(int(h) for h in half),
                               K._ptr(node_u), *_geometry(node_u, origin, spacing), K._ptr(C), K._ptr(stats),
                               K.stream_handle()),
        "aldvc_warp_znssd",
    )  # fmt: skip
    return C, stats


class WarpResidual(torch.autograd.Function):
    """``C (S,) float64 = WarpResidual.apply(node_u, g, f, centers, half, origin, spacing, mode)``.

    ``node_u`` (gnz, gny, gnx, 3) float32 node displacements [u, v, w] on the grid of nodes
    ``origin + (i, j, k) * spacing`` (x, y, z voxels); ``g``, ``f`` (nz, ny, nx) float32 on the
    GPU; ``centers`` (S, 3) int64 subset centres [x, y, z]; ``half`` (hx, hy, hz). Subsets that
    leave the volume or the sampling domain give NaN and no gradient."""

    @staticmethod
    def forward(ctx, node_u, g, f, centers, half, origin, spacing, mode=1):
        node_u = node_u.detach().to(torch.float32).contiguous()
        gd = g.detach().to(torch.float32).contiguous()
        fd = f.detach().to(torch.float32).contiguous()
        cd = centers.to(device=fd.device, dtype=torch.int64).contiguous()
        C, stats = _launch_forward(fd, gd, cd, half, node_u, origin, spacing, mode)
        ctx.save_for_backward(node_u, gd, fd, cd, stats)
        ctx.geom = (tuple(half), tuple(origin), tuple(spacing), int(mode))
        ctx.need_g = bool(ctx.needs_input_grad[1])
        return C

    @staticmethod
    def backward(ctx, grad_C):
        node_u, g, f, centers, stats = ctx.saved_tensors
        half, origin, spacing, mode = ctx.geom
        nz, ny, nx = (int(v) for v in f.shape)
        gC = torch.nan_to_num(grad_C.to(torch.float64), nan=0.0).contiguous()
        grad_u = torch.zeros(node_u.shape, dtype=torch.float64, device=f.device)
        grad_g = torch.zeros_like(g) if ctx.need_g else None
        S = int(centers.shape[0])
        gnz, gny, gnx = (int(v) for v in node_u.shape[:3])
        nbytes = int(lib().aldvc_warp_backward_workspace_bytes(nz, ny, nx, S, *(int(h) for h in half), gnx, gny, gnz))
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=f.device)
        check(
            lib().aldvc_warp_znssd_backward(
                K._ptr(f), K._ptr(g), nz, ny, nx, mode, K._ptr(centers), S, *half, K._ptr(node_u),
                *_geometry(node_u, origin, spacing), K._ptr(stats), K._ptr(gC), K._ptr(grad_u), K._ptr(grad_g),
                K._ptr(ws), nbytes, K.stream_handle(),
            ),
            "aldvc_warp_znssd_backward",
        )  # fmt: skip
        return grad_u.to(torch.float32), grad_g, None, None, None, None, None, None


def warp_residual(node_u, g, f, centers, half, origin, spacing, mode: int = 1) -> torch.Tensor:
    """Per-subset ZNSSD under the node field (differentiable in ``node_u`` and ``g``)."""
    return WarpResidual.apply(node_u, g, f, centers, half, origin, spacing, mode)


def warp_znssd_stats(node_u, g, f, centers, half, origin, spacing, mode: int = 1):
    """Forward only: ``(C (S,), stats (S, 8))`` with the statistics n, m_f, b_f, m_g, b_g, ZNCC,
    valid flag and C per subset."""
    dev = K.device()
    t = lambda x, dt: (x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))).to(dev, dt).contiguous()  # noqa: E731
    return _launch_forward(t(f, torch.float32), t(g, torch.float32), t(centers, torch.int64), half,
                           t(node_u, torch.float32), origin, spacing, mode)  # fmt: skip
