"""Per-node displacement uncertainty on the GPU (solver/uncertainty.py:83-133).

``displacement_uncertainty(ctx, zncc, status, noise_correction)`` keeps the reference's signature
and result: the standard deviation ``(N, 3)`` of ``u, v, w`` in voxels from each node's IC-GN
Hessian, final ZNCC, ``n_valid`` and ``bottomf``, NaN where no estimate exists. The Hessians stay
in HBM (``LocalContext`` keeps them there); ``aldvc_displacement_uncertainty`` solves the three
translation columns of the corrected Hessian per node. When a node's corrected Hessian is exactly
singular the batch is recomputed with the pseudo-inverse, as the reference does when its batched
inverse fails (uncertainty.py:122-125), with torch on the device.
"""

from __future__ import annotations

import numpy as np
import torch

from . import kernels as K
from ._lib import check, lib
from .noise import STENCIL_NOISE_GAIN, noise_hessian_pattern, subset_moment_matrix

N_DOF = 12  # uncertainty.py:41
MIN_EXTRA_DOF = 2  # uncertainty.py:42
MIN_CORRECTED_FRACTION = 0.1  # uncertainty.py:44
STATUS_CONVERGED = 0


def _dev(x, dtype, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x)).to(device=dev, dtype=dtype)


def _context_arrays(ctx, dev):
    d = getattr(ctx, "device", None)
    if isinstance(d, dict) and "H" in d:
        return d["H"], d["valid"], d["n_valid"], d["bottomf"]
    return (_dev(ctx.H_all, torch.float64, dev), _dev(np.asarray(ctx.valid, dtype=np.uint8), torch.uint8, dev),
            _dev(ctx.n_valid, torch.int64, dev), _dev(ctx.bottomf, torch.float64, dev))  # fmt: skip


def _pinv_batch(H, z, st, valid, nv, bf, pattern, noise_correction):
    """The reference's pseudo-inverse branch, on the device in float64."""
    ok = valid.bool() & torch.isfinite(z) & (nv > N_DOF + MIN_EXTRA_DOF)
    if st is not None:
        ok &= st == STATUS_CONVERGED
    out = torch.full((H.shape[0], 3), float("nan"), dtype=torch.float64, device=H.device)
    idx = torch.nonzero(ok).flatten()
    if idx.numel() == 0:
        return out
    om = torch.clamp(1.0 - z[idx], min=0.0)
    nvi = nv[idx].double()
    bf2 = bf[idx] ** 2
    factor = 2.0 * om / (nvi - N_DOF) * bf2
    s2 = om * bf2 / nvi
    pat = torch.from_numpy(pattern).to(H)
    scale = nvi / pattern[9, 9]
    Hs = H[idx].clone()
    t = torch.arange(9, 12, device=H.device)
    if noise_correction:
        corr = (STENCIL_NOISE_GAIN * s2 * scale)[:, None, None] * pat[None]
        td, cd = Hs[:, t, t], corr[:, t, t]
        lim = torch.where(cd > 0, (1.0 - MIN_CORRECTED_FRACTION) * td / torch.clamp(cd, min=1e-300), torch.ones_like(cd)).amin(1)
        Hs = Hs - corr * torch.clamp(lim, 0.0, 1.0)[:, None, None]
    inv = torch.linalg.pinv(Hs)
    var = factor[:, None] * inv[:, t, t]
    if noise_correction:
        cross = inv @ (pat[None] * scale[:, None, None]) @ inv
        var = var + (STENCIL_NOISE_GAIN * s2**2)[:, None] * cross[:, t, t]
    out[idx] = torch.sqrt(torch.clamp(var, min=0.0))
    return out


def displacement_uncertainty(ctx, zncc, status=None, noise_correction: bool = True):
    """Standard deviation ``(N, 3)`` of ``u, v, w`` in voxels; NaN where undefined
    (uncertainty.py:94-133). Host ``zncc`` in -> host array out; a device tensor -> device tensor."""
    dev = K.device()
    H, valid, nv, bf = _context_arrays(ctx, dev)
    N = int(H.shape[0])
    z = _dev(zncc, torch.float64, dev)
    st = None if status is None else _dev(status, torch.int8, dev)
    stride = int(getattr(ctx, "stride", 1))
    moments = np.ascontiguousarray(np.diagonal(subset_moment_matrix(ctx.half, stride)), dtype=np.float64)
    out = torch.empty((N, 3), dtype=torch.float64, device=dev)
    singular = torch.zeros(1, dtype=torch.int32, device=dev)
    check(lib().aldvc_displacement_uncertainty(K._ptr(H), K._ptr(z), K._ptr(st), K._ptr(valid), K._ptr(nv), K._ptr(bf), N,
                                               moments.ctypes.data, float(STENCIL_NOISE_GAIN), int(bool(noise_correction)),
                                               K._ptr(out), K._ptr(singular), K.stream_handle()),
          "aldvc_displacement_uncertainty")  # fmt: skip
    if int(singular.item()):
        out = _pinv_batch(H, z, st, valid, nv, bf, noise_hessian_pattern(ctx.half, stride), noise_correction)
    return out if isinstance(zncc, torch.Tensor) else out.cpu().numpy()
