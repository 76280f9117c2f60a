"""Volume preparation on the GPU (io/volume_ops.py:53-292): VOI statistics and normalisation, the
cubic B-spline prefilter, the deformed frame's NaN mask, and the reference bundle.

Same signatures and results as the reference functions; host arrays in give host arrays out,
torch GPU tensors in give GPU tensors out. Every voxel-level step runs in the HIP library
(``csrc/aldvc_volume.hip``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import kernels as K
from ._lib import check, lib

_DTYPE_CODE = {torch.float32: 0, torch.uint16: 1, torch.uint8: 2, torch.int16: 3, torch.float64: 4}
GRADIENT_AUTO_BYTES = 8 * 1024**3  # "auto" stores gradients below this many bytes (volume_ops.py:247)


def _device_volume(vol) -> tuple[torch.Tensor, bool]:
    """The volume on the GPU in one of the kernels' input types, and whether it came from the host."""
    host = not isinstance(vol, torch.Tensor)
    if host:
        a = np.ascontiguousarray(vol)
        if a.ndim != 3:
            raise ValueError(f"Expected a 3-D volume, got shape {a.shape}.")
        t = torch.from_numpy(a if a.dtype != np.uint16 else a.view(np.int16)).to(K.device())
        if a.dtype == np.uint16:
            t = t.view(torch.uint16)
    else:
        if vol.dim() != 3:
            raise ValueError(f"Expected a 3-D volume, got shape {tuple(vol.shape)}.")
        t = vol.to(K.device()).contiguous()
    if t.dtype not in _DTYPE_CODE:
        t = t.to(torch.float64)
    return t, host


def _voi_bounds(shape, voi):
    """Half-open ``(z0, z1, y0, y1, x0, x1)`` of the VOI clamped to ``shape``; a range ``(lo, hi)``
    is inclusive and ``hi < 0`` means the last voxel (data_structures.py:100-124)."""
    if voi is None:
        return 0, shape[0], 0, shape[1], 0, shape[2]

    def one(rng, n):
        lo, hi = int(rng[0]), int(rng[1])
        if hi < 0:
            hi = n - 1
        lo = max(0, min(lo, n - 1))
        hi = max(lo, min(hi, n - 1))
        return lo, hi + 1

    z0, z1 = one(voi.z, shape[0])
    y0, y1 = one(voi.y, shape[1])
    x0, x1 = one(voi.x, shape[2])
    return z0, z1, y0, y1, x0, x1


def voi_mean_std(vol, voi=None) -> tuple[float, float]:
    """Mean and population standard deviation of ``vol`` inside the clamped VOI, in float64: sums
    of ``v - pilot`` and its square with the VOI's centre voxel as pilot (volume_ops.py:83-102)."""
    t, _ = _device_volume(vol)
    nz, ny, nx = t.shape
    z0, z1, y0, y1, x0, x1 = _voi_bounds((nz, ny, nx), voi)
    n = (z1 - z0) * (y1 - y0) * (x1 - x0)
    if n <= 0:
        raise ValueError("VOI is empty after clamping to the volume.")
    pilot = float(t[(z0 + z1) // 2, (y0 + y1) // 2, (x0 + x1) // 2].cpu().numpy())
    sums = torch.zeros(2, dtype=torch.float64, device=t.device)
    ws = torch.empty(int(lib().aldvc_voi_moments_workspace_bytes(nz, ny)), dtype=torch.uint8, device=t.device)
    check(
        lib().aldvc_voi_moments(K._ptr(t), _DTYPE_CODE[t.dtype], nz, ny, nx, z0, z1, y0, y1, x0, x1, pilot,
                                K._ptr(sums), K._ptr(ws), int(ws.numel()), K.stream_handle()),
        "aldvc_voi_moments",
    )  # fmt: skip
    s, ss = (float(v) for v in sums.cpu())
    shift = s / n
    mean, std = pilot + shift, float(np.sqrt(max(ss / n - shift * shift, 0.0)))
    if not np.isfinite(mean) or not np.isfinite(std):
        raise ValueError("Volume contains NaN/Inf inside the VOI; cannot normalise.")
    return mean, std


def normalize_volume(vol, voi=None):
    """``(vol - mean_voi) / std_voi`` as float32, statistics in float64 (volume_ops.py:105-121)."""
    t, host = _device_volume(vol)
    mean, std = voi_mean_std(t, voi)
    if std < 1e-12:
        std = 1.0
    out = torch.empty(t.shape, dtype=torch.float32, device=t.device)
    check(
        lib().aldvc_normalize(K._ptr(t), _DTYPE_CODE[t.dtype], int(t.numel()), float(mean), 1.0 / float(std),
                              K._ptr(out), K.stream_handle()),
        "aldvc_normalize",
    )  # fmt: skip
    return out.cpu().numpy() if host else out


def prefilter_bspline(vol):
    """Cubic B-spline coefficients of ``vol`` (whole-sample mirror boundaries): sampling them with
    the cubic B-spline basis reproduces the samples (volume_ops.py:215-222)."""
    t, host = _device_volume(vol)
    t = t.to(torch.float32).contiguous()
    nz, ny, nx = t.shape
    out = torch.empty_like(t)
    scratch = torch.empty_like(t)
    check(lib().aldvc_bspline_prefilter(K._ptr(t), nz, ny, nx, K._ptr(out), K._ptr(scratch), K.stream_handle()),
          "aldvc_bspline_prefilter")  # fmt: skip
    return out.cpu().numpy() if host else out


def prepare_deformed(vol, interp_method: str, mask=None):
    """The array the interpolation samples for ``interp_method``: B-spline coefficients for
    ``"bspline"``, the volume otherwise; NaN where the deformed-frame ``mask`` is False
    (volume_ops.py:225-244)."""
    t, host = _device_volume(vol)
    out = prefilter_bspline(t) if interp_method == "bspline" else t.to(torch.float32).clone()
    if mask is not None:
        m = mask if isinstance(mask, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(mask, dtype=bool))
        m = m.to(device=out.device).to(torch.uint8).contiguous()
        if tuple(m.shape) != tuple(out.shape):
            raise ValueError(f"deformed-frame mask shape {tuple(m.shape)} != volume shape {tuple(out.shape)}")
        check(lib().aldvc_mask_nan(K._ptr(out), K._ptr(m), int(out.numel()), K.stream_handle()), "aldvc_mask_nan")
    return out.cpu().numpy() if host else out


@dataclass(frozen=True)
class ReferenceBundle:
    """Everything the local solver needs about a reference frame (data_structures.py:291-335):
    ``f``, gradients ``gx, gy, gz`` (1x1x1 placeholders when evaluated on the fly) and ``mask``
    (1x1x1 placeholder: no mask), here as GPU tensors."""

    f: torch.Tensor
    gx: torch.Tensor
    gy: torch.Tensor
    gz: torch.Tensor
    mask: torch.Tensor

    @property
    def shape(self):
        return tuple(int(s) for s in self.f.shape)

    @property
    def has_mask(self) -> bool:
        return tuple(self.mask.shape) == tuple(self.f.shape)

    @property
    def stored_gradients(self) -> bool:
        return tuple(self.gx.shape) == tuple(self.f.shape)


def resolve_gradient_mode(mode: str, shape) -> str:
    """``"auto"`` becomes ``"on_the_fly"`` when the three gradient volumes would pass 8 GB."""
    if mode != "auto":
        return mode
    return "on_the_fly" if 12 * int(np.prod(shape)) > GRADIENT_AUTO_BYTES else "stored"


def build_reference_bundle(f, mask, gradient_mode: str = "stored") -> ReferenceBundle:
    """Gradients and mask of a normalised reference volume, on the GPU (volume_ops.py:265-292)."""
    fd = (f if isinstance(f, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(f, dtype=np.float32)))
    fd = fd.to(device=K.device(), dtype=torch.float32).contiguous()
    mode = resolve_gradient_mode(gradient_mode, tuple(fd.shape))
    if mode == "on_the_fly":
        gx = gy = gz = torch.zeros((1, 1, 1), dtype=torch.float32, device=fd.device)
    elif mode == "stored":
        gx, gy, gz = K.compute_gradients(fd)
    else:
        raise ValueError(f"gradient_mode must be 'stored' or 'on_the_fly' (got {gradient_mode!r})")
    if mask is None:
        m = torch.ones((1, 1, 1), dtype=torch.uint8, device=fd.device)
    else:
        m = mask if isinstance(mask, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(mask))
        if tuple(m.shape) != tuple(fd.shape):
            raise ValueError(f"mask shape {tuple(m.shape)} != volume shape {tuple(fd.shape)}")
        m = (m.to(fd.device) != 0).to(torch.uint8).contiguous()
    return ReferenceBundle(f=fd, gx=gx, gy=gy, gz=gz, mask=m)
