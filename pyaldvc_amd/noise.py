"""Noise pattern of the IC-GN Hessian (solver/uncertainty.py:47-80).

With white noise of unit variance on the reference gradients, the expected Hessian of the 12
warp parameters is ``E[J J^T]_pq = delta(c_p, c_q) * M[b_p, b_q]``: parameter p multiplies
gradient component ``c_p`` (x, y, z) by subset coordinate ``b_p`` (X, Y, Z, or 1 for the
translations), and ``M = sum_voxels q q^T`` with ``q = (X, Y, Z, 1)`` is the moment matrix of the
sampled subset. The kernels scale it by the 7-point stencil's noise gain (the sum of its squared
weights) when the noise-corrected Hessian is on.
"""

from __future__ import annotations

import numpy as np

# 7-point central difference (volume_ops.py:25-26): weights of f[x-3 .. x+3]
STENCIL7 = np.array([-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0, 0.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0])
STENCIL_NOISE_GAIN = float(np.dot(STENCIL7, STENCIL7))

# P = [F row-major (9), u, v, w]: gradient component and subset coordinate of each parameter
_COMPONENT = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2, 0, 1, 2])
_COORD = np.array([0, 1, 2, 0, 1, 2, 0, 1, 2, 3, 3, 3])


def subset_moment_matrix(half, stride: int = 1) -> np.ndarray:
    """``sum q q^T`` over the sampled subset, ``q = (X, Y, Z, 1)``; diagonal by symmetry."""
    axes = [np.arange(-int(h), int(h) + 1, int(stride), dtype=np.float64) for h in half]
    counts = [a.size for a in axes]
    n = float(np.prod(counts))
    diag = [float(np.sum(a * a)) * n / counts[i] for i, a in enumerate(axes)]
    return np.diag(diag + [n])


def noise_hessian_pattern(half, stride: int = 1) -> np.ndarray:
    """``(12, 12)`` expected Hessian for unit-variance white gradient noise on the subset."""
    M = subset_moment_matrix(half, stride)
    same = _COMPONENT[:, None] == _COMPONENT[None, :]
    return np.where(same, M[_COORD[:, None], _COORD[None, :]], 0.0)
