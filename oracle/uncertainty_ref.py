"""Test infrastructure (the oracle): per-node displacement uncertainty in float64 NumPy.

Restates solver/uncertainty.py:47-133 (``subset_moment_matrix``, ``noise_hessian_pattern``,
``parameter_covariance_factor``, ``displacement_uncertainty``) operation for operation; only the
tests, ``__graft_entry__.smoke()`` and the bench's CPU leg use it. Pinned in
tests/test_oracle_uncertainty.py by hand-derived closed forms (diagonal Hessians with and without
the noise correction, the 10 % floor of the correction active and inactive) and by a Monte-Carlo
least-squares experiment whose empirical parameter covariance the uncorrected model must match.
"""

from __future__ import annotations

import numpy as np

N_DOF = 12  # uncertainty.py:41
MIN_EXTRA_DOF = 2  # uncertainty.py:42
MIN_CORRECTED_FRACTION = 0.1  # uncertainty.py:44
STATUS_CONVERGED = 0
STENCIL7 = np.array([-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0, 0.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0])
STENCIL_NOISE_GAIN = float(np.sum(STENCIL7**2))  # uncertainty.py:43


def subset_moment_matrix(half, stride: int = 1) -> np.ndarray:
    """uncertainty.py:47-58: diag(sum X^2, sum Y^2, sum Z^2, n) over the sampled subset."""
    ax = [np.arange(-int(h), int(h) + 1, int(stride)) for h in half]
    counts = [a.size for a in ax]
    n = counts[0] * counts[1] * counts[2]
    sq = [float(np.sum(a.astype(np.float64) ** 2)) * (n // counts[i]) for i, a in enumerate(ax)]
    return np.diag(sq + [float(n)])


def noise_hessian_pattern(half, stride: int = 1) -> np.ndarray:
    """uncertainty.py:61-80: E[J J^T] for unit white gradient noise, P = [F row-major, u, v, w]."""
    M = subset_moment_matrix(half, stride)
    comp = [p // 3 for p in range(9)] + [0, 1, 2]
    coord = [p % 3 for p in range(9)] + [3, 3, 3]
    pat = np.zeros((N_DOF, N_DOF))
    for p in range(N_DOF):
        for q in range(N_DOF):
            if comp[p] == comp[q]:
                pat[p, q] = M[coord[p], coord[q]]
    return pat


def displacement_uncertainty(H, zncc, valid, n_valid, bottomf, half, stride=1, status=None, noise_correction=True,
                             gain=STENCIL_NOISE_GAIN):  # fmt: skip
    """uncertainty.py:83-133 on explicit arrays: (N, 3) standard deviations, NaN where undefined."""
    H = np.asarray(H, dtype=np.float64)
    z = np.asarray(zncc, dtype=np.float64)
    nvalid = np.asarray(n_valid)
    n = H.shape[0]
    out = np.full((n, 3), np.nan)
    ok = np.asarray(valid, dtype=bool) & np.isfinite(z) & (nvalid > N_DOF + MIN_EXTRA_DOF)
    if status is not None:
        ok &= np.asarray(status) == STATUS_CONVERGED
    idx = np.flatnonzero(ok)
    if idx.size == 0:
        return out
    om = np.clip(1.0 - z[idx], 0.0, None)
    nv = nvalid[idx].astype(np.float64)
    bf2 = np.asarray(bottomf, dtype=np.float64)[idx] ** 2
    factor = 2.0 * om / (nv - N_DOF) * bf2
    s2 = om * bf2 / nv
    pat = noise_hessian_pattern(half, stride)
    scale = nv / pat[9, 9]
    Hs = H[idx].copy()
    t = np.arange(9, 12)
    if noise_correction:
        corr = (gain * s2 * scale)[:, None, None] * pat[None]
        td = Hs[:, t, t]
        cd = corr[:, t, t]
        lim = np.min(np.where(cd > 0, (1.0 - MIN_CORRECTED_FRACTION) * td / np.maximum(cd, 1e-300), 1.0), axis=1)
        Hs = Hs - corr * np.clip(lim, 0.0, 1.0)[:, None, None]
    try:
        inv = np.linalg.inv(Hs)
    except np.linalg.LinAlgError:
        inv = np.linalg.pinv(Hs)
    var = factor[:, None] * inv[:, t, t]
    if noise_correction:
        cross = inv @ (pat[None] * scale[:, None, None]) @ inv
        var = var + (gain * s2**2)[:, None] * cross[:, t, t]
    out[idx] = np.sqrt(np.clip(var, 0.0, None))
    return out
