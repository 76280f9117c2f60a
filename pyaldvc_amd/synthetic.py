"""Synthetic bead volume pairs with a known imposed displacement field.

A bead is ``A * exp(-|x - c|^2 / (2 sigma^2))`` rendered over a cube of half-width
``ceil(3 sigma)`` around its centre. The deformed volume re-renders the same beads at the
mapped centres ``x = X + u(X)`` (Lagrangian convention of the solver), so the error of a
solved node is measured against ``u`` evaluated at the node. Seeds: bead placement ``seed``,
noise on the reference ``seed + 1``, noise on the deformed volume ``seed + 2``.

Volumes are ``(nz, ny, nx)`` float32; bead centres are ``(n, 3)`` float64 in ``[x, y, z]``
order, the node-coordinate order of the solver.

The reference ships its own generators (synthetic.py:20-57 of al_dvc); they normalise to unit
maximum and resample the deformed frame instead of re-rendering it, so they do not produce the
benchmark's bead pairs and are not used here.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable

import numpy as np

DispFunc = Callable[[np.ndarray, np.ndarray, np.ndarray], tuple[np.ndarray, np.ndarray, np.ndarray]]


@dataclass(frozen=True)
class BeadSet:
    centres: np.ndarray  # (n, 3) float64, [x, y, z]
    amps: np.ndarray  # (n,) float64
    sigma: float


def random_beads(shape, n_beads: int, sigma: float, seed: int, amp_range=(0.5, 1.0)) -> BeadSet:
    """``n_beads`` centres uniform over the volume and amplitudes uniform in ``amp_range``."""
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, size=(int(n_beads), 3)) * np.array([nx, ny, nz], dtype=np.float64)
    amps = rng.uniform(amp_range[0], amp_range[1], size=int(n_beads))
    return BeadSet(centres=centres, amps=amps, sigma=float(sigma))


def moved(beads: BeadSet, disp: DispFunc) -> BeadSet:
    """The same beads at ``X + u(X)``."""
    c = beads.centres
    u, v, w = disp(c[:, 0], c[:, 1], c[:, 2])
    return BeadSet(centres=c + np.column_stack([u, v, w]), amps=beads.amps, sigma=beads.sigma)


def render_np(shape, beads: BeadSet, chunk: int = 20000) -> np.ndarray:
    """Render ``beads`` into a float32 volume on the host (float64 accumulation)."""
    nz, ny, nx = shape
    r = int(math.ceil(3.0 * beads.sigma))
    off = np.arange(-r, r + 1)
    oz, oy, ox = np.meshgrid(off, off, off, indexing="ij")
    oz, oy, ox = oz.ravel(), oy.ravel(), ox.ravel()
    vol = np.zeros(nz * ny * nx, dtype=np.float64)
    inv2s2 = 1.0 / (2.0 * beads.sigma * beads.sigma)
    for s in range(0, beads.centres.shape[0], chunk):
        c = beads.centres[s : s + chunk]
        a = beads.amps[s : s + chunk]
        bx = np.rint(c[:, 0]).astype(np.int64)[:, None] + ox[None, :]
        by = np.rint(c[:, 1]).astype(np.int64)[:, None] + oy[None, :]
        bz = np.rint(c[:, 2]).astype(np.int64)[:, None] + oz[None, :]
        d2 = (bx - c[:, 0:1]) ** 2 + (by - c[:, 1:2]) ** 2 + (bz - c[:, 2:3]) ** 2
        val = a[:, None] * np.exp(-d2 * inv2s2)
        ok = (bx >= 0) & (bx < nx) & (by >= 0) & (by < ny) & (bz >= 0) & (bz < nz)
        lin = (bz * ny + by) * nx + bx
        vol += np.bincount(lin[ok], weights=val[ok], minlength=vol.size)
    return vol.reshape(shape).astype(np.float32)


def add_noise(vol: np.ndarray, sigma: float, seed: int) -> np.ndarray:
    if sigma <= 0:
        return vol
    rng = np.random.default_rng(seed)
    return (vol + rng.normal(0.0, sigma, size=vol.shape)).astype(np.float32)


# --------------------------------------------------------------------------- imposed fields


def translation(t) -> DispFunc:
    tx, ty, tz = (float(v) for v in t)

    def fn(x, y, z):
        x = np.asarray(x, dtype=np.float64)
        return np.full_like(x, tx), np.full_like(x, ty), np.full_like(x, tz)

    return fn


def affine(F, centre, t=(0.0, 0.0, 0.0)) -> DispFunc:
    """``u(X) = t + F (X - c)``, ``F[i, j] = du_i / dx_j``."""
    Fm = np.asarray(F, dtype=np.float64)
    cx, cy, cz = centre
    tx, ty, tz = t

    def fn(x, y, z):
        dx, dy, dz = np.asarray(x) - cx, np.asarray(y) - cy, np.asarray(z) - cz
        return (
            tx + Fm[0, 0] * dx + Fm[0, 1] * dy + Fm[0, 2] * dz,
            ty + Fm[1, 0] * dx + Fm[1, 1] * dy + Fm[1, 2] * dz,
            tz + Fm[2, 0] * dx + Fm[2, 1] * dy + Fm[2, 2] * dz,
        )

    return fn


def strain_rotation_F(strain_z: float = 0.02, angle_deg: float = 5.0) -> np.ndarray:
    """``F = R E - I`` for a uniaxial stretch ``E = diag(1, 1, 1 + strain_z)`` then a rotation about z."""
    a = math.radians(angle_deg)
    R = np.array([[math.cos(a), -math.sin(a), 0.0], [math.sin(a), math.cos(a), 0.0], [0.0, 0.0, 1.0]])
    E = np.diag([1.0, 1.0, 1.0 + strain_z])
    return R @ E - np.eye(3)


def sinusoid_z(amplitude: float, wavelength: float) -> DispFunc:
    """``u_z = A sin(2 pi x / L)``, other components zero."""
    k = 2.0 * math.pi / wavelength

    def fn(x, y, z):
        x = np.asarray(x, dtype=np.float64)
        return np.zeros_like(x), np.zeros_like(x), amplitude * np.sin(k * x)

    return fn


def hertz_dimple(depth: float, radius: float, centre_xy, z_top: float) -> DispFunc:
    """``u_z = -depth * max(0, 1 - r^2 / radius^2)^1.5`` under the top surface ``z_top``."""
    cx, cy = centre_xy

    def fn(x, y, z):
        x = np.asarray(x, dtype=np.float64)
        r2 = (x - cx) ** 2 + (np.asarray(y) - cy) ** 2
        prof = np.maximum(0.0, 1.0 - r2 / (radius * radius)) ** 1.5
        w = -depth * prof * np.ones_like(np.asarray(z, dtype=np.float64))
        return np.zeros_like(x), np.zeros_like(x), w

    return fn


def node_lattice(shape, winsize, step) -> np.ndarray:
    """``(N, 3)`` int64 [x, y, z] node centres ``w/2 + step*i``; node order ``iz*ny*nx + iy*nx + ix``.

    There are ``(L - w) // step + 1`` nodes along an axis of length ``L``; the first subset starts
    at voxel 0.
    """
    nz, ny, nx = shape
    wx, wy, wz = (int(w) for w in winsize)
    sx, sy, sz = (int(s) for s in step)
    ax = [wx // 2 + sx * np.arange((nx - wx) // sx + 1), wy // 2 + sy * np.arange((ny - wy) // sy + 1), wz // 2 + sz * np.arange((nz - wz) // sz + 1)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return np.column_stack([X.ravel(), Y.ravel(), Z.ravel()]).astype(np.int64)


def normalize_np(vol: np.ndarray) -> np.ndarray:
    """``(v - mean) / std`` over the whole volume, float64 statistics, float32 result."""
    v = np.asarray(vol)
    mean = float(np.mean(v, dtype=np.float64))
    std = float(np.std(v, dtype=np.float64))
    if std < 1e-12:
        std = 1.0
    return ((v.astype(np.float64) - mean) / std).astype(np.float32)


def bead_pair_np(shape, n_beads, sigma, disp: DispFunc, seed=0, noise=0.01, amp_range=(0.5, 1.0)):
    """Host-rendered ``(f, g)`` raw float32 pair (not normalised) for small volumes."""
    beads = random_beads(shape, n_beads, sigma, seed, amp_range)
    f = add_noise(render_np(shape, beads), noise, seed + 1)
    g = add_noise(render_np(shape, moved(beads, disp)), noise, seed + 2)
    return f, g
