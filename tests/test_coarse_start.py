"""Host logic of the coarse-lattice start (coarse_init.py:33-78): the nearest-finite fill against
scipy's Euclidean distance transform (the reference's own call), ties included, and the separable
lift against scipy's RegularGridInterpolator, on the CPU (torch CPU tensors)."""

from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy.interpolate import RegularGridInterpolator
from scipy.ndimage import distance_transform_edt

from pyaldvc_amd import init_guess as IG
from pyaldvc_amd import kernels as K
from pyaldvc_amd.local import NodeMesh


def test_nearest_fill_equals_the_distance_transform_on_random_grids():
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(300):
        shape = tuple(int(n) for n in rng.integers(1, 9, size=3))
        t = rng.normal(size=shape + (3,))
        holes = rng.random(shape) < rng.uniform(0.05, 0.95)
        if holes.all() or not holes.any():
            continue
        t[holes, rng.integers(0, 3)] = np.nan
        got = IG._fill_holes_from_nearest(torch.from_numpy(t)).numpy()
        _, idx = distance_transform_edt(holes, return_distances=True, return_indices=True)
        assert np.array_equal(got, t[tuple(idx)])
        checked += 1
    assert checked > 200


def test_nearest_fill_edge_cases():
    t = np.full((2, 2, 2, 1), np.nan)
    assert np.array_equal(IG._fill_holes_from_nearest(torch.from_numpy(t)).numpy(), np.zeros_like(t))
    t = np.arange(8.0).reshape(2, 2, 2, 1)
    assert np.array_equal(IG._fill_holes_from_nearest(torch.from_numpy(t)).numpy(), t)


@pytest.fixture
def cpu_device(monkeypatch):
    monkeypatch.setattr(K, "device", lambda: torch.device("cpu"))


def test_separable_lift_equals_trilinear_interpolation(cpu_device):
    """Non-uniform node axes, the last fine node beyond the lattice (extrapolation), holes with ties."""
    x0 = np.array([4.0, 11, 20, 27, 36, 45, 52, 61])
    y0 = np.array([3.0, 12, 19, 28, 35, 44, 51])
    z0 = np.array([7.0, 15, 24, 31])
    Z, Y, X = np.meshgrid(z0, y0, x0, indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
    mesh = NodeMesh(coords, (4, 7, 8), np.ones(coords.shape[0], dtype=bool))
    coarse, rows = IG.coarse_lattice(mesh, 3)
    assert tuple(coarse.grid_shape) == (2, 3, 3) and np.array_equal(coarse.coordinates, coords[rows])
    rng = np.random.default_rng(5)
    for holes in ([], [3], [3, 7], [0], [17], [1, 2, 5]):
        vals = rng.normal(size=(coarse.n_nodes, 12))
        vals[holes, 4] = np.nan
        got = IG.interpolate_to_mesh(coarse, vals, mesh)
        grid = vals.reshape(2, 3, 3, 12)
        bad = ~np.isfinite(grid).all(-1)
        _, idx = distance_transform_edt(bad, return_distances=True, return_indices=True)
        interp = RegularGridInterpolator((z0[::3], y0[::3], x0[::3]), grid[tuple(idx)], method="linear",
                                         bounds_error=False, fill_value=None)  # fmt: skip
        ref = interp(np.column_stack([coords[:, 2], coords[:, 1], coords[:, 0]]))
        assert np.max(np.abs(got - ref)) < 1e-12, holes
