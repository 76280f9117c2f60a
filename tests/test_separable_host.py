"""Host logic of the separable passes (no GPU): which calls take them, and the workspace they get.

The distinct node coordinates and the z range are counted with torch on whatever device holds the
coordinates (the CPU here), and the library's sizing entries are plain host arithmetic."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pyaldvc_amd import _lib
from pyaldvc_amd import kernels as K


def test_separable_wanted_follows_the_read_count():
    dims = (512, 512, 512)
    assert K.separable_wanted(226_981, (16, 16, 16), 1, dims, True)  # the 512^3 headline lattice
    assert not K.separable_wanted(343, (16, 16, 16), 1, (128, 128, 128), True)  # 343 nodes at 128^3
    assert not K.separable_wanted(226_981, (16, 16, 16), 2, dims, True)  # stride 2: per node
    assert not K.separable_wanted(226_981, (16, 16, 16), 1, dims, False)  # gradients on the fly
    # enough reads in absolute terms but fewer than four times the voxels: per node
    assert not K.separable_wanted(20_000, (16, 16, 16), 1, (2048, 1024, 1024), True)


def test_extent_counts_distinct_centres_and_the_z_range():
    rng = np.random.default_rng(0)
    dims = (60, 70, 80)
    hx, hy, hz = 5, 6, 4
    lat = np.stack(np.meshgrid(np.arange(hx, 80 - hx, 7), np.arange(hy, 70 - hy, 5), np.arange(hz, 60 - hz, 6),
                               indexing="ij"), -1).reshape(-1, 3)  # fmt: skip
    out = np.array([[2, 30, 30], [30, 69, 30], [30, 30, 59]])  # subsets leaving the volume: not counted
    c = np.concatenate([lat, out]).astype(np.int64)
    vd = np.ones(len(c), np.uint8)
    vd[rng.choice(len(lat), 5, replace=False)] = 0  # invalid nodes are not counted either
    ext = K._separable_extent(torch.from_numpy(c), torch.from_numpy(vd), hx, hy, hz, 1, dims, True, True)
    keep = (vd != 0)[: len(lat)]
    ins = lat[keep]
    assert ext == (len(np.unique(ins[:, 0])), len(np.unique(ins[:, 1])), int(np.ptp(ins[:, 2])) + 2 * hz + 1)
    # forced off, on-the-fly gradients or stride 2: the per-node pass
    assert K._separable_extent(torch.from_numpy(c), None, hx, hy, hz, 1, dims, True, False) is None
    assert K._separable_extent(torch.from_numpy(c), None, hx, hy, hz, 1, dims, False, True) is None
    assert K._separable_extent(torch.from_numpy(c), None, hx, hy, hz, 2, dims, True, True) is None
    # left to the size rule, this small set stays per node
    assert K._separable_extent(torch.from_numpy(c), None, hx, hy, hz, 1, dims, True) is None


@pytest.mark.parametrize("entry,nq", [("aldvc_icgn12_row_moments_bytes", 21), ("aldvc_precompute_row_bytes", 39)])
def test_workspace_sizes(entry, nq):
    lib = _lib.lib()
    f = getattr(lib, entry)
    ny, nx = 512, 512
    fixed = 4 * 64 + (4 * 2 * (nx + ny) + 255) // 256 * 256  # header, marks and centre lists
    assert f(61, 61, 512, ny, nx) == fixed + 8 * nq * 512 * 61 * 61
    assert f(0, 61, 512, ny, nx) == 0 and f(61, 61, 0, ny, nx) == 0
