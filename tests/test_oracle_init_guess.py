"""The oracle's coarse-to-fine initial-guess pieces (oracle/ncc_ref.py) pinned by closed forms and by
the reference's own test cases (tests/test_integer_search.py:28-38, :74-90)."""

from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from oracle import ncc_ref
from oracle import cleanup_ref as cleanup
from pyaldvc_amd import synthetic as syn


def test_block_downsample_matches_explicit_block_means():
    v = np.arange(4 * 6 * 8, dtype=np.float32).reshape(4, 6, 8)
    d = ncc_ref.block_downsample(v, 2)
    assert d.shape == (2, 3, 4) and d.dtype == np.float32
    for z in range(2):
        for y in range(3):
            for x in range(4):
                assert d[z, y, x] == np.float32(v[2 * z : 2 * z + 2, 2 * y : 2 * y + 2, 2 * x : 2 * x + 2].astype(np.float64).mean())
    # trailing voxels are dropped; factor 1 returns the volume
    w = np.random.default_rng(0).normal(size=(5, 7, 9)).astype(np.float32)
    assert ncc_ref.block_downsample(w, 2).shape == (2, 3, 4)
    assert ncc_ref.block_downsample(w, 3).shape == (1, 2, 3)
    assert np.array_equal(ncc_ref.block_downsample(w, 1), w)
    # a block mean of a constant is the constant
    assert np.all(ncc_ref.block_downsample(np.full((6, 6, 6), 2.5, np.float32), 3) == 2.5)


@pytest.mark.parametrize("shift", [(3, -2, 5), (-7, 4, 0), (0, 0, 0)])
def test_phase_correlation_recovers_an_exact_cyclic_shift(shift):
    rng = np.random.default_rng(1)
    f = ndimage.gaussian_filter(rng.normal(size=(40, 48, 56)), 1.5).astype(np.float32)
    dx, dy, dz = shift
    g = np.roll(f, (dz, dy, dx), axis=(0, 1, 2))
    assert np.array_equal(ncc_ref.phase_correlation_shift(f, g), np.array(shift, dtype=float))


def test_phase_correlation_on_the_reference_shifted_pair():
    """tests/test_integer_search.py:35-38 with a bead pair: within one voxel of round(shift)."""
    shift = (3.4, -2.6, 1.2)
    f, g = syn.bead_pair_np((72, 64, 64), 4000, 2.0, syn.translation(shift), seed=11)
    s = ncc_ref.phase_correlation_shift(syn.normalize_np(f), syn.normalize_np(g))
    assert np.allclose(s, np.round(shift), atol=1.0)


def test_coarse_radius_and_levels_by_hand():
    assert ncc_ref.coarse_radius((4, 4, 4), 2) == (3, 3, 3)  # max(2, ceil(4/2) + 1)
    assert ncc_ref.coarse_radius((4, 4, 4), 4) == (2, 2, 2)  # max(2, 1 + 1)
    assert ncc_ref.coarse_radius((16, 5, 1), 8) == (3, 2, 2)
    # (nz, ny, nx) = (80, 72, 72), 16^3 templates, radius 4: level 2 needs 8 + 1 + 2*2 + 4 = 17 <= 18
    # coarse voxels, level 3 needs 17 > 9
    assert ncc_ref.auto_pyramid_levels((80, 72, 72), (16, 16, 16), (4, 4, 4)) == 2
    # nx = 40: level 1 needs 8 + 1 + 2*3 + 4 = 19 <= 20, level 2 needs 17 > 10; nx = 36 fits none (19 > 18)
    assert ncc_ref.auto_pyramid_levels((80, 72, 40), (16, 16, 16), (4, 4, 4)) == 1
    assert ncc_ref.auto_pyramid_levels((80, 72, 36), (16, 16, 16), (4, 4, 4)) == 0
    # white noise: a block mean over f^3 voxels keeps 1/sqrt(f^3) of the standard deviation, 0.354 at
    # f = 2 (above the 0.35 bar) and 0.125 at f = 4, so exactly one level survives
    noise = np.random.default_rng(2).normal(size=(80, 72, 72)).astype(np.float32)
    assert ncc_ref.auto_pyramid_levels((80, 72, 72), (16, 16, 16), (4, 4, 4), f=noise) == 1


def _clean_host(d, grid_shape, threshold):
    """integer_search.py:545-556 on the host pieces of the product (pinned by reference vectors)."""
    d = np.array(d, dtype=np.float64).reshape(tuple(grid_shape) + (3,))
    valid = np.all(np.isfinite(d), axis=-1)
    if threshold > 0 and valid.sum() > 27:
        d[cleanup.universal_median_test(d, valid, threshold)] = np.nan
    return np.stack([cleanup.fill_nan_grid(d[..., c]) for c in range(3)], -1).reshape(-1, 3)


def test_pyramid_recovers_the_reference_large_shift():
    """tests/test_integer_search.py:74-90 with a bead pair: every node within 0.3 voxel, with and
    without the global pre-shift."""
    shift = (11.3, -9.6, 6.4)
    f, g = syn.bead_pair_np((80, 72, 72), 5000, 2.0, syn.translation(shift), seed=12, noise=0.0)
    f, g = syn.normalize_np(f), syn.normalize_np(g)
    ax = np.arange(24, 50, 8)
    gz, gy, gx = np.meshgrid(ax, ax, ax, indexing="ij")
    coords = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1).astype(np.int64)
    levels = ncc_ref.auto_pyramid_levels(f.shape, (16, 16, 16), (4, 4, 4))
    assert levels >= 1
    info = ncc_ref.pyramid_search(f, g, coords, (4, 4, 4), (16, 16, 16), (4, 4, 4), levels, _clean_host)
    assert np.all(np.abs(info["disp"] - np.array(shift)) < 0.3)
    info2 = ncc_ref.pyramid_search(f, g, coords, (4, 4, 4), (16, 16, 16), (4, 4, 4), 0, _clean_host,
                                   global_shift=ncc_ref.phase_correlation_shift(f, g))  # fmt: skip
    assert np.all(np.abs(info2["disp"] - np.array(shift)) < 0.3)


@pytest.fixture(scope="module")
def gp():
    import os

    with np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_pyramid.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("key", ["plain", "gshift"])
def test_pyramid_matches_the_reference_vectors(gp, key):
    """The reference's own pyramid_search (make_golden_pyramid.py): same level count, per-level radii
    and solved counts, expansions, ok/clipped flags; displacements within 2e-4 voxel (its maps come
    from float32 FFTs, the oracle's from float64 direct sums)."""
    f, g = gp["f"], gp["g"]
    assert ncc_ref.auto_pyramid_levels(f.shape, (16, 16, 16), (4, 4, 4), f=f) == int(gp["levels_auto"])
    gs = ncc_ref.phase_correlation_shift(f, g)
    assert np.array_equal(gs, gp["global_shift"])
    info = ncc_ref.pyramid_search(f, g, gp["coords"], (4, 4, 4), (16, 16, 16), (4, 4, 4), 0, _clean_host,
                                  global_shift=gs if key == "gshift" else None)  # fmt: skip
    assert len(info["per_level"]) == int(gp[f"{key}_levels"]) + 1
    assert [tuple(p["radius"]) for p in info["per_level"]] == [tuple(r) for r in gp[f"{key}_radii"]]
    assert [p["n_ok"] for p in info["per_level"]] == gp[f"{key}_n_ok"].tolist()
    assert np.array_equal(info["ok"], gp[f"{key}_ok"]) and np.array_equal(info["clipped"], gp[f"{key}_clipped"])
    assert np.max(np.abs(info["disp"] - gp[f"{key}_disp"])) < 2e-4
    assert np.max(np.abs(info["cc"] - gp[f"{key}_cc"])) < 1e-5
