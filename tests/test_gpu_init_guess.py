"""Initial-guess orchestration on the GPU (pyaldvc_amd/init_guess.py) against the oracle
(oracle/ncc_ref.py) and the reference's own test cases."""

from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage
from types import SimpleNamespace

from oracle import ncc_ref
from oracle import cleanup_ref as cleanup
from pyaldvc_amd import init_guess as IG
from pyaldvc_amd import synthetic as syn
from pyaldvc_amd.local import NodeMesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("factor", [2, 3, 4])
def test_block_downsample_matches_oracle(factor):
    v = np.random.default_rng(factor).normal(size=(37, 42, 51)).astype(np.float32)
    got = IG.block_downsample(v, factor).cpu().numpy()
    ref = ncc_ref.block_downsample(v, factor)
    assert got.shape == ref.shape
    # float64 block sums on the device, NumPy's float32 mean in the oracle: a few ulps apart
    assert np.max(np.abs(got - ref)) <= 4 * np.finfo(np.float32).eps * max(1.0, float(np.max(np.abs(ref))))


@pytest.mark.parametrize("shift", [(3, -2, 5), (-7, 4, 0)])
def test_phase_correlation_matches_oracle_on_a_cyclic_shift(shift):
    rng = np.random.default_rng(1)
    f = ndimage.gaussian_filter(rng.normal(size=(40, 48, 56)), 1.5).astype(np.float32)
    g = np.roll(f, (shift[2], shift[1], shift[0]), axis=(0, 1, 2))
    assert np.array_equal(IG.phase_correlation_shift(f, g), np.array(shift, dtype=float))


def _pair(shift, shape=(80, 72, 72), seed=12):
    f, g = syn.bead_pair_np(shape, 5000, 2.0, syn.translation(shift), seed=seed, noise=0.0)
    return syn.normalize_np(f), syn.normalize_np(g)


def _clean_host(d, grid_shape, threshold):
    d = np.array(d, dtype=np.float64).reshape(tuple(grid_shape) + (3,))
    valid = np.all(np.isfinite(d), axis=-1)
    if threshold > 0 and valid.sum() > 27:
        d[cleanup.universal_median_test(d, valid, threshold)] = np.nan
    return np.stack([cleanup.fill_nan_grid(d[..., c]) for c in range(3)], -1).reshape(-1, 3)


def test_phase_correlation_and_levels_match_oracle_on_a_bead_pair():
    f, g = _pair((11.3, -9.6, 6.4))
    assert np.array_equal(IG.phase_correlation_shift(f, g), ncc_ref.phase_correlation_shift(f, g))
    for shape in ((80, 72, 72), (80, 72, 40)):
        assert IG.auto_pyramid_levels(shape, (16, 16, 16), (4, 4, 4)) == ncc_ref.auto_pyramid_levels(shape, (16, 16, 16), (4, 4, 4))
    assert IG.auto_pyramid_levels(f.shape, (16, 16, 16), (4, 4, 4), f=f) == ncc_ref.auto_pyramid_levels(
        f.shape, (16, 16, 16), (4, 4, 4), f=f)  # fmt: skip


@pytest.mark.parametrize("use_shift", [False, True])
def test_pyramid_search_matches_oracle_and_recovers_the_large_shift(use_shift):
    """tests/test_integer_search.py:74-90: every node within 0.3 voxel of the imposed shift; the
    device path (HIP maps, device median test and spring fill) agrees with the oracle."""
    shift = (11.3, -9.6, 6.4)
    f, g = _pair(shift)
    ax = np.arange(24, 50, 8)
    gz, gy, gx = np.meshgrid(ax, ax, ax, indexing="ij")
    coords = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1).astype(np.int64)
    gs = ncc_ref.phase_correlation_shift(f, g) if use_shift else None
    levels = 0 if use_shift else ncc_ref.auto_pyramid_levels(f.shape, (16, 16, 16), (4, 4, 4))
    ref = ncc_ref.pyramid_search(f, g, coords, (4, 4, 4), (16, 16, 16), (4, 4, 4), levels, _clean_host, global_shift=gs)
    got = IG.pyramid_search(f, g, coords, (4, 4, 4), (16, 16, 16), (4, 4, 4), levels=levels, global_shift=gs)
    assert np.all(np.abs(got["disp"] - np.array(shift)) < 0.3)
    assert np.array_equal(got["ok"], ref["ok"])
    assert np.max(np.abs(got["disp"] - ref["disp"])) < 1e-3  # float32 maps on the device, float64 in the oracle
    assert [p["radius"] for p in got["per_level"]] == [p["radius"] for p in ref["per_level"]]


def test_split_nodes_take_their_initial_guess_from_their_own_side():
    """tests/test_subset_split.py:220-248: a node whose subset the cut reduced below 90 % is filled
    from its own side; without the split information its contaminated guess stays."""
    axis = np.arange(5.0)
    gz, gy, gx = np.meshgrid(axis, axis, axis, indexing="ij")
    coordinates = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    grid = (5, 5, 5)
    n = coordinates.shape[0]
    ix = (np.arange(n) % grid[2]).reshape(grid)
    edge_ok = np.ones(grid + (3,), dtype=bool)
    edge_ok[:, :, 1, 0] = False  # the +x edge between ix = 1 and ix = 2 is cut everywhere
    mesh = NodeMesh(coordinates, grid, np.ones(n, dtype=bool), edge_ok.reshape(n, 3))
    disp = np.zeros((n, 3))
    disp[(ix <= 1).ravel(), 0] = 1.0
    disp[(ix >= 2).ravel(), 0] = -1.0
    disp[(ix == 1).ravel(), 0] = 0.0  # the contaminated peak: halfway between the two sides
    split_fraction = np.ones(n, dtype=np.float32)
    split_fraction[(ix == 1).ravel()] = 0.6
    para = SimpleNamespace(init_min_pce=0.0, init_outlier_threshold=0.0)
    ok = np.ones(n, dtype=bool)
    out, bad = IG.clean_initial_guess(disp.copy(), ok, None, mesh, para, split_fraction)
    cut = (ix == 1).ravel()
    assert bad[cut].all() and not bad[~cut].any()
    assert np.allclose(out[cut, 0], 1.0, atol=1e-6)
    assert np.allclose(out[(ix >= 2).ravel(), 0], -1.0)
    out2, bad2 = IG.clean_initial_guess(disp.copy(), ok, None, mesh, para, None)
    assert not bad2.any() and np.allclose(out2[cut, 0], 0.0)


def test_clean_initial_guess_median_test_matches_host():
    """Outliers planted in a smooth field are flagged and refilled as the host restatement does."""
    grid = (6, 7, 8)
    n = int(np.prod(grid))
    gz, gy, gx = np.meshgrid(*(np.arange(s, dtype=float) for s in grid), indexing="ij")
    coordinates = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    disp = np.stack([0.1 * gx.ravel(), -0.05 * gy.ravel(), 0.02 * gz.ravel() + 1.0], 1)
    rng = np.random.default_rng(5)
    out_idx = rng.choice(n, 12, replace=False)
    disp[out_idx] += rng.normal(0, 4.0, (12, 3))
    ok = np.ones(n, dtype=bool)
    ok[rng.choice(n, 5, replace=False)] = False
    mesh = NodeMesh(coordinates, grid, np.ones(n, dtype=bool))
    para = SimpleNamespace(init_min_pce=0.0, init_outlier_threshold=2.0)
    out, bad = IG.clean_initial_guess(disp.copy(), ok, None, mesh, para)
    d = disp.copy()
    hb = ~ok
    d[hb] = np.nan
    flag = cleanup.universal_median_test(d.reshape(grid + (3,)), (~hb).reshape(grid), 2.0).ravel()
    hb |= flag
    d[hb] = np.nan
    ref = np.stack([cleanup.fill_nan_grid(d[:, c].reshape(grid)).ravel() for c in range(3)], 1)
    assert np.array_equal(bad, hb)
    assert np.max(np.abs(out - ref)) < 1e-8


def _fine_mesh(grid=(5, 7, 8), step=(6.0, 5.0, 4.0), origin=(10.0, 12.0, 14.0)):
    nz, ny, nx = grid
    z0 = origin[2] + step[2] * np.arange(nz)
    y0 = origin[1] + step[1] * np.arange(ny)
    x0 = origin[0] + step[0] * np.arange(nx)
    Z, Y, X = np.meshgrid(z0, y0, x0, indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
    return NodeMesh(coords, grid, np.ones(coords.shape[0], dtype=bool)), (z0, y0, x0)


def test_coarse_lattice_and_interpolation_match_scipy():
    """coarse_init.py:33-67: every k-th node per axis; trilinear interpolation with linear
    extrapolation past the lattice (the fine grid's last x node lies beyond it) and nearest-node
    filling of non-finite coarse values, against scipy's RegularGridInterpolator."""
    from scipy.interpolate import RegularGridInterpolator
    from scipy.ndimage import distance_transform_edt

    mesh, (z0, y0, x0) = _fine_mesh()
    coarse, fine_index = IG.coarse_lattice(mesh, 3)
    assert tuple(coarse.grid_shape) == (2, 3, 3)
    assert np.array_equal(coarse.coordinates, mesh.coordinates[fine_index])
    assert IG.coarse_lattice(mesh, 1) is None and IG.coarse_lattice(mesh, 5) is None  # 5 leaves 1 z node
    rng = np.random.default_rng(4)
    vals = rng.normal(size=(coarse.n_nodes, 4))
    vals[3, 1] = np.nan
    got = IG.interpolate_to_mesh(coarse, vals, mesh)
    grid = vals.reshape(2, 3, 3, 4).copy()
    bad = ~np.all(np.isfinite(grid), axis=-1)
    _, idx = distance_transform_edt(bad, return_distances=True, return_indices=True)
    grid = grid[tuple(idx)]
    interp = RegularGridInterpolator((z0[::3], y0[::3], x0[::3]), grid, method="linear", bounds_error=False, fill_value=None)
    ref = interp(np.column_stack([mesh.coordinates[:, 2], mesh.coordinates[:, 1], mesh.coordinates[:, 0]]))
    assert np.max(np.abs(got - ref)) < 1e-12
    # an affine field is reproduced exactly, extrapolated nodes included
    A = np.array([[0.01, -0.02, 0.03], [0.0, 0.02, -0.01], [0.015, 0.0, -0.02]])
    aff = coarse.coordinates @ A.T + np.array([0.3, -0.4, 0.5])
    fine = IG.interpolate_to_mesh(coarse, aff, mesh)
    assert np.max(np.abs(fine - (mesh.coordinates @ A.T + np.array([0.3, -0.4, 0.5])))) < 1e-12


def test_coarse_initial_guess_starts_near_the_affine_field():
    """coarse_init.py:81-128 on a bead pair under an affine field: the guess from every second
    node's solve, interpolated, is within 0.05 voxel of u and 5e-3 of F at every node."""
    from pyaldvc_amd import local as LOC

    shape = (80, 88, 96)
    F = syn.strain_rotation_F(0.01, 2.0)
    centre = tuple((s - 1) / 2 for s in shape[::-1])
    disp = syn.affine(F, centre, t=(1.3, -0.8, 0.6))
    f_raw, g_raw = syn.bead_pair_np(shape, 3000, 1.5, disp, seed=8, noise=0.005)
    f, g = syn.normalize_np(f_raw), syn.normalize_np(g_raw)
    # nodes 24 voxels clear of every face, so that every subset (and every coarse one) is solvable
    axes = [np.arange(24, n - 23, 8) for n in shape[::-1]]  # x, y, z
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()]).astype(np.int64)
    grid = (axes[2].size, axes[1].size, axes[0].size)
    mesh = NodeMesh(coords.astype(np.float64), grid, np.ones(coords.shape[0], dtype=bool))
    para = SimpleNamespace(init_guess_method="pyramid", winsize=(16, 16, 16), init_subset=None, search_radius=(4, 4, 4),
                           global_shift=True, voi=None, pyramid_levels=0, pyramid_fine_radius=2, ncc_auto_expand=True,
                           ncc_max_expand=3, init_outlier_threshold=2.0, init_min_pce=0.0, init_coarse_factor=2,
                           subset_stride=1, min_valid_ratio=0.5, hessian_cond_max=1e12, subset_split=False,
                           interp_method="cubic", icgn_tol=1e-2, icgn_dp_tol=1e-3, icgn_max_iter=100, icgn_patience=5,
                           icgn_predictive_stop=True, icgn_noise_hessian=False, local_outlier_threshold=2.0,
                           tile_local=0)  # fmt: skip
    ref = SimpleNamespace(f=f, mask=None, gradient_mode="stored")
    U0, F0, info = IG.coarse_initial_guess(ref, g, g, mesh, para)
    assert info["coarse_factor"] == 2 and info["coarse_nodes"] == int(np.prod([(n + 1) // 2 for n in grid]))
    u, v, w = disp(*(coords[:, a].astype(float) for a in range(3)))
    inner = np.ones(coords.shape[0], dtype=bool)
    err = np.abs(U0 - np.column_stack([u, v, w]))
    assert np.max(err[inner]) < 0.05, (info, np.median(err[inner], axis=0), coords[inner][np.argmax(err[inner].max(1))])
    assert np.max(np.abs(F0[inner] - np.asarray(F))) < 5e-3  # F is the displacement gradient, as P stores it


# ---------------------------------------------------------------------------- reference vectors
@pytest.fixture(scope="module")
def gp():
    import os

    with np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_pyramid.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("key", ["plain", "gshift"])
def test_pyramid_matches_the_reference_vectors(gp, key):
    """The device pyramid against the reference's own pyramid_search (make_golden_pyramid.py): level
    count, per-level radii and solved counts, expansions and flags identical; displacements within
    1e-3 voxel."""
    f, g = gp["f"], gp["g"]
    assert IG.auto_pyramid_levels(f.shape, (16, 16, 16), (4, 4, 4), f=f) == int(gp["levels_auto"])
    gs = IG.phase_correlation_shift(f, g)
    assert np.array_equal(gs, gp["global_shift"])
    got = IG.pyramid_search(f, g, gp["coords"], (4, 4, 4), (16, 16, 16), (4, 4, 4), levels=0,
                            global_shift=gs if key == "gshift" else None)  # fmt: skip
    assert got["levels"] == int(gp[f"{key}_levels"])
    assert [tuple(p["radius"]) for p in got["per_level"]] == [tuple(r) for r in gp[f"{key}_radii"]]
    assert [p["n_ok"] for p in got["per_level"]] == gp[f"{key}_n_ok"].tolist()
    assert got["expansions"] == int(gp[f"{key}_expansions"])
    assert np.array_equal(got["ok"], gp[f"{key}_ok"]) and np.array_equal(got["clipped"], gp[f"{key}_clipped"])
    assert np.max(np.abs(got["disp"] - gp[f"{key}_disp"])) < 1e-3


@pytest.mark.parametrize("key", ["ncc_r3", "ncc_shift", "ncc_exp"])
def test_ncc_search_matches_the_reference_vectors(golden, key):
    """integer_search.ncc_search / ncc_search_expanding on the device against the reference's outputs
    (golden_local_step.npz): flags, expansions and radius identical, integer peaks identical,
    sub-voxel displacements within 1e-3 voxel."""
    from pyaldvc_amd import integer_search as IS

    f, g, sc = golden["f"], golden["g"], golden["ncc_coords"]
    if key == "ncc_r3":
        res = IS.ncc_search(f, g, sc, (8, 8, 8), (3, 3, 3))
    elif key == "ncc_shift":
        res = IS.ncc_search(f, g, sc, (8, 6, 8), (2, 3, 2), golden["ncc_shift0"])
    else:
        res = IS.ncc_search_expanding(f, g, sc, (8, 8, 8), (1, 1, 1))
        assert res["expansions"] == int(golden["ncc_exp_expansions"])
        assert tuple(res["radius"]) == tuple(golden["ncc_exp_radius"])
    assert np.array_equal(res["ok"], golden[f"{key}_ok"])
    assert np.array_equal(res["clipped"], golden[f"{key}_clipped"])
    ok = res["ok"]
    assert np.array_equal(np.round(res["disp"][ok]), np.round(golden[f"{key}_disp"][ok]))
    assert np.max(np.abs(res["disp"][ok] - golden[f"{key}_disp"][ok])) < 1e-3
    assert np.max(np.abs(res["cc"][ok] - golden[f"{key}_cc"][ok])) < 1e-4
