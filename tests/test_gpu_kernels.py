"""HIP kernels against the oracle, through the C ABI (needs an MI355X).

Criteria: the reference's own GPU-vs-CPU parity bar (tests/test_cuda_backend.py:149-161) with the
sub-voxel tolerance tightened to 1e-3 voxel; interpolation and gradients within float32 round-off;
integer NCC peaks identical.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from oracle import cpu_ref, ncc_ref
from tests.parity import MASK_NONE, OTF, assert_icgn_parity, bead_case, lattice, truth_at

pytestmark = pytest.mark.gpu

K = pytest.importorskip("pyaldvc_amd.kernels")


@pytest.fixture(scope="module")
def case():
    return bead_case()


def test_gradients_match_oracle(golden):
    gx, gy, gz = K.compute_gradients(golden["f"])
    for a, b in ((gx, golden["gx"]), (gy, golden["gy"]), (gz, golden["gz"])):
        assert np.max(np.abs(a - b)) < 2e-6 * max(1.0, float(np.max(np.abs(b))))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_sample_points_match_oracle(golden, mode):
    src = golden["g_bspline"] if mode == 2 else golden["g"]
    args = (golden["pts_z"], golden["pts_y"], golden["pts_x"], mode)
    got = K.sample_points(src, *args)
    ref = cpu_ref.sample_points(src, *args)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.max(np.abs(got[ok] - ref[ok])) < 2e-5 * max(1.0, float(np.max(np.abs(ref[ok]))))


PRE = {
    "plain": dict(half=(6, 6, 6), stride=1, grads="stored", mask=None),
    "mask": dict(half=(6, 6, 6), stride=1, grads="stored", mask="mask"),
    "stride2": dict(half=(6, 5, 4), stride=2, grads="stored", mask=None),
    "otf": dict(half=(6, 6, 6), stride=1, grads="otf", mask=None),
}


def _pre_args(golden, key):
    c = PRE[key]
    grads = (golden["gx"], golden["gy"], golden["gz"]) if c["grads"] == "stored" else (OTF, OTF, OTF)
    mask = golden[c["mask"]] if c["mask"] else MASK_NONE
    return c["half"], c["stride"], grads, mask


@pytest.mark.parametrize("key", list(PRE))
def test_precompute_matches_oracle(golden, key):
    half, stride, grads, mask = _pre_args(golden, key)
    args = (golden["coords"], *half, golden["f"], *grads, mask, 0.5, 1e12, stride)
    H, L, mf, bf, nv, valid = K.precompute_nodes(*args)
    Hr, Lr, mfr, bfr, nvr, validr = cpu_ref.precompute_nodes(*args)
    assert np.array_equal(valid, validr)
    assert np.array_equal(nv, nvr)
    assert np.max(np.abs(H - Hr)) <= 2e-6 * np.max(np.abs(Hr))
    assert np.max(np.abs(mf - mfr)) < 1e-5
    assert np.max(np.abs(bf - bfr) / bfr) < 1e-5
    v = validr
    assert np.max(np.abs(L[v] - Lr[v])) <= 1e-5 * np.sqrt(np.max(np.abs(Hr)))


CASES12 = {
    "cubic": dict(pre="plain", g="g", mode=1),
    "linear": dict(pre="plain", g="g", mode=0),
    "bspline": dict(pre="plain", g="g_bspline", mode=2),
    "mask": dict(pre="mask", g="g", mode=1),
    "defmask": dict(pre="plain", g="g_nan", mode=1),
    "stride2": dict(pre="stride2", g="g", mode=1),
    "otf": dict(pre="otf", g="g", mode=1),
    "status": dict(pre="plain", g="g", mode=1, far=True),
    "noise": dict(pre="plain", g="g", mode=1, noise=True),
    "nopred": dict(pre="plain", g="g", mode=1, predictive=False),
    "stall": dict(pre="plain", g="g_noise", mode=1),
}


def _start(golden, far=False):
    P0 = np.zeros((golden["coords"].shape[0], 12))
    P0[:, 9:] = np.round(golden["U_true"])
    if far:
        P0[0, 9] = 60.0
        P0[1, 10] = np.nan
    return P0


@pytest.mark.parametrize("variant", ["plain", "fused", "fused_gather"])
@pytest.mark.parametrize("key", list(CASES12))
def test_icgn12_matches_oracle(golden, key, variant):
    c = CASES12[key]
    half, stride, grads, mask = _pre_args(golden, c["pre"])
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(golden["coords"], *half, golden["f"], *grads, mask, 0.5, 1e12, stride)
    noise = c.get("noise", False)
    args = (
        golden["coords"], _start(golden, c.get("far", False)), *half, golden["f"], *grads, mask, golden[c["g"]],
        c["mode"], L, mf, bf, valid, 1e-2, 1e-3, 100, 5, stride, H, golden["pattern_666_s1"] if noise else None,
        float(golden["stencil_noise_gain"]) if noise else 0.0, c.get("predictive", True),
    )  # fmt: skip
    ref = cpu_ref.icgn_12dof_parallel(*args)
    got = K.icgn_12dof_parallel(*args, variant=variant)
    if key == "stall":  # a pure-noise deformed volume: the trajectories are chaotic, compare outcomes
        assert np.array_equal(ref[2] == 0, got[2] == 0)
        return
    assert_icgn_parity(ref, got)


@pytest.mark.parametrize("variant", ["plain", "fused", "fused_gather"])
def test_icgn12_split_rows_match_oracle(golden, variant):
    coords = golden["coords"]
    si = np.arange(coords.shape[0], dtype=np.int64)
    rows = golden["split_m0_rows"]
    f, gx, gy, gz, m = golden["f"], golden["gx"], golden["gy"], golden["gz"], golden["split_mask"]
    H, L, mf, bf, nv, valid = K.precompute_nodes(coords, 6, 6, 6, f, gx, gy, gz, m, 0.5, 1e12, 1, si, rows)
    assert np.array_equal(valid, golden["pre_split_valid"])
    assert np.array_equal(nv, golden["pre_split_nvalid"])
    args = (coords, _start(golden), 6, 6, 6, f, gx, gy, gz, m, golden["g"], 1, golden["pre_split_L"],
            golden["pre_split_meanf"], golden["pre_split_bottomf"], golden["pre_split_valid"],
            1e-2, 1e-3, 100, 5, 1, None, None, 0.0, True, si, rows)  # fmt: skip
    ref = (golden["i12_split_P"], golden["i12_split_n_iter"], golden["i12_split_status"], golden["i12_split_zncc"])
    assert_icgn_parity(ref, K.icgn_12dof_parallel(*args, variant=variant))


@pytest.mark.parametrize("variant", ["plain", "fused", "fused_gather"])
@pytest.mark.parametrize("mode", [1, 0])
def test_icgn12_lattice_matches_oracle_and_truth(case, variant, mode):
    """A full node lattice (subset 16, step 8) on a 64x72x80 bead pair under an affine field."""
    coords = lattice(case["shape"], 16, 8)
    half = (8, 8, 8)
    f, g, gx, gy, gz = case["f"], case["g"], case["gx"], case["gy"], case["gz"]
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, *half, f, gx, gy, gz, MASK_NONE, 0.5, 1e12)
    U = truth_at(case, coords)
    P0 = np.zeros((coords.shape[0], 12))
    P0[:, 9:] = np.round(U)
    args = (coords, P0, *half, f, gx, gy, gz, MASK_NONE, g, mode, L, mf, bf, valid, 1e-2, 1e-3, 100, 5)
    ref = cpu_ref.icgn_12dof_parallel(*args)
    got = K.icgn_12dof_parallel(*args, variant=variant)
    assert_icgn_parity(ref, got)
    conv = got[2] == 0
    assert conv.sum() > 0.5 * valid.sum()
    err = np.abs(got[0][conv, 9:] - U[conv])
    assert np.median(err) < (0.05 if mode == 1 else 0.1)


def test_workgroup_clock_records_every_node_and_changes_nothing(case):
    """aldvc_workgroup_clock: one start / end pair per entry of the valid-node list, ends after
    starts, and the solve's outputs bit-identical to a run without the clock."""
    import ctypes

    import torch

    coords = lattice(case["shape"], 16, 8)
    half = (8, 8, 8)
    f, g, gx, gy, gz = case["f"], case["g"], case["gx"], case["gy"], case["gz"]
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, *half, f, gx, gy, gz, MASK_NONE, 0.5, 1e12)
    P0 = np.zeros((coords.shape[0], 12))
    P0[:, 9:] = np.round(truth_at(case, coords))
    args = (coords, P0, *half, f, gx, gy, gz, MASK_NONE, g, 1, L, mf, bf, valid, 1e-2, 1e-3, 100, 5)
    plain = K.icgn_12dof_parallel(*args, variant="fused")
    n = (coords.shape[0] + 1023) // 1024 * 1024  # the launch's grid: the node count rounded up to 1024
    buf = torch.zeros(2 * n, dtype=torch.int64, device=K.device())
    khz = ctypes.c_int(0)
    K.check(K.lib().aldvc_workgroup_clock(ctypes.c_void_p(buf.data_ptr()), n, ctypes.byref(khz)), "clock")
    try:
        clocked = K.icgn_12dof_parallel(*args, variant="fused")
    finally:
        K.lib().aldvc_workgroup_clock(None, 0, None)
    for a, b in zip(plain, clocked):
        assert np.array_equal(a, b, equal_nan=True)
    c = buf.cpu().numpy().reshape(n, 2)
    rec = c[:, 0] > 0
    assert rec.sum() == int(np.sum(valid)) and np.all(c[rec, 1] >= c[rec, 0]) and khz.value > 0
    assert np.all(c[~rec] == 0)


def _rot_z(deg, stretch_z=0.0):
    c, s_ = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    return np.array([[c - 1.0, -s_, 0.0], [s_, c - 1.0, 0.0], [0.0, 0.0, stretch_z]])


# Warps that exercise the fused kernel's LDS bricks: 32^3 subsets that fit a brick, a rotation too
# large for the brick's rows (every chunk gathers from memory), a z stretch that shrinks the chunks,
# nodes whose warped subsets leave the sampling domain (status OUT_OF_BOUNDS), 35 x 33 x 27 subsets
# (rows of 35 voxels in plane pairs, an odd plane count per chunk) in the three interpolation modes,
# a 15^3 subset (rows too short for plane pairs), and 49-voxel rows cut into x/y boxes.
BRICK_CASES = {
    "fits": dict(F=_rot_z(3.0), t=(0.6, -0.4, 0.7)),
    "rotation": dict(F=_rot_z(12.0), t=(0.3, 0.5, -0.6)),
    "stretch": dict(F=_rot_z(2.0, 0.3), t=(-0.4, 0.2, 0.5)),
    "escape": dict(F=_rot_z(1.0), t=(7.4, -1.6, 6.2)),
    "rows35_linear": dict(F=_rot_z(2.0), t=(0.4, 0.3, -0.5), half=(17, 16, 13), mode=0),
    "rows35_cubic": dict(F=_rot_z(2.0), t=(0.4, 0.3, -0.5), half=(17, 16, 13), mode=1),
    "rows35_bspline": dict(F=_rot_z(2.0), t=(0.4, 0.3, -0.5), half=(17, 16, 13), mode=2),
    "rows15": dict(F=_rot_z(3.0), t=(0.6, -0.4, 0.7), half=(7, 7, 7)),
    # subsets wider than a brick row: x/y boxes of 32 x 24 and 17 x 17 voxel columns
    "tiled_cubic": dict(F=_rot_z(3.0), t=(0.6, -0.4, 0.7), half=(24, 20, 10), ax=28 + 14 * np.arange(5)),
    "tiled_linear": dict(F=_rot_z(2.0), t=(0.4, 0.3, -0.5), half=(24, 20, 10), mode=0, ax=28 + 14 * np.arange(5)),
    # corner boxes rotate out of the brick's rows and gather, the others stay in bricks
    "tiled_rotation": dict(F=_rot_z(9.0), t=(0.3, 0.5, -0.6), half=(24, 24, 8), ax=28 + 14 * np.arange(5)),
    # a cylinder mask on the reference and NaN outside it in the deformed volume: selection and the
    # NaN correction on x/y boxes
    "tiled_masked": dict(F=_rot_z(3.0), t=(0.6, -0.4, 0.7), half=(24, 20, 10), ax=28 + 14 * np.arange(5), mask=0.42),
}


def _cylinder(shape, radius_frac):
    nz, ny, nx = shape
    y, x = np.mgrid[0:ny, 0:nx]
    disk = (x - (nx - 1) / 2) ** 2 + (y - (ny - 1) / 2) ** 2 < (radius_frac * min(ny, nx)) ** 2
    return np.broadcast_to(disk, shape).astype(np.uint8).copy()


@pytest.mark.parametrize("variant", ["fused", "fused_gather"])
@pytest.mark.parametrize("key", list(BRICK_CASES))
def test_icgn12_brick_paths_match_oracle(key, variant):
    c = BRICK_CASES[key]
    cs = bead_case(shape=(112, 112, 112), F=c["F"], t=c["t"], seed=9)
    ax = c.get("ax", 24 + 16 * np.arange(5))
    coords = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3).astype(np.int64)
    half, mode = c.get("half", (16, 16, 16)), c.get("mode", 1)
    f, g, gx, gy, gz = cs["f"], cs["g"], cs["gx"], cs["gy"], cs["gz"]
    if mode == 2:  # B-spline coefficients of g (both sides sample the same array)
        g = ndimage.spline_filter(g, order=3, output=np.float32, mode="mirror")
    mask = MASK_NONE
    if "mask" in c:
        mask = _cylinder(f.shape, c["mask"])
        g = np.where(mask > 0, g, np.float32(np.nan)).astype(np.float32)
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, *half, f, gx, gy, gz, mask, 0.5, 1e12)
    U = truth_at(cs, coords)
    P0 = np.zeros((coords.shape[0], 12))
    P0[:, :9] = c["F"].reshape(-1)
    P0[:, 9:] = np.round(U)
    args = (coords, P0, *half, f, gx, gy, gz, mask, g, mode, L, mf, bf, valid, 1e-2, 1e-3, 100, 5)
    ref = cpu_ref.icgn_12dof_parallel(*args)
    got = K.icgn_12dof_parallel(*args, variant=variant)
    assert_icgn_parity(ref, got)
    if key == "escape":
        assert np.any(ref[2] == 2)
    else:
        assert np.sum(ref[2] == 0) > (0.25 if "mask" in c else 0.5) * valid.sum()


@pytest.mark.parametrize("key,mu,noise", [("i3_plain", 1e-3, False), ("i3_mu", 0.5, False), ("i3_noise", 1e-3, True)])
def test_icgn3_matches_oracle(golden, key, mu, noise):
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(
        golden["coords"], 6, 6, 6, golden["f"], golden["gx"], golden["gy"], golden["gz"], MASK_NONE, 0.5, 1e12
    )
    pattern = golden["pattern_666_s1"]
    args = (
        golden["coords"], golden["i3_U_old"], golden["i3_F"], golden["i3_vdual"], 6, 6, 6, golden["f"],
        golden["gx"], golden["gy"], golden["gz"], MASK_NONE, golden["g"], 1, H, mf, bf, valid, mu, 1e-2, 1e-3,
        100, 5, 1, float(pattern[9, 9]), float(golden["stencil_noise_gain"]) if noise else 0.0, True,
    )  # fmt: skip
    assert_icgn_parity(cpu_ref.icgn_3dof_parallel(*args), K.icgn_3dof_parallel(*args))


@pytest.mark.parametrize("variant", ["fused", "fused_gather", "auto"])
@pytest.mark.parametrize("half", [(16, 16, 16), (24, 20, 10)], ids=["whole", "tiled"])
def test_icgn3_brick_paths_match_oracle(half, variant, tmp_path, monkeypatch):
    """Translation-only solves on the fused kernel's brick path: 33^3 subsets in one box, 49 x 41 x
    21 subsets in x/y boxes; fixed F from the imposed field, U_old one voxel off, a small dual."""
    F = _rot_z(3.0)
    cs = bead_case(shape=(112, 112, 112), F=F, t=(0.6, -0.4, 0.7), seed=9)
    ax = 28 + 14 * np.arange(5)
    coords = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3).astype(np.int64)
    f, g, gx, gy, gz = cs["f"], cs["g"], cs["gx"], cs["gy"], cs["gz"]
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, *half, f, gx, gy, gz, MASK_NONE, 0.5, 1e12)
    N = coords.shape[0]
    U_old = np.round(truth_at(cs, coords)) + np.array([0.4, -0.3, 0.2])
    Ff = np.tile(F.reshape(1, 9), (N, 1))
    vdual = np.random.default_rng(3).normal(0.0, 0.05, (N, 3))
    args = (coords, U_old, Ff, vdual, *half, f, gx, gy, gz, MASK_NONE, g, 1, H, mf, bf, valid, 1e-3, 1e-2, 1e-3, 100, 5)
    from pyaldvc_amd import autotune as AT

    monkeypatch.setenv("PYALDVC_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    AT.clear(memory_only=True)
    ref = cpu_ref.icgn_3dof_parallel(*args)
    got = K.icgn_3dof_parallel(*args, variant=variant)
    assert_icgn_parity(ref, got)
    assert np.sum(ref[2] == 0) > 0.5 * valid.sum()
    if variant == "auto":
        assert AT.last_report["key"].startswith("icgn3|")
        assert len(AT.last_report["seconds"]) == len(AT.CANDIDATES)


@pytest.mark.parametrize("variant", ["fused", "fused_gather"])
@pytest.mark.parametrize("split", [False, True], ids=["mask", "mask+split"])
def test_icgn3_masked_and_split_match_oracle(golden, variant, split):
    """Translation-only solves with the reference mask (and split keep rows) on both sampling
    paths of the fused kernel, against the oracle on the same inputs."""
    coords = golden["coords"]
    f, gx, gy, gz, m = golden["f"], golden["gx"], golden["gy"], golden["gz"], golden["split_mask"]
    kw = dict(split_index=np.arange(coords.shape[0], dtype=np.int64), split_keep=golden["split_m0_rows"]) if split else {}
    extra = (kw["split_index"], kw["split_keep"]) if split else ()
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, 6, 6, 6, f, gx, gy, gz, m, 0.5, 1e12, 1, *extra)
    args = (coords, golden["i3_U_old"], golden["i3_F"], golden["i3_vdual"], 6, 6, 6, f, gx, gy, gz, m, golden["g"], 1,
            H, mf, bf, valid, 1e-3, 1e-2, 1e-3, 100, 5)  # fmt: skip
    ref = cpu_ref.icgn_3dof_parallel(*args, **kw)
    assert_icgn_parity(ref, K.icgn_3dof_parallel(*args, **kw, variant=variant))
    assert np.sum(ref[2] == 0) > 0


def test_evaluate_zncc_matches_oracle(golden):
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(
        golden["coords"], 6, 6, 6, golden["f"], golden["gx"], golden["gy"], golden["gz"], MASK_NONE, 0.5, 1e12
    )
    args = (golden["coords"], golden["eval_P"], 6, 6, 6, golden["f"], MASK_NONE, golden["g"], 1, mf, bf, valid)
    z = K.evaluate_zncc_parallel(*args)
    zr = cpu_ref.evaluate_zncc_parallel(*args)
    assert np.array_equal(np.isnan(z), np.isnan(zr))
    ok = ~np.isnan(zr)
    assert np.max(np.abs(z[ok] - zr[ok])) < 1e-5


def test_zncc_map_matches_oracle(golden):
    hx, hy, hz, rx, ry, rz = (int(v) for v in golden["zd_params"])
    args = (golden["zd_f"], golden["g"], golden["zd_coords"], hx, hy, hz, rx, ry, rz,
            golden["zd_wx0"], golden["zd_wy0"], golden["zd_wz0"], golden["zd_tok"])  # fmt: skip
    got = K.zncc_direct(*args)
    ref = np.empty(golden["zd_maps"].shape)
    cpu_ref.zncc_direct(*args, ref)
    assert np.max(np.abs(got - ref)) < 2e-5


@pytest.mark.parametrize("h,r", [((8, 8, 8), (2, 2, 2)), ((4, 4, 4), (4, 4, 4)), ((3, 5, 2), (1, 3, 0)),
                                 ((5, 2, 4), (3, 1, 2)), ((8, 8, 8), (5, 5, 5)), ((2, 2, 2), (4, 0, 1))])
def test_zncc_map_small_and_large_windows_match_oracle(h, r):
    """Float32 maps against the float64 oracle: x radii up to 4 take the small-window kernel
    (window in LDS, work split over template planes, separable box sums), radius 5 the row
    kernel. A flat template, t_ok = 0 and a window at the volume's faces are included."""
    from pyaldvc_amd import synthetic as syn

    shape = (48, 52, 56)
    f_raw, g_raw = syn.bead_pair_np(shape, 600, 1.5, syn.translation((0.4, -1.3, 0.7)), seed=4)
    f, g = syn.normalize_np(f_raw), syn.normalize_np(g_raw)
    hx, hy, hz = h
    rx, ry, rz = r
    f[0:12, 0:12, 0:12] = 1.0  # flat template around the first node
    coords = np.array([[6, 6, 6], [28, 26, 24], [hx + rx, hy + ry, hz + rz],
                       [55 - hx - rx, 51 - hy - ry, 47 - hz - rz], [20, 30, 22]], dtype=np.int64)  # fmt: skip
    wx0, wy0, wz0 = coords[:, 0] - hx - rx, coords[:, 1] - hy - ry, coords[:, 2] - hz - rz
    ok = np.array([1, 1, 1, 1, 0], np.uint8)
    keep = (wx0 >= 0) & (wy0 >= 0) & (wz0 >= 0)
    coords, wx0, wy0, wz0, ok = coords[keep], wx0[keep], wy0[keep], wz0[keep], ok[keep]
    got = K.zncc_direct(f, g, coords, hx, hy, hz, rx, ry, rz, wx0, wy0, wz0, ok)
    ref = np.empty(got.shape)
    cpu_ref.zncc_direct(f, g, coords, hx, hy, hz, rx, ry, rz, wx0, wy0, wz0, ok, ref)
    assert np.max(np.abs(got - ref)) < 2e-5
    assert np.all(got[ok == 0] == -2.0)


@pytest.mark.parametrize("radius,shift", [((3, 3, 3), False), ((2, 3, 2), True), ((4, 4, 4), False)])
def test_ncc_search_matches_oracle(golden, radius, shift):
    from pyaldvc_amd import integer_search as gis

    sc = golden["ncc_coords"]
    sh = golden["ncc_shift0"] if shift else None
    subset = (8, 6, 8) if shift else (8, 8, 8)
    got = gis.ncc_search(golden["f"], golden["g"], sc, subset, radius, sh)
    ref = ncc_ref.ncc_search(golden["f"], golden["g"], sc, subset, radius, sh)
    assert np.array_equal(got["ok"], ref["ok"])
    assert np.array_equal(got["clipped"], ref["clipped"])
    ok = ref["ok"]
    assert np.array_equal(np.round(got["disp"][ok]), np.round(ref["disp"][ok]))
    assert np.max(np.abs(got["disp"][ok] - ref["disp"][ok])) < 1e-3
    assert np.max(np.abs(got["cc"][ok] - ref["cc"][ok])) < 1e-5


def test_ncc_search_lattice_integer_peaks(case):
    """Integer peaks of a whole lattice equal the float64 oracle's; sub-voxel within 1e-3 voxel."""
    from pyaldvc_amd import integer_search as gis

    coords = lattice(case["shape"], 16, 8)
    got = gis.ncc_search(case["f"], case["g"], coords, (16, 16, 16), (3, 3, 3))
    ref = ncc_ref.ncc_search(case["f"], case["g"], coords, (16, 16, 16), (3, 3, 3))
    assert np.array_equal(got["ok"], ref["ok"])
    ok = ref["ok"]
    assert np.max(np.abs(got["disp"][ok] - ref["disp"][ok])) < 1e-3


# --------------------------------------------------------------------------- bfloat16 MFMA maps


def test_zncc_map_bf16_close_to_oracle(golden):
    """BF16 operands with float32 accumulation and float32 box sums: maps within bf16 rounding."""
    hx, hy, hz, rx, ry, rz = (int(v) for v in golden["zd_params"])
    args = (golden["zd_f"], golden["g"], golden["zd_coords"], hx, hy, hz, rx, ry, rz,
            golden["zd_wx0"], golden["zd_wy0"], golden["zd_wz0"], golden["zd_tok"])  # fmt: skip
    got = K.zncc_direct(*args, precision="bf16")
    ref = np.empty(golden["zd_maps"].shape)
    cpu_ref.zncc_direct(*args, ref)
    assert np.array_equal(got == -2.0, ref == -2.0)  # t_ok = 0 and flat templates
    assert np.max(np.abs(got - ref)) < 1e-2
    for n in range(ref.shape[0]):
        if ref[n, 0, 0, 0] > -1.5:
            assert np.argmax(got[n]) == np.argmax(ref[n])


def test_ncc_search_bf16_lattice_peaks(case):
    """Integer peaks of a whole lattice from the BF16 maps equal the float64 oracle's."""
    from pyaldvc_amd import integer_search as gis

    coords = lattice(case["shape"], 16, 8)
    got = gis.ncc_search(case["f"], case["g"], coords, (16, 16, 16), (3, 3, 3), precision="bf16")
    ref = ncc_ref.ncc_search(case["f"], case["g"], coords, (16, 16, 16), (3, 3, 3))
    assert np.array_equal(got["ok"], ref["ok"])
    ok = ref["ok"]
    d = np.abs(got["disp"][ok] - ref["disp"][ok])
    assert np.mean(np.all(d < 0.5, axis=1)) >= 0.99  # same integer peak
    assert np.median(d) < 0.02


def test_zncc_map_bf16_skewed_planes_small_shape_matches_fp32():
    """13^3 templates at radius 8 (17^3 offsets: two column tiles, a last group of one output
    plane) through the skewed-plane tiling, and a node without a template (t_ok = 0, map -2):
    within bf16 rounding of the float32 maps, same integer peaks."""
    from pyaldvc_amd import synthetic as syn

    shape = (96, 96, 96)
    f_raw, g_raw = syn.bead_pair_np(shape, int(0.004 * np.prod(shape)), 1.5, syn.translation((2.3, -1.6, 3.2)), seed=5)
    f, g = syn.normalize_np(f_raw), syn.normalize_np(g_raw)
    coords = np.array([[48, 48, 48], [30, 40, 60], [62, 55, 35], [40, 66, 50], [50, 30, 30]], dtype=np.int64)
    h, r = 6, 8
    wx0, wy0, wz0 = (coords[:, k] - h - r for k in range(3))
    ok = np.array([1, 1, 0, 1, 1], np.uint8)
    m32 = K.zncc_direct(f, g, coords, h, h, h, r, r, r, wx0, wy0, wz0, ok)
    m16 = K.zncc_direct(f, g, coords, h, h, h, r, r, r, wx0, wy0, wz0, ok, precision="bf16")
    assert np.all(m16[2] == -2.0) and np.all(m32[2] == -2.0)
    assert np.max(np.abs(m16 - m32)) < 1e-2
    for n in (0, 1, 3, 4):
        assert np.argmax(m16[n]) == np.argmax(m32[n])


@pytest.mark.parametrize("radius", [(16, 16, 16), (9, 12, 5)])
def test_zncc_map_bf16_large_template_matches_fp32(radius):
    """24^3 subsets (25^3 templates, two K slices) with 16-voxel radii, the shape of the BF16
    NCC configuration, on a 128^3 bead pair: same integer peaks as the float32 maps."""
    from pyaldvc_amd import synthetic as syn

    shape = (128, 128, 128)
    disp = syn.sinusoid_z(6.0, 256.0)
    f_raw, g_raw = syn.bead_pair_np(shape, int(0.004 * np.prod(shape)), 1.5, disp, seed=21)
    f, g = syn.normalize_np(f_raw), syn.normalize_np(g_raw)
    coords = np.array([[64, 64, 64], [40, 50, 60], [80, 70, 48], [56, 88, 72], [72, 40, 80]], dtype=np.int64)
    rx, ry, rz = radius
    h = 12
    wx0 = coords[:, 0] - h - rx
    wy0 = coords[:, 1] - h - ry
    wz0 = coords[:, 2] - h - rz
    ok = np.ones(coords.shape[0], np.uint8)
    m32 = K.zncc_direct(f, g, coords, h, h, h, rx, ry, rz, wx0, wy0, wz0, ok)
    m16 = K.zncc_direct(f, g, coords, h, h, h, rx, ry, rz, wx0, wy0, wz0, ok, precision="bf16")
    assert np.max(np.abs(m16 - m32)) < 1e-2
    for n in range(coords.shape[0]):
        assert np.argmax(m16[n]) == np.argmax(m32[n])
    # the float32 map of one node against the float64 oracle
    ref = np.empty((1,) + m32.shape[1:])
    cpu_ref.zncc_direct(f, g, coords[:1], h, h, h, rx, ry, rz, wx0[:1], wy0[:1], wz0[:1], ok[:1], ref)
    assert np.max(np.abs(m32[0] - ref[0])) < 2e-5


def test_autotuned_launch_shapes_agree_and_cache(tmp_path, monkeypatch):
    """Every candidate launch shape of the fused kernel gives the oracle's statuses and the same
    translations to 1e-4 voxel; the tuner picks one of them, caches it, and does not time again."""
    from pyaldvc_amd import autotune as AT

    monkeypatch.setenv("PYALDVC_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    AT.clear(memory_only=True)
    c = BRICK_CASES["fits"]
    cs = bead_case(shape=(112, 112, 112), F=c["F"], t=c["t"], seed=9)
    ax = 24 + 16 * np.arange(5)
    coords = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3).astype(np.int64)
    half = (16, 16, 16)
    f, g, gx, gy, gz = cs["f"], cs["g"], cs["gx"], cs["gy"], cs["gz"]
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, *half, f, gx, gy, gz, MASK_NONE, 0.5, 1e12)
    P0 = np.zeros((coords.shape[0], 12))
    P0[:, :9] = c["F"].reshape(-1)
    P0[:, 9:] = np.round(truth_at(cs, coords))
    args = (coords, P0, *half, f, gx, gy, gz, MASK_NONE, g, 1, L, mf, bf, valid, 1e-2, 1e-3, 100, 5)
    ref = cpu_ref.icgn_12dof_parallel(*args)
    base = K.icgn_12dof_parallel(*args, tuning=AT.DEFAULT)
    for cand in AT.CANDIDATES:
        got = K.icgn_12dof_parallel(*args, tuning=cand)
        assert_icgn_parity(ref, got)
        ok = (got[2] == 0) & (base[2] == 0)
        assert np.array_equal(got[2], base[2]), cand
        assert np.max(np.abs(got[0][ok] - base[0][ok])) < 1e-4, cand
    auto = K.icgn_12dof_parallel(*args, variant="auto")
    assert_icgn_parity(ref, auto)
    rep = dict(AT.last_report)
    assert rep["choice"] in {cand.label() for cand in AT.CANDIDATES} and len(rep["seconds"]) == len(AT.CANDIDATES)
    assert (tmp_path / "tune.json").exists()
    AT.last_report.clear()
    K.icgn_12dof_parallel(*args, variant="auto")
    assert AT.last_report == {}  # served from the cache
