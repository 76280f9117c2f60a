"""The oracle against the reference's own outputs (golden vectors written by tests/golden/make_golden.py).

The vectors are the reference CPU kernels' results on small seeded bead pairs; the oracle must
reproduce statuses and iteration counts exactly and values to float64 round-off.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu_ref, ncc_ref

MASK_NONE = np.ones((1, 1, 1), dtype=np.uint8)
OTF = np.zeros((1, 1, 1), dtype=np.float32)


def test_gradients_match_reference(golden):
    gx, gy, gz = cpu_ref.compute_gradients(golden["f"])
    for a, b in ((gx, golden["gx"]), (gy, golden["gy"]), (gz, golden["gz"])):
        # float64 stencil then float32 rounding; scipy's correlate1d sums the same taps in another order
        assert np.max(np.abs(a - b)) <= 2e-7 * max(1.0, float(np.max(np.abs(b))))


def test_normalize_matches_reference(golden):
    from pyaldvc_amd.synthetic import normalize_np

    for raw, ref in (("f_raw", "f"), ("g_raw", "g")):
        assert np.allclose(normalize_np(golden[raw]), golden[ref], atol=2e-6)


@pytest.mark.parametrize("mode,name", [(0, "linear"), (1, "cubic"), (2, "bspline")])
def test_interp_matches_reference(golden, mode, name):
    src = golden["g_bspline"] if mode == 2 else golden["g"]
    got = cpu_ref.sample_points(src, golden["pts_z"], golden["pts_y"], golden["pts_x"], mode)
    ref = golden[f"interp_{name}"]
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.max(np.abs(got[ok] - ref[ok])) < 1e-12


def test_compose_and_cholesky_match_reference(golden):
    for i in range(golden["compose_P"].shape[0]):
        P = golden["compose_P"][i].copy()
        ok = cpu_ref.compose_warp_inplace(P, golden["compose_dP"][i])
        assert ok == bool(golden["compose_ok"][i])
        assert np.max(np.abs(P - golden["compose_out"][i])) < 1e-15
    for i in range(golden["chol_H"].shape[0]):
        ok, L = cpu_ref.cholesky12(golden["chol_H"][i])
        assert ok == bool(golden["chol_ok"][i])
        if ok:
            assert np.max(np.abs(L - golden["chol_L"][i])) < 1e-12


PRE_CASES = {
    "pre_plain": dict(half=(6, 6, 6), stride=1, grads="stored", mask=None),
    "pre_mask": dict(half=(6, 6, 6), stride=1, grads="stored", mask="mask"),
    "pre_stride2": dict(half=(6, 5, 4), stride=2, grads="stored", mask=None),
    "pre_otf": dict(half=(6, 6, 6), stride=1, grads="otf", mask=None),
}


def _pre_args(golden, key):
    c = PRE_CASES[key]
    grads = (golden["gx"], golden["gy"], golden["gz"]) if c["grads"] == "stored" else (OTF, OTF, OTF)
    mask = golden[c["mask"]] if c["mask"] else MASK_NONE
    return c["half"], c["stride"], grads, mask


@pytest.mark.parametrize("key", list(PRE_CASES))
def test_precompute_matches_reference(golden, key):
    half, stride, grads, mask = _pre_args(golden, key)
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(golden["coords"], *half, golden["f"], *grads, mask, 0.5, 1e12, stride)
    assert np.array_equal(valid, golden[f"{key}_valid"])
    assert np.array_equal(nv, golden[f"{key}_nvalid"])
    scale = np.max(np.abs(golden[f"{key}_H"]))
    assert np.max(np.abs(H - golden[f"{key}_H"])) <= 1e-12 * scale
    assert np.max(np.abs(mf - golden[f"{key}_meanf"])) < 1e-12
    assert np.max(np.abs(bf - golden[f"{key}_bottomf"])) < 1e-10
    assert np.max(np.abs(L[valid] - golden[f"{key}_L"][valid])) <= 1e-10 * np.sqrt(scale)


I12_CASES = {
    "i12_cubic": dict(pre="pre_plain", g="g", mode=1),
    "i12_linear": dict(pre="pre_plain", g="g", mode=0),
    "i12_bspline": dict(pre="pre_plain", g="g_bspline", mode=2),
    "i12_mask": dict(pre="pre_mask", g="g", mode=1),
    "i12_defmask": dict(pre="pre_plain", g="g_nan", mode=1),
    "i12_stride2": dict(pre="pre_stride2", g="g", mode=1),
    "i12_otf": dict(pre="pre_otf", g="g", mode=1),
    "i12_status": dict(pre="pre_plain", g="g", mode=1, far=True),
    "i12_noise": dict(pre="pre_plain", g="g", mode=1, noise=True),
    "i12_nopred": dict(pre="pre_plain", g="g", mode=1, predictive=False),
    "i12_stall": dict(pre="pre_plain", g="g_noise", mode=1),
}


def start_params(golden, far=False):
    P0 = np.zeros((golden["coords"].shape[0], 12))
    P0[:, 9:] = np.round(golden["U_true"])
    if far:
        P0[0, 9] = 60.0
        P0[1, 10] = np.nan
    return P0


@pytest.mark.parametrize("key", list(I12_CASES))
def test_icgn12_matches_reference(golden, key):
    c = I12_CASES[key]
    half, stride, grads, mask = _pre_args(golden, c["pre"])
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(golden["coords"], *half, golden["f"], *grads, mask, 0.5, 1e12, stride)
    noise = c.get("noise", False)
    pattern = golden["pattern_666_s1"]
    P, it, st, zc = cpu_ref.icgn_12dof_parallel(
        golden["coords"], start_params(golden, c.get("far", False)), *half, golden["f"], *grads, mask, golden[c["g"]],
        c["mode"], L, mf, bf, valid, 1e-2, 1e-3, 100, 5, stride, H, pattern if noise else None,
        float(golden["stencil_noise_gain"]) if noise else 0.0, c.get("predictive", True),
    )  # fmt: skip
    assert np.array_equal(st, golden[f"{key}_status"])
    assert np.array_equal(it, golden[f"{key}_n_iter"])
    ref_P = golden[f"{key}_P"]
    fin = np.isfinite(ref_P)
    assert np.array_equal(fin, np.isfinite(P))
    assert np.max(np.abs(P[fin] - ref_P[fin])) < 1e-10
    ref_z = golden[f"{key}_zncc"]
    assert np.array_equal(np.isnan(zc), np.isnan(ref_z))
    ok = ~np.isnan(ref_z)
    assert np.max(np.abs(zc[ok] - ref_z[ok]), initial=0.0) < 1e-10


@pytest.mark.parametrize("key,mu,noise", [("i3_plain", 1e-3, False), ("i3_mu", 0.5, False), ("i3_noise", 1e-3, True)])
def test_icgn3_matches_reference(golden, key, mu, noise):
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(
        golden["coords"], 6, 6, 6, golden["f"], golden["gx"], golden["gy"], golden["gz"], MASK_NONE, 0.5, 1e12
    )
    pattern = golden["pattern_666_s1"]
    U, it, st, zc = cpu_ref.icgn_3dof_parallel(
        golden["coords"], golden["i3_U_old"], golden["i3_F"], golden["i3_vdual"], 6, 6, 6, golden["f"],
        golden["gx"], golden["gy"], golden["gz"], MASK_NONE, golden["g"], 1, H, mf, bf, valid, mu, 1e-2, 1e-3,
        100, 5, 1, float(pattern[9, 9]), float(golden["stencil_noise_gain"]) if noise else 0.0, True,
    )  # fmt: skip
    assert np.array_equal(st, golden[f"{key}_status"])
    assert np.array_equal(it, golden[f"{key}_n_iter"])
    assert np.max(np.abs(U - golden[f"{key}_U"])) < 1e-10
    ref_z = golden[f"{key}_zncc"]
    ok = ~np.isnan(ref_z)
    assert np.array_equal(np.isnan(zc), np.isnan(ref_z))
    assert np.max(np.abs(zc[ok] - ref_z[ok]), initial=0.0) < 1e-10


def test_evaluate_zncc_matches_reference(golden):
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(
        golden["coords"], 6, 6, 6, golden["f"], golden["gx"], golden["gy"], golden["gz"], MASK_NONE, 0.5, 1e12
    )
    z = cpu_ref.evaluate_zncc_parallel(golden["coords"], golden["eval_P"], 6, 6, 6, golden["f"], MASK_NONE, golden["g"], 1, mf, bf, valid)
    ref = golden["eval_zncc"]
    assert np.array_equal(np.isnan(z), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.max(np.abs(z[ok] - ref[ok])) < 1e-12


@pytest.mark.parametrize("key,half,stride,margin", [("split_m0", (6, 6, 6), 1, 0), ("split_m3", (6, 6, 6), 1, 3), ("split_s2", (6, 5, 4), 2, 0)])
def test_split_rows_match_reference(golden, key, half, stride, margin):
    cand = np.arange(golden["coords"].shape[0], dtype=np.int64)
    rows, nk, nm = cpu_ref.build_split_rows(golden["coords"], cand, *half, stride, golden["split_mask"], margin)
    assert np.array_equal(rows, golden[f"{key}_rows"])
    assert np.array_equal(nk, golden[f"{key}_nkeep"])
    assert np.array_equal(nm, golden[f"{key}_ninmask"])


def test_split_precompute_and_icgn_match_reference(golden):
    coords = golden["coords"]
    si = np.arange(coords.shape[0], dtype=np.int64)
    rows = golden["split_m0_rows"]
    f, gx, gy, gz, m = golden["f"], golden["gx"], golden["gy"], golden["gz"], golden["split_mask"]
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, 6, 6, 6, f, gx, gy, gz, m, 0.5, 1e12, 1, si, rows)
    assert np.array_equal(valid, golden["pre_split_valid"])
    assert np.array_equal(nv, golden["pre_split_nvalid"])
    assert np.max(np.abs(H - golden["pre_split_H"])) <= 1e-12 * np.max(np.abs(H))
    P, it, st, zc = cpu_ref.icgn_12dof_parallel(
        coords, start_params(golden), 6, 6, 6, f, gx, gy, gz, m, golden["g"], 1, L, mf, bf, valid,
        1e-2, 1e-3, 100, 5, 1, None, None, 0.0, True, si, rows,
    )  # fmt: skip
    assert np.array_equal(st, golden["i12_split_status"])
    assert np.array_equal(it, golden["i12_split_n_iter"])
    assert np.max(np.abs(P - golden["i12_split_P"])) < 1e-10


def test_zncc_direct_matches_reference(golden):
    hx, hy, hz, rx, ry, rz = (int(v) for v in golden["zd_params"])
    out = np.empty(golden["zd_maps"].shape)
    cpu_ref.zncc_direct(
        golden["zd_f"], golden["g"], golden["zd_coords"], hx, hy, hz, rx, ry, rz,
        golden["zd_wx0"], golden["zd_wy0"], golden["zd_wz0"], golden["zd_tok"], out,
    )  # fmt: skip
    assert np.max(np.abs(out - golden["zd_maps"])) < 1e-12


@pytest.mark.parametrize("key", ["ncc_r3", "ncc_shift"])
def test_ncc_search_matches_reference(golden, key):
    """The reference ran its FFT engine here (float32 transforms); integer peaks must be identical."""
    sc = golden["ncc_coords"]
    if key == "ncc_r3":
        res = ncc_ref.ncc_search(golden["f"], golden["g"], sc, (8, 8, 8), (3, 3, 3))
    else:
        res = ncc_ref.ncc_search(golden["f"], golden["g"], sc, (8, 6, 8), (2, 3, 2), golden["ncc_shift0"])
    assert np.array_equal(res["ok"], golden[f"{key}_ok"])
    assert np.array_equal(res["clipped"], golden[f"{key}_clipped"])
    ok = res["ok"]
    d, dr = res["disp"][ok], golden[f"{key}_disp"][ok]
    assert np.array_equal(np.round(d), np.round(dr))
    assert np.max(np.abs(d - dr)) < 2e-4
    assert np.max(np.abs(res["cc"][ok] - golden[f"{key}_cc"][ok])) < 1e-5
    assert np.max(np.abs(res["pce"][ok] - golden[f"{key}_pce"][ok]) / golden[f"{key}_pce"][ok]) < 1e-4


def test_ncc_search_expanding_matches_reference(golden):
    res = ncc_ref.ncc_search_expanding(golden["f"], golden["g"], golden["ncc_coords"], (8, 8, 8), (1, 1, 1))
    assert res["expansions"] == int(golden["ncc_exp_expansions"])
    assert tuple(res["radius"]) == tuple(golden["ncc_exp_radius"])
    assert np.array_equal(res["ok"], golden["ncc_exp_ok"])
    assert np.array_equal(res["clipped"], golden["ncc_exp_clipped"])
    ok = res["ok"]
    assert np.max(np.abs(res["disp"][ok] - golden["ncc_exp_disp"][ok])) < 2e-4
