"""Tiled local steps on the GPU (pyaldvc_amd/tiling.py + local.py): tiling changes the memory and
not the answer (the reference's tests/test_tiling.py), including a volume beyond 2^31 voxels that
the kernels only take in boxes."""

from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from pyaldvc_amd import kernels as K
from pyaldvc_amd import local as LOC
from pyaldvc_amd import tiling as T

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def gd():
    return np.load(os.path.join(HERE, "golden", "golden_driver.npz"))


def _para(**kw):
    base = dict(winsize=(12, 12, 12), subset_stride=1, min_valid_ratio=0.5, hessian_cond_max=1e12, subset_split=False,
                interp_method="cubic", icgn_tol=1e-2, icgn_dp_tol=1e-3, icgn_max_iter=100, icgn_patience=5,
                icgn_predictive_stop=True, icgn_noise_hessian=False, local_outlier_threshold=2.0, tile_local=0)  # fmt: skip
    base.update(kw)
    return SimpleNamespace(**base)


def _case(gd, p):
    mask = gd[f"{p}_mask"].astype(bool) if p == "drvm" else None
    edge = gd[f"{p}_edge_ok"] if p == "drvm" else None
    mesh = LOC.NodeMesh(gd[f"{p}_coords"], tuple(gd[f"{p}_grid_shape"]), gd[f"{p}_node_valid_in"].copy(), edge)
    ref = SimpleNamespace(f=gd["drv_f"], mask=mask, gradient_mode="stored")
    return mesh, ref


def _same_context(a, b):
    for k in ("H_all", "L_all", "meanf", "bottomf", "n_valid", "valid"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert (a.split_fraction is None) == (b.split_fraction is None)
    if a.split_fraction is not None:
        assert np.array_equal(a.split_fraction, b.split_fraction, equal_nan=True)
        ia, ib = a.split_index, b.split_index
        assert np.array_equal(ia >= 0, ib >= 0)
        rows = ia >= 0
        assert np.array_equal(a.split_keep[ia[rows]], b.split_keep[ib[rows]])


@pytest.mark.parametrize("p", ["drv", "drvm"])
@pytest.mark.parametrize("mode", ["stored", "on_the_fly", "stored_separable"])
def test_tiled_local_steps_equal_the_whole_volume(gd, p, mode, monkeypatch):
    """Precompute, the 12-DOF local step and the 3-DOF ADMM step in boxes of at most 38 voxels per
    edge (several boxes with open faces) give bit-identical results to the whole-volume run; also
    with the separable precompute and reference-moment passes, which these node counts would not
    take by themselves (each node's sums come from its own voxels in a fixed order)."""
    if mode == "stored_separable":
        monkeypatch.setattr(K, "separable_wanted", lambda n, half, stride, dims, stored: bool(stored) and stride == 1)
    mesh, ref = _case(gd, p)
    if mode == "on_the_fly":
        ph = np.zeros((1, 1, 1), np.float32)
        ref = SimpleNamespace(f=ref.f, gx=ph, gy=ph, gz=ph, mask=ref.mask if ref.mask is not None else np.ones((1, 1, 1), np.uint8))
    split = p == "drvm"
    whole_para, tiled_para = _para(subset_split=split), _para(subset_split=split, tile_local=38)
    plan = T.plan_tiles(mesh.grid_shape, np.round(mesh.coordinates).astype(np.int64), gd["drv_f"].shape, tiled_para,
                        (6, 6, 6), disp=gd[f"{p}_U0"])  # fmt: skip
    assert len(plan) > 1 and any(b.open for _n, b in plan)
    c0 = LOC.precompute_local_context(mesh, ref, whole_para)
    c1 = LOC.precompute_local_context(mesh, ref, tiled_para)
    assert getattr(c1, "n_tiles", 1) > 1
    assert c0.separable is c1.separable is (mode == "stored_separable")
    _same_context(c0, c1)
    r0 = LOC.local_icgn(c0, ref, gd["drv_g"], gd[f"{p}_U0"], whole_para, mesh)
    r1 = LOC.local_icgn(c1, ref, gd["drv_g"], gd[f"{p}_U0"], tiled_para, mesh)
    for a, b in ((r0[0], r1[0]), (r0[1], r1[1]), (r0[3], r1[3])):
        assert np.array_equal(a, b)
    for k in ("status", "n_iter"):
        assert np.array_equal(getattr(r0[2], k), getattr(r1[2], k))
    assert np.array_equal(r0[2].zncc, r1[2].zncc, equal_nan=True)
    mesh.node_valid = c0.valid.copy()
    s0 = LOC.subpb1_solver(c0, ref, gd["drv_g"], gd[f"{p}_U"], gd[f"{p}_F"], gd[f"{p}_sp1_vdual"], 1e-3, whole_para, mesh)
    s1 = LOC.subpb1_solver(c1, ref, gd["drv_g"], gd[f"{p}_U"], gd[f"{p}_F"], gd[f"{p}_sp1_vdual"], 1e-3, tiled_para, mesh)
    assert np.array_equal(s0[0], s1[0]) and np.array_equal(s0[2], s1[2])
    assert np.array_equal(s0[1].status, s1[1].status) and np.array_equal(s0[1].n_iter, s1[1].n_iter)


def test_nodes_that_leave_a_short_halo_are_solved_again(gd, monkeypatch):
    """With the deformed halo cut to the reference one, samples that move past a tile's open face
    come back out of bounds; they are solved again on the whole volume and the answer is unchanged
    (local_icgn.py:440-452)."""
    mesh, ref = _case(gd, "drv")
    para = _para(tile_local=24)
    c0 = LOC.precompute_local_context(mesh, ref, _para())
    r0 = LOC.local_icgn(c0, ref, gd["drv_g"], gd["drv_U0"], _para(), mesh)
    real = T.halo
    monkeypatch.setattr(T, "halo", lambda para_, half, disp=None, deformed=False: real(para_, half))
    seen = {}
    orig = LOC._solve_tiled

    def spy(ctx, mesh_, para_, disp, run, n_out):
        def run_spy(nodes_d, box):
            out = run(nodes_d, box)
            seen.setdefault("oob", 0)
            seen["oob"] += int((out[2] == K.STATUS_OUT_OF_BOUNDS).sum()) if box.open else 0
            return out

        return orig(ctx, mesh_, para_, disp, run_spy, n_out)

    monkeypatch.setattr(LOC, "_solve_tiled", spy)
    r1 = LOC.local_icgn(c0, ref, gd["drv_g"], gd["drv_U0"], para, mesh)
    assert seen.get("oob", 0) > 0  # the short halo did cut some subsets
    assert np.array_equal(r0[0], r1[0]) and np.array_equal(r0[2].status, r1[2].status)


def test_a_volume_beyond_two_to_the_31_voxels_is_solved_in_boxes():
    """80 x 2048 x 13200 voxels (2.16e9, past the kernels' 32-bit offsets): the whole-volume solve is
    refused, the driver plans boxes below the limit, and every node equals the same node solved
    untiled on a crop that holds it. One layer of nodes at z = 40 (windows clear of the z faces)."""
    dev = K.device()
    shape = (80, 2048, 13200)
    assert int(np.prod(shape)) > T.KERNEL_MAX_VOXELS
    gen = torch.Generator(device=dev)
    gen.manual_seed(3)
    block = torch.randn((80, 2048, 1320), device=dev, generator=gen)
    for _ in range(2):  # a smooth texture: box filters along each axis
        block = torch.nn.functional.avg_pool3d(block[None, None], 3, stride=1, padding=1, count_include_pad=False)[0, 0]
    block = (block - block.mean()) / block.std()
    f = block.repeat(1, 1, 10).contiguous()
    del block
    g = torch.roll(f, shifts=(1, -2, 1), dims=(0, 1, 2))  # u = (1, -2, 1) voxels
    w, step = 16, 64
    gx_n, gy_n = (shape[2] - w) // step + 1, (shape[1] - w) // step + 1
    Y, X = np.meshgrid(w // 2 + step * np.arange(gy_n), w // 2 + step * np.arange(gx_n), indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel(), np.full(X.size, 40)]).astype(np.int64)
    grid = (1, gy_n, gx_n)
    N = coords.shape[0]
    mesh = LOC.NodeMesh(coords.astype(np.float64), grid, np.ones(N, dtype=bool))
    para = _para(winsize=(w, w, w), winstepsize=(step,) * 3, local_outlier_threshold=0.0)
    ref = SimpleNamespace(f=f, mask=None, gradient_mode="on_the_fly")
    ctx = LOC.precompute_local_context(mesh, ref, para)
    assert getattr(ctx, "n_tiles", 1) > 1
    U0 = np.tile(np.array([1.4, -1.6, 0.7]), (N, 1))
    U, F, info, bad = LOC.local_icgn(ctx, ref, g, U0, para, mesh)
    inner = np.all((coords[:, :2] >= 24) & (coords[:, :2] <= np.array([13200, 2048]) - 24), axis=1)  # away from the roll's wrap
    conv = (info.status == 0) & inner
    st_counts = dict(zip(*(a.tolist() for a in np.unique(info.status[inner], return_counts=True))))
    assert conv.sum() > 0.9 * inner.sum(), (st_counts, int(ctx.valid[inner].sum()), U[inner][:3], info.n_iter[inner][:5])
    assert np.max(np.abs(U[conv] - np.array([1.0, -2.0, 1.0]))) < 1e-2
    # the same nodes untiled on a crop of 4000 voxels along x that holds them
    sel = np.flatnonzero(coords[:, 0] < 3900)
    assert int(np.prod(f[:, :, :4000].shape)) < T.KERNEL_MAX_VOXELS
    crop = (slice(None), slice(None), slice(0, 4000))
    fc, gc = f[crop].contiguous(), g[crop].contiguous()
    P0 = np.zeros((sel.size, 12))
    P0[:, 9:] = U0[sel]
    nomask = np.ones((1, 1, 1), np.uint8)
    ph = torch.zeros((1, 1, 1), device=dev)
    cd = torch.from_numpy(coords[sel]).to(dev)
    H, L, mf, bf, nv, valid = K.precompute_nodes(cd, 8, 8, 8, fc, ph, ph, ph, nomask, 0.5, 1e12)
    P, it, st, zn = K.icgn_12dof_parallel(cd, torch.from_numpy(P0).to(dev), 8, 8, 8, fc, ph, ph, ph, nomask, gc, 1, L, mf,
                                          bf, valid, 1e-2, 1e-3, 100, 5, 1, H, variant="fused_gather")  # fmt: skip
    st = st.cpu().numpy()
    assert np.array_equal(st, info.status[sel])
    # the boxes of the big volume exceed 2^30 voxels and take the pointer-gather form of the kernel,
    # the 4000-voxel crop (under 2^30) the buffer-gather form: the same samples in the same order, so
    # the boxes must reproduce the crop bit for bit. (The LDS-brick form rounds its taps differently,
    # and a node whose last step lies at a stopping threshold may then stop one update apart.)
    ok = ~bad[sel]
    assert ok.sum() > 0.9 * sel.size
    assert np.array_equal(P.cpu().numpy()[ok, 9:], U[sel][ok])
