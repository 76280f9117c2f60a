"""ADMM iterations on the GPU (needs an MI355X).

Each device piece against its host restatement, which the CPU tests pin to the reference and
SciPy: the spring fill against the oracle's ``fill_nan_grid`` (uncut and cut meshes), the local-step
cleanup against the oracle's ``universal_median_test`` (identical flags: index work), the global step
against the host-controlled PCG and nodal gradient, the dual update against NumPy. The reference's
own ``subpb1_solver`` vectors cover the cleanup inside ``local.subpb1_solver`` (test_gpu_driver).
The whole loop (``AdmmLoop``, eager and replayed from a HIP graph, several replays with other work
between) against the reference loop of core/pipeline.py:290-358 restated on the host pieces.
"""

from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = pytest.importorskip("pyaldvc_amd.kernels")
from pyaldvc_amd import admm as A  # noqa: E402
from oracle import cleanup_ref as C  # noqa: E402
from pyaldvc_amd import fem as FEM  # noqa: E402
from pyaldvc_amd import local as LOC  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def _field(shape, seed, outliers=0.03):
    """A smooth (N, 3) node field with planted outliers."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    f = np.stack([gaussian_filter(rng.normal(0, 3, shape), 2.0) for _ in range(3)], -1)
    f += np.linspace(0, 1, shape[2])[None, None, :, None]
    out = rng.random(shape) < outliers
    f[out] += rng.normal(0, 5, (int(out.sum()), 3))
    return f.reshape(-1, 3)


def _edges(shape, seed, frac=0.15):
    rng = np.random.default_rng(seed)
    return (rng.random(shape + (3,)) > frac).astype(bool)


@pytest.mark.parametrize("cut", [False, True])
@pytest.mark.parametrize("shape,holes", [((9, 11, 13), 0.3), ((6, 7, 5), 0.8), ((1, 12, 17), 0.25), ((24, 30, 36), 0.5)])
def test_spring_fill_matches_host(shape, holes, cut):
    rng = np.random.default_rng(3)
    vals = _field(shape, 1, 0.0)
    for c in range(3):  # per-column holes, plus a hole shared by all columns
        vals[rng.random(vals.shape[0]) < holes / 3, c] = np.nan
    vals[rng.random(vals.shape[0]) < holes / 2] = np.nan
    edge = _edges(shape, 4) if cut else None
    got = A.spring_fill(vals, shape, edge_ok=edge)
    for c in range(3):
        ref = C.fill_nan_grid(vals[:, c].reshape(shape), edge_ok=edge).ravel()
        assert np.max(np.abs(got[:, c] - ref)) < 1e-8 * max(1.0, np.max(np.abs(ref))), c
    known = ~np.isnan(vals)
    assert np.array_equal(got[known], vals[known])


@pytest.mark.parametrize("n_unknown", [7168, 7169])
def test_spring_fill_either_side_of_the_one_workgroup_limit(n_unknown):
    """Up to 7,168 unknown nodes the fill runs in one workgroup per column, beyond that over the
    whole grid's launch shape: both against the oracle."""
    shape = (20, 24, 30)
    N = int(np.prod(shape))
    rng = np.random.default_rng(n_unknown)
    vals = _field(shape, 2, 0.0)
    vals[rng.choice(N, n_unknown, replace=False)] = np.nan
    got = A.spring_fill(vals, shape)
    for c in range(3):
        ref = C.fill_nan_grid(vals[:, c].reshape(shape)).ravel()
        assert np.max(np.abs(got[:, c] - ref)) < 1e-8 * max(1.0, np.max(np.abs(ref))), c
    known = ~np.isnan(vals)
    assert np.array_equal(got[known], vals[known])


def _host_cleanup(U, U_hat, status, shape, threshold, edge):
    good = status == 0
    bad = ~good
    if threshold > 0 and good.sum() > 27:
        bad |= C.universal_median_test(U.reshape(shape + (3,)), good.reshape(shape), threshold, edge_ok=edge).ravel()
    out = U.copy()
    out[bad] = U_hat[bad]
    return out, bad


@pytest.mark.parametrize("cut", [False, True])
@pytest.mark.parametrize("seed,shape", [(0, (10, 12, 14)), (1, (10, 12, 14)), (2, (10, 12, 14)), (3, (24, 30, 36))])
def test_cleanup_matches_host_median_test(seed, shape, cut):
    N = int(np.prod(shape))
    rng = np.random.default_rng(seed)
    U = _field(shape, seed)
    U[rng.random(N) < 0.02, 1] = np.nan  # a NaN component at a few converged nodes
    U_hat = U + rng.normal(0, 0.1, U.shape)
    status = np.where(rng.random(N) < 0.2, rng.integers(1, 8, N), 0).astype(np.int8)
    edge = _edges(shape, seed + 10) if cut else None
    flagged = 0
    for thr in (2.0, 0.5, 0.0):
        out, bad = A.subpb1_cleanup(U, U_hat, status, shape, thr, edge_ok=edge)
        ref_out, ref_bad = _host_cleanup(U, U_hat, status, shape, thr, edge)
        assert np.array_equal(bad, ref_bad), (thr, int(np.sum(bad != ref_bad)))
        assert np.array_equal(out, ref_out, equal_nan=True)
        flagged += int(np.sum(ref_bad & (status == 0)))
    assert flagged > 0  # the median test flags some converged nodes


def test_cleanup_few_good_nodes_skips_the_median_test():
    shape = (4, 5, 6)
    N = int(np.prod(shape))
    U = _field(shape, 5, 0.3)
    status = np.full(N, 1, np.int8)
    status[:27] = 0  # exactly 27 good nodes: no median test (subpb1_solver.py:143)
    out, bad = A.subpb1_cleanup(U, np.zeros_like(U), status, shape, 2.0)
    assert np.array_equal(bad, status != 0)
    assert np.array_equal(out[:27], U[:27]) and np.all(out[27:] == 0.0)


@pytest.fixture(scope="module")
def gd():
    return np.load(os.path.join(HERE, "golden", "golden_driver.npz"))


@pytest.fixture(scope="module")
def ops(gd):
    mesh = SimpleNamespace(grid_shape=tuple(gd["gs_grid_shape"]), elements=gd["gs_elements"], spacing=tuple(gd["gs_spacing"]))
    return FEM.build_global_operators(mesh, "fem", 2)


def test_global_step_matches_host_solve(gd, ops):
    import torch

    para = SimpleNamespace(global_solver="pcg", pcg_tol=1e-8, pcg_max_iter=500)
    system = FEM.build_global_system(ops, 0.5, 2e-3, 0.05, para)
    U1, F1, W, v = gd["gs_U1"], gd["gs_F1"], gd["gs_W"], gd["gs_v"]
    U2_ref, info = FEM.solve_subpb2(system, U1, F1, W, v)
    F2_ref = FEM.nodal_gradient(ops, U2_ref)
    N = ops.n_nodes
    nz, ny, nx = ops.grid_shape
    dev = K.device()
    t = lambda a, s: torch.from_numpy(np.ascontiguousarray(np.reshape(a, s), dtype=np.float64)).to(dev)  # noqa: E731
    U2 = torch.empty((N, 3), dtype=torch.float64, device=dev)
    F2 = torch.empty((N, 9), dtype=torch.float64, device=dev)
    it = torch.zeros(1, dtype=torch.int32, device=dev)
    stop = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = torch.empty(int(K.lib().aldvc_admm_global_workspace_bytes(nx, ny, nz)), dtype=torch.uint8, device=dev)
    L = K.lib()
    ins = [t(U1, (N, 3)), t(F1, (N, 9)), t(W, (N, 9)), t(v, (N, 3))]  # held until the calls are done
    args = lambda st: (K._ptr(ops.mats), K._ptr(system.Ae), nx, ny, nz, K._ptr(ops.elem_active), K._ptr(ops.node_active),  # noqa: E731
                       0.5, 2e-3, 0.05, *(K._ptr(x) for x in ins), 1e-8, 500, K._ptr(U2), K._ptr(F2), K._ptr(it),
                       K._ptr(st), 0, None, K._ptr(ws), int(ws.numel()), K.stream_handle())  # fmt: skip
    K.check(L.aldvc_admm_global_step(*args(stop)), "global step")
    assert int(it.item()) == info["iterations"]
    assert np.max(np.abs(U2.cpu().numpy() - U2_ref)) < 1e-9
    F = F2.cpu().numpy().reshape(N, 3, 3)
    assert np.array_equal(np.isnan(F), np.isnan(F2_ref))
    ok = ~np.isnan(F2_ref)
    assert np.max(np.abs(F[ok] - F2_ref[ok])) < 1e-9
    # a set stop flag leaves the outputs alone
    U2.fill_(7.0)
    stop.fill_(1)
    K.check(L.aldvc_admm_global_step(*args(stop)), "global step (stopped)")
    assert bool(torch.all(U2 == 7.0))


def _dual_ref(U1n, U2n, F2n, U1, U2, F2, W, v, accumulate):
    F2f = np.where(np.isfinite(F2n), F2n, F2)
    rms = lambda a: float(np.sqrt(np.mean(a * a)))  # noqa: E731
    h = [rms(U2n - U2), rms(U1n - U1), rms(U2n - U1n), rms(F2f - F2)]
    W2 = W + (F2f - F2) if accumulate else F2f - F2
    v2 = v + (U2n - U1n) if accumulate else U2n - U1n
    return h, W2, v2, F2f


@pytest.mark.parametrize("accumulate", [1, 0])
def test_dual_update_matches_numpy(accumulate):
    import torch

    rng = np.random.default_rng(6)
    N = 5000
    U1n, U2n, U1, U2, v = (rng.normal(0, 1, (N, 3)) for _ in range(5))
    F2n, F2, W = (rng.normal(0, 0.01, (N, 9)) for _ in range(3))
    F2n[rng.random((N, 9)) < 0.05] = np.nan
    h, W2, v2, F2f = _dual_ref(U1n, U2n, F2n, U1, U2, F2, W, v, accumulate)
    dev = K.device()
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    d = {k: t(a) for k, a in dict(U1n=U1n, U2n=U2n, F2n=F2n, U1=U1, U2=U2, F2=F2, W=W, v=v).items()}
    hist = torch.full((4, 4), np.nan, dtype=torch.float64, device=dev)
    stop = torch.zeros(1, dtype=torch.int32, device=dev)
    steps = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = torch.empty(int(K.lib().aldvc_dual_update_workspace_bytes(N)), dtype=torch.uint8, device=dev)
    K.check(
        K.lib().aldvc_dual_update(N, *(K._ptr(d[k]) for k in ("U1n", "U2n", "F2n", "U1", "U2", "F2", "W", "v")),
                                  accumulate, 1, max(h[0], h[1]) * 1.01, 2, K._ptr(hist), K._ptr(stop), K._ptr(steps),
                                  K._ptr(ws), int(ws.numel()), K.stream_handle()),
        "dual update",
    )  # fmt: skip
    assert np.allclose(hist.cpu().numpy()[2], h, rtol=1e-12, atol=0)
    assert int(stop.item()) == 1 and int(steps.item()) == 2
    assert np.allclose(d["W"].cpu().numpy(), W2, rtol=0, atol=1e-15)
    assert np.allclose(d["v"].cpu().numpy(), v2, rtol=0, atol=1e-15)
    assert np.array_equal(d["F2"].cpu().numpy(), F2f)
    assert np.array_equal(d["U1"].cpu().numpy(), U1n) and np.array_equal(d["U2"].cpu().numpy(), U2n)


# ---------------------------------------------------------------------------- the whole loop
def _para(**kw):
    base = dict(subset_stride=1, interp_method="cubic", icgn_tol=1e-2, icgn_dp_tol=1e-3, icgn_max_iter=100, icgn_patience=5,
                icgn_predictive_stop=True, icgn_noise_hessian=False, min_valid_ratio=0.5, hessian_cond_max=1e12,
                local_outlier_threshold=2.0, subset_split=False, gradient_mode="stored", winsize=(12, 12, 12),
                mu=1e-3, admm_max_iter=4, admm_tol=1e-6, dual_update="accumulate", global_solver="pcg",
                pcg_tol=1e-8, pcg_max_iter=500, alpha=0.0)  # fmt: skip
    base.update(kw)
    return SimpleNamespace(**base)


def _host_loop(ctx, ref, g, mesh, system, para, U1, F1):
    """core/pipeline.py:290-358 on the host pieces (GPU 3-DOF solve, host-controlled PCG)."""
    ops = system.ops
    N = ctx.n_nodes
    W = np.zeros((N, 3, 3))
    v = np.zeros((N, 3))
    U2, _ = FEM.solve_subpb2(system, U1, F1, W, v)
    F2 = FEM.nodal_gradient(ops, U2)
    F2 = np.where(np.isfinite(F2), F2, F1)
    W, v = F2 - F1, U2 - U1
    rms = lambda a: float(np.sqrt(np.mean(a * a)))  # noqa: E731
    res_u, res_f, upd_g, upd_l = [rms(U2 - U1)], [rms(F2 - F1)], [], []
    n_steps, status, bad = 1, None, None
    for step in range(2, para.admm_max_iter + 1):
        n_steps = step
        U1_prev = U1
        U1, info, bad = LOC.subpb1_solver(ctx, ref, g, U2, F2, v, system.mu, para, mesh)
        status = info.status
        F1 = F2
        U2n, _ = FEM.solve_subpb2(system, U1, F1, W, v)
        F2n = FEM.nodal_gradient(ops, U2n)
        F2n = np.where(np.isfinite(F2n), F2n, F1)
        upd_g.append(rms(U2n - U2))
        upd_l.append(rms(U1 - U1_prev))
        W = W + (F2n - F1)
        v = v + (U2n - U1)
        U2, F2 = U2n, F2n
        res_u.append(rms(U2 - U1))
        res_f.append(rms(F2 - F1))
        if upd_g[-1] < para.admm_tol or upd_l[-1] < para.admm_tol:
            break
    return dict(U=U2, F=F2, n=n_steps, upd_g=upd_g, upd_l=upd_l, res_u=res_u, res_f=res_f, status=status, bad=bad)


@pytest.fixture(scope="module", params=["drv", "drvm", "drvm-fd"])
def frame(gd, request):
    """A golden driver pair (40x44x48): local 12-DOF solve, then the global system. ``drv``: no mask;
    ``drvm``: a cylinder mask with voids and a crack, split subsets and a cut mesh (the elements that
    bridge the mask dropped, cut edges excluded from the median test and the fills); ``drvm-fd``:
    the same frame with the finite-difference global step (valid nodes, one-sided differences at
    the cut edges)."""
    p, _, method = request.param.partition("-")
    method = method or "fem"
    coords = gd[f"{p}_coords"]
    grid = tuple(int(v) for v in gd[f"{p}_grid_shape"])
    masked = p == "drvm"
    para = _para(subset_split=masked)
    if masked:
        from pyaldvc_amd.mesh_cut import apply_mask_to_mesh

        mm = apply_mask_to_mesh(coords, grid, gd["drvm_mask"].astype(bool), para.winsize, para.min_valid_ratio,
                                cut_bridging=True)  # fmt: skip
        assert np.array_equal(mm.node_valid, gd["drvm_node_valid_in"])
        assert np.array_equal(mm.edge_ok, gd["drvm_edge_ok"])
        mesh = LOC.NodeMesh(coords, grid, gd["drvm_node_valid_in"].copy(), gd["drvm_edge_ok"])
        ref = SimpleNamespace(f=gd["drv_f"], mask=gd["drvm_mask"].astype(bool), gradient_mode="stored")
        elements = mm.elements
    else:
        mesh = LOC.NodeMesh(coords, grid, gd[f"{p}_node_valid_in"].copy())
        ref = SimpleNamespace(f=gd["drv_f"], mask=None, gradient_mode="stored")
        elements = FEM.hex8_elements(grid)
    ctx = LOC.precompute_local_context(mesh, ref, para)
    mesh.node_valid = ctx.valid.copy()
    U1, F1, info, bad = LOC.local_icgn(ctx, ref, gd["drv_g"], gd[f"{p}_U0"], para, mesh)
    spacing = tuple(float(np.diff(np.unique(coords[:, k]))[0]) for k in range(3))
    el = np.array(elements, copy=True)
    el[(el < 0).any(1) | ~mesh.node_valid[np.maximum(el, 0)].all(1)] = -1
    gm = SimpleNamespace(grid_shape=grid, elements=el, spacing=spacing, node_valid=mesh.node_valid, edge_ok=mesh.edge_ok)
    ops = FEM.build_global_operators(gm, method, 2)
    system = FEM.build_global_system(ops, 1e-3 * np.mean(spacing) ** 2, para.mu, 0.0, para)
    return SimpleNamespace(mesh=mesh, ref=ref, para=para, ctx=ctx, U1=U1, F1=F1, info=info, system=system, g=gd["drv_g"],
                           name=request.param)  # fmt: skip


def _stopping_tol(fr):
    """An admm_tol at which the reference loop stops at step 2 (just above the smaller update)."""
    host = _host_loop(fr.ctx, fr.ref, fr.g, fr.mesh, fr.system, _para(admm_tol=1e-12, subset_split=fr.para.subset_split),
                      fr.U1, fr.F1)  # fmt: skip
    return 1.01 * min(host["upd_g"][0], host["upd_l"][0])


@pytest.mark.parametrize("stops", [False, True])
def test_admm_loop_matches_the_reference_loop(frame, stops):
    fr = frame
    tol = _stopping_tol(fr) if stops else 1e-12
    para = _para(admm_tol=tol, subset_split=fr.para.subset_split)
    host = _host_loop(fr.ctx, fr.ref, fr.g, fr.mesh, fr.system, para, fr.U1, fr.F1)
    loop = A.AdmmLoop(fr.ctx, fr.g, fr.mesh, fr.system, para)
    loop.load(fr.U1, fr.F1)
    loop.run(graph=False)
    U, F, info, status, zncc, bad = loop.results(fr.info)
    assert info.n_steps == host["n"]
    assert np.array_equal(status, host["status"]) and np.array_equal(bad, host["bad"])
    assert np.max(np.abs(U - host["U"])) < 1e-6
    assert np.max(np.abs(F - host["F"])) < 1e-8
    assert np.allclose(info.update_global, host["upd_g"], rtol=1e-6, atol=1e-12)
    assert np.allclose(info.primal_residual_u, host["res_u"], rtol=1e-6, atol=1e-12)
    assert info.n_steps == (2 if stops else para.admm_max_iter)


@pytest.mark.parametrize("stops", [False, True])
def test_admm_loop_graph_replays_equal_eager(frame, stops):
    import torch

    fr = frame
    para = _para(admm_tol=_stopping_tol(fr) if stops else 1e-12, subset_split=fr.para.subset_split)
    loop = A.AdmmLoop(fr.ctx, fr.g, fr.mesh, fr.system, para)
    loop.load(fr.U1, fr.F1)
    loop.run(graph=False)
    eager = loop.results(fr.info)
    for rep in range(3):
        loop.load(fr.U1, fr.F1)
        loop.run(graph=True)
        got = loop.results(fr.info)
        assert got[2].n_steps == eager[2].n_steps, rep
        assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1]), rep
        assert np.array_equal(got[3], eager[3]) and np.array_equal(got[5], eager[5]), rep
        assert got[2].pcg_iterations == eager[2].pcg_iterations, rep
        assert not loop.fell_back and loop.graph_budgets == loop.budgets, rep
        _ = torch.randn(1000, device=K.device()).sum().item()  # other work between replays
    assert loop.budgets[0] < loop.pcg_max  # the graph was recorded with a learned, smaller budget


def test_admm_loop_graph_budget_too_small_falls_back(frame):
    """A graph recorded with CG budgets below what the solves need raises the device flag; the run
    is redone eagerly with the full limits and equals the eager result, and the next call records
    the graph again with the budgets of that run."""
    fr = frame
    para = _para(admm_tol=1e-12, subset_split=fr.para.subset_split)
    loop = A.AdmmLoop(fr.ctx, fr.g, fr.mesh, fr.system, para)
    loop.load(fr.U1, fr.F1)
    loop.run(graph=False)
    eager = loop.results(fr.info)
    learned = loop.budgets
    loop.budgets = (2, 2)  # far too few iterations for the global and fill solves
    loop.load(fr.U1, fr.F1)
    loop.run(graph=True)
    assert loop.fell_back and loop.graph is None and loop.budgets == learned
    got = loop.results(fr.info)
    assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1])
    assert got[2].pcg_iterations == eager[2].pcg_iterations
    loop.load(fr.U1, fr.F1)
    loop.run(graph=True)
    assert not loop.fell_back and loop.graph_budgets == learned
    again = loop.results(fr.info)
    assert np.array_equal(again[0], eager[0]) and np.array_equal(again[1], eager[1])
