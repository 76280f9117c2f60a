"""The ADMM iterations of a frame on the GPU, with no host round trip inside an iteration, so the
whole loop can be recorded into one HIP graph and replayed (core/pipeline.py:286-366).

One iteration (steps 2..``admm_max_iter`` of the reference loop) is:

1. subproblem 1: translation IC-GN of every node with ``F`` held at ``F2`` and the proximal pull
   towards ``U2 + v`` (``aldvc_icgn3_tuned``, gated by the stop flag);
2. its cleanup: bad = not converged, or flagged by the universal median test (unknown nodes
   spring-filled first); bad nodes keep ``U2`` (``aldvc_subpb1_cleanup``,
   solver/subpb1_solver.py:140-164);
3. subproblem 2: the right-hand side, the float64 PCG solve and the nodal gradient with the FEM or
   the FD operators (``aldvc_admm_global_step`` / ``aldvc_admm_global_step_fd``,
   solver/subpb2_solver.py:108-155, global_operators.py:147-170);
4. the dual update, the RMS updates and the stopping rule (``aldvc_dual_update``).

The first pass after the local 12-DOF solve (the global step from ``W = 0``, ``v = 0``) is the same
sequence without 1-2. Every launch reads a device flag ``stop`` that the dual update sets when the
reference loop would break (RMS update of ``U`` global or local below ``admm_tol``), so the
iterations recorded after that point change nothing and the results equal the reference loop's.
The whole sequence is one call of the C ABI (``aldvc_admm_loop_enqueue`` with the buffers in an
``AldvcAdmmLoop`` descriptor), and the library records it into a graph itself
(``aldvc_admm_loop_graph``), so callers other than Python get the same replayable loop.

The spring fill and the median test are also exposed on their own (``spring_fill``,
``subpb1_cleanup``) with the reference's host signatures.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import autotune as AT
from . import kernels as K
from ._lib import AdmmLoopDesc, check, lib
from .local import INTERP_MODE_BY_NAME, LocalSolveInfo

FILL_TOL = 1e-12  # relative residual of the spring fill's CG (the reference solves it directly up to 60k unknowns)
FILL_MAX_ITER = 2000  # the reference's CG limit for the fill (inpaint.py:118)
MEDIAN_EPS = 0.1  # noise level of the universal median test (outlier_detection.py:138)
DIRECT_GRAPH_ITER = 2000  # PCG iterations recorded for the "direct" global solver (driven to round-off)
SYNC_EVERY = 8  # eager runs: the host reads a CG solve's stop flag every 8 iterations
# A recorded graph enqueues a fixed number of CG slots per solve; idle slots after convergence still
# cost a dispatch each, so a graph is recorded with a budget learned from the eager run (twice the
# largest iteration count seen, plus a margin, at most the solver's own limit). A solve that uses
# its whole budget raises a device flag, and the loop then reruns eagerly with the full limits.
BUDGET_FACTOR, BUDGET_MARGIN = 2, 16
STATUS_CONVERGED, STATUS_INVALID_SUBSET, STATUS_SKIPPED = 0, 3, 6


def _dev(a, dtype=torch.float64, shape=None):
    """A device tensor of ``a`` (host array or tensor), optionally reshaped."""
    if isinstance(a, torch.Tensor):
        t = a.to(device=K.device(), dtype=dtype)
    else:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(device=K.device(), dtype=dtype)
    return (t if shape is None else t.reshape(shape)).contiguous()


def _addr(t):
    """Device address of a tensor for a descriptor field (None for a missing buffer)."""
    return None if t is None else int(t.data_ptr())


class _Graph:
    """An executable HIP graph made by ``aldvc_admm_loop_graph``; freed with the object."""

    def __init__(self, handle: int):
        self.handle = handle

    def launch(self):
        check(lib().aldvc_graph_launch(ctypes.c_void_p(self.handle), K.stream_handle()), "aldvc_graph_launch")

    def __del__(self):
        if self.handle:
            lib().aldvc_graph_destroy(ctypes.c_void_p(self.handle))
            self.handle = None


def _ws(nbytes: int) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=K.device())


def _edges(edge_ok, grid_shape):
    if edge_ok is None:
        return None
    e = np.asarray(edge_ok, dtype=np.uint8).reshape(int(np.prod(grid_shape)), 3)
    return _dev(e, torch.uint8)


# ---------------------------------------------------------------------------- stand-alone pieces
def spring_fill(values, grid_shape, edge_ok=None, tol: float = FILL_TOL, max_iter: int = FILL_MAX_ITER) -> np.ndarray:
    """``values`` (N, 3) with NaN holes on ``grid_shape`` (nz, ny, nx), each column filled by the
    spring model (utils/inpaint.py:35-121 with method "spring"); host arrays in and out, or device
    tensors in and out."""
    nz, ny, nx = (int(v) for v in grid_shape)
    vals = _dev(values, shape=(nz * ny * nx, 3))
    ws = _ws(lib().aldvc_spring_fill_workspace_bytes(nx, ny, nz))
    edges = _edges(edge_ok, grid_shape)
    it = ctypes.c_int(0)
    check(
        lib().aldvc_spring_fill(nx, ny, nz, K._ptr(edges), K._ptr(vals), float(tol), int(max_iter),
                                ctypes.byref(it), K._ptr(ws), int(ws.numel()), K.stream_handle()),
        "aldvc_spring_fill",
    )  # fmt: skip
    return vals if isinstance(values, torch.Tensor) else vals.cpu().numpy()


def subpb1_cleanup(U, U_hat, status, grid_shape, threshold: float, edge_ok=None, eps: float = MEDIAN_EPS):
    """``(U_out, bad)``: bad = not converged, or (threshold > 0 and more than 27 converged) flagged
    by the universal median test; bad nodes take ``U_hat`` (solver/subpb1_solver.py:140-164). Host
    arrays out, or device tensors (``bad`` as bool) when ``U`` is a device tensor."""
    nz, ny, nx = (int(v) for v in grid_shape)
    N = nz * ny * nx
    Ud, Uh = _dev(U, shape=(N, 3)), _dev(U_hat, shape=(N, 3))
    st = _dev(status, torch.int8, (N,))
    out = torch.empty_like(Ud)
    bad = torch.empty(N, dtype=torch.uint8, device=Ud.device)
    ws = _ws(lib().aldvc_subpb1_cleanup_workspace_bytes(nx, ny, nz))
    edges = _edges(edge_ok, grid_shape)
    check(
        lib().aldvc_subpb1_cleanup(nx, ny, nz, K._ptr(edges), K._ptr(st), K._ptr(Ud), K._ptr(Uh),
                                   float(threshold), float(eps), FILL_TOL, FILL_MAX_ITER, K._ptr(out), K._ptr(bad),
                                   None, SYNC_EVERY, None, None, K._ptr(ws), int(ws.numel()), K.stream_handle()),
        "aldvc_subpb1_cleanup",
    )  # fmt: skip
    if isinstance(U, torch.Tensor):
        return out, bad.bool()
    return out.cpu().numpy(), bad.cpu().numpy().astype(bool)


# ---------------------------------------------------------------------------- the loop
@dataclass
class ADMMInfo:
    """The reference's ``ADMMInfo`` (core/pipeline.py:359-370)."""

    beta: float
    mu: float
    n_steps: int
    update_global: tuple
    update_local: tuple
    primal_residual_u: tuple
    primal_residual_f: tuple
    local_info: tuple
    subpb2_time: float
    pcg_iterations: tuple = field(default=())
    beta_sweep: object = None


class AdmmLoop:
    """Device state of the ADMM iterations of one frame.

    ``ctx`` (LocalContext), ``g`` (deformed volume), ``mesh`` (NodeMesh), ``system``
    (fem.GlobalSystem) and ``para`` as for the reference loop; ``U_local`` (N, 3) and ``F_local``
    (N, 3, 3) the output of the local 12-DOF step. ``run(graph=...)`` executes the first pass and
    steps 2..``para.admm_max_iter``; ``graph=True`` records them once into a HIP graph (made by
    the library, ``aldvc_admm_loop_graph``) and replays it."""

    def __init__(self, ctx, g, mesh, system, para):
        self.ctx, self.mesh, self.system, self.para = ctx, mesh, system, para
        ops = system.ops
        self.grid = tuple(int(v) for v in ops.grid_shape)
        nz, ny, nx = self.grid
        N = nz * ny * nx
        if N != ctx.n_nodes:
            raise ValueError("the operators' node grid and the local context disagree")
        from .local import _tiled

        if _tiled(ctx, para):
            raise ValueError("the ADMM loop on the device runs on one volume box; a tiled reference (tile_local, or a "
                             "volume beyond 2^31 voxels) takes local.subpb1_solver and fem.solve_subpb2 per step")  # fmt: skip
        self.N = N
        self.max_steps = int(para.admm_max_iter)
        dev = K.device()
        f64 = dict(dtype=torch.float64, device=dev)
        z3 = lambda: torch.zeros((N, 3), **f64)  # noqa: E731
        z9 = lambda: torch.zeros((N, 9), **f64)  # noqa: E731
        self.U1, self.U2, self.v = z3(), z3(), z3()
        self.F2, self.W = z9(), z9()
        self.U1raw, self.U1n, self.U2n = z3(), z3(), z3()
        self.F2n = z9()
        S = self.max_steps + 1
        self.status = torch.zeros((S, N), dtype=torch.int8, device=dev)
        self.zncc = torch.zeros((S, N), **f64)
        self.n_iter = torch.zeros((S, N), dtype=torch.int32, device=dev)
        self.bad = torch.zeros((S, N), dtype=torch.uint8, device=dev)
        self.stop = torch.zeros(1, dtype=torch.int32, device=dev)
        self.steps = torch.zeros(1, dtype=torch.int32, device=dev)
        self.hist = torch.full((S, 4), float("nan"), **f64)
        self.pcg_iters = torch.zeros(S, dtype=torch.int32, device=dev)
        self.fill_iters = torch.zeros(S, dtype=torch.int32, device=dev)  # CG iterations of each cleanup's fill
        self.capped = torch.zeros(1, dtype=torch.int32, device=dev)  # a budgeted solve used its whole budget
        L = lib()
        self.ws_global = _ws(L.aldvc_admm_global_workspace_bytes(nx, ny, nz))
        self.ws_clean = _ws(L.aldvc_subpb1_cleanup_workspace_bytes(nx, ny, nz))
        self.ws_dual = _ws(L.aldvc_dual_update_workspace_bytes(N))
        self.ws_icgn3 = _ws(L.aldvc_icgn3_workspace_bytes(N))
        self.edge_ok = _edges(mesh.edges_ok_grid(), self.grid)
        # 3-DOF inputs (solver/subpb1_solver.py:77-133)
        vols = ctx._volumes
        self.vols = vols
        self.gd = K._Args.vol(g)
        self.mode = INTERP_MODE_BY_NAME[para.interp_method]
        self.pattern, self.gain = ctx.noise_args(para)
        sp = ctx.split_args()
        self.split_args = sp
        self.split = K._split(K._Args(self.gd), N, sp.get("split_index"), sp.get("split_keep"))
        # the first 3-DOF pass's reference moments by the separable pass, where it applies
        d = ctx.device
        extra = K.row_moments_bytes(d["coords"], d["valid"], *ctx.half, ctx.stride, tuple(vols["f"].shape),
                                    vols["stored"], getattr(ctx, "separable", None))  # fmt: skip
        self.row_flag = K.TUNE_ROW_MOMENTS if extra else 0
        if extra:
            self.ws_icgn3 = _ws(L.aldvc_icgn3_workspace_bytes(N) + extra)
        self.tuning3 = None  # launch shape of the 3-DOF solves, chosen by the autotuner in load()
        if system.solver == "direct":
            self.pcg_tol, self.pcg_max = 1e-13, int(min(max(system.pcg_max_iter, 10 * ops.n_active), DIRECT_GRAPH_ITER))
        else:
            self.pcg_tol, self.pcg_max = float(system.pcg_tol), int(system.pcg_max_iter)
        self.graph = None
        self.graph_budgets = None  # (PCG, fill) iteration budgets the recorded graph was built with
        self.budgets = None  # learned from the last eager run
        self.fell_back = False  # the last graph run was redone eagerly (a budget was too small)
        self._loaded = None
        self._sync = 0  # sync_every of the launches being enqueued
        self._limits = None  # (PCG, fill) iteration limits of the launches being enqueued
        self.time_s = 0.0

    # -- launches (all on the current stream, no host synchronisation) -------------------------
    def _desc(self) -> AdmmLoopDesc:
        """The loop's buffers and settings as the C ABI's ``AldvcAdmmLoop``, with the host polling
        interval and the CG limits of the run being enqueued (``_sync``, ``_limits``)."""
        ctx, p, d, vols, s = self.ctx, self.para, self.ctx.device, self.vols, self.system
        ops = s.ops
        nz, ny, nx = self.grid
        si, sk, kb = self.split
        t = self.tuning3 if self.tuning3 is not None else AT.DEFAULT
        stored = vols["stored"]
        D = AdmmLoopDesc()
        D.struct_bytes = ctypes.sizeof(AdmmLoopDesc)
        # local 3-DOF step (solver/subpb1_solver.py:77-133)
        D.coords, D.n_nodes = _addr(d["coords"]), self.N
        D.hx, D.hy, D.hz = ctx.half
        D.stride = int(ctx.stride)
        D.f, D.mask, D.g = _addr(vols["f"]), _addr(vols["mask"]), _addr(self.gd)
        D.gx, D.gy, D.gz = (_addr(vols[k]) if stored else None for k in ("gx", "gy", "gz"))
        D.vol_nz, D.vol_ny, D.vol_nx = (int(v) for v in vols["f"].shape)
        D.mode = self.mode
        D.H_all, D.meanf, D.bottomf, D.valid = (_addr(d[k]) for k in ("H", "meanf", "bottomf", "valid"))
        D.icgn_tol, D.icgn_dp_tol = float(p.icgn_tol), float(p.icgn_dp_tol)
        D.icgn_max_iter, D.icgn_patience = int(p.icgn_max_iter), int(p.icgn_patience)
        D.noise_gain = float(self.gain) if self.gain > 0 else 0.0
        D.predictive = int(bool(p.icgn_predictive_stop))
        D.split_index, D.split_keep, D.keep_bytes = _addr(si), _addr(sk), int(kb)
        D.variant, D.lds_budget, D.flags = int(t.variant), int(t.lds_budget), int(t.flags) | self.row_flag
        # cleanup (solver/subpb1_solver.py:140-164)
        D.nx, D.ny, D.nz = nx, ny, nz
        D.edge_ok = _addr(self.edge_ok)
        D.outlier_threshold, D.median_eps, D.fill_tol = float(p.local_outlier_threshold), MEDIAN_EPS, FILL_TOL
        D.fill_budget, D.fill_limit = int(self._limits[1]), FILL_MAX_ITER
        # global step (solver/subpb2_solver.py:108-155)
        if ops.method == "fd":
            D.method, D.codes = 1, _addr(ops.codes)
            D.spacing_x, D.spacing_y, D.spacing_z = ops.spacing
        else:
            D.method, D.mats, D.Ae = 0, _addr(ops.mats), _addr(s.Ae)
            D.elem_active, D.node_active = _addr(ops.elem_active), _addr(ops.node_active)
        D.beta, D.mu, D.alpha, D.pcg_tol = s.beta, s.mu, s.alpha, self.pcg_tol
        D.pcg_budget, D.pcg_limit = int(self._limits[0]), self.pcg_max
        # dual update and stopping rule (core/pipeline.py:300-358)
        D.accumulate = int(p.dual_update == "accumulate")
        D.admm_tol, D.max_steps = float(p.admm_tol), self.max_steps
        for k in ("U1", "U2", "v", "U1raw", "U1n", "U2n", "F2", "W", "F2n", "status", "zncc", "n_iter", "bad", "stop",
                  "steps", "capped"):  # fmt: skip
            setattr(D, k, _addr(getattr(self, k)))
        D.history, D.pcg_iterations, D.fill_iterations = _addr(self.hist), _addr(self.pcg_iters), _addr(self.fill_iters)
        for k, ws in (("icgn3", self.ws_icgn3), ("cleanup", self.ws_clean), ("global", self.ws_global),
                      ("dual", self.ws_dual)):  # fmt: skip
            setattr(D, f"ws_{k}", _addr(ws))
            setattr(D, f"ws_{k}_bytes", int(ws.numel()))
        return D

    def _enqueue(self):
        """The first pass and every later step, in stream order (one call of the C ABI)."""
        D = self._desc()
        check(lib().aldvc_admm_loop_enqueue(ctypes.byref(D), self._sync, K.stream_handle()), "aldvc_admm_loop_enqueue")

    # -- public ----------------------------------------------------------------------------------
    def load(self, U_local, F_local, status_local=None, zncc_local=None):
        """The local 12-DOF result as the loop's start: U1 = U_local, F2 = F_local (the F1 of the
        first pass), W = v = 0."""
        N = self.N
        self.U1.copy_(_dev(U_local, shape=(N, 3)))
        self.U2.copy_(self.U1)
        self.F2.copy_(_dev(F_local, shape=(N, 9)))
        self.W.zero_()
        self.v.zero_()
        self.stop.zero_()
        self.steps.zero_()
        self.hist.fill_(float("nan"))
        self.pcg_iters.zero_()
        self.fill_iters.zero_()
        self.capped.zero_()
        self.status.zero_()
        self.zncc.zero_()
        self.n_iter.zero_()
        self.bad.zero_()
        self._loaded = (U_local, F_local)
        if self.tuning3 is None:
            self.tuning3 = self._choose_tuning3()

    def _choose_tuning3(self):
        """The autotuner's launch shape for this loop's 3-DOF solves, timed (eagerly, before any
        graph capture) on a node sample of the loaded state."""
        ctx, p, d, vols = self.ctx, self.para, self.ctx.device, self.vols
        stored = vols["stored"]
        gx, gy, gz = (vols[k] if stored else None for k in ("gx", "gy", "gz"))
        args = (d["coords"], self.U2, self.F2, self.v, *ctx.half, vols["f"], gx, gy, gz, vols["mask"], self.gd,
                self.mode, d["H"], d["meanf"], d["bottomf"], d["valid"], float(self.system.mu), float(p.icgn_tol),
                float(p.icgn_dp_tol), int(p.icgn_max_iter), int(p.icgn_patience))  # fmt: skip
        kw = dict(stride=int(ctx.stride), noise_gain=float(self.gain) if self.gain > 0 else 0.0,
                  predictive=bool(p.icgn_predictive_stop), split_index=self.split_args.get("split_index"),
                  split_keep=self.split_args.get("split_keep"))  # fmt: skip
        return AT.choose_icgn3(args, kw)

    def run(self, graph: bool = False):
        """Executes the loop on the current state (call ``load`` first). Eagerly (the default) the
        CG solves are polled by the host and enqueue nothing after convergence; the run also sets
        the iteration budgets of later graph recordings. With ``graph`` the sequence is recorded
        once (the first call, or again after the budgets changed) and replayed; if a budgeted solve
        used its whole budget, the loop reloads and reruns eagerly with the full limits, so the
        result is always that of the unbudgeted loop."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.fell_back = False
        if not graph:
            self._run_eager()
        else:
            if self.graph is None or self.graph_budgets != self._graph_budgets():
                self.capture()
            self.graph.launch()
            if int(self.capped.item()):
                self.fell_back = True
                self.load(*self._loaded)
                self._run_eager()
                self.graph = None  # recorded again with the larger budgets on the next call
        torch.cuda.synchronize()
        self.time_s = time.perf_counter() - t0

    def _run_eager(self):
        self._sync, self._limits = SYNC_EVERY, (self.pcg_max, FILL_MAX_ITER)
        self._enqueue()
        torch.cuda.synchronize()
        n = int(self.steps.item())
        pcg = int(self.pcg_iters[1 : n + 1].max().item()) if n >= 1 else 0
        fill = int(self.fill_iters[2 : n + 1].max().item()) if n >= 2 else 0
        budget = lambda used, limit: int(min(limit, BUDGET_FACTOR * used + BUDGET_MARGIN))  # noqa: E731
        self.budgets = (budget(pcg, self.pcg_max), budget(fill, FILL_MAX_ITER))

    def _graph_budgets(self):
        return self.budgets if self.budgets is not None else (self.pcg_max, FILL_MAX_ITER)

    def capture(self):
        """Records the loop into a HIP graph through the C ABI (``aldvc_admm_loop_graph``; the
        state buffers are fixed, so a replay runs the loop on whatever ``load`` put there), with
        the iteration budgets of the last eager run (the full limits before any)."""
        self._sync, self._limits = 0, self._graph_budgets()
        D = self._desc()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        h = ctypes.c_void_p()
        check(lib().aldvc_admm_loop_graph(ctypes.byref(D), K.stream_handle(s), ctypes.byref(h)), "aldvc_admm_loop_graph")
        torch.cuda.current_stream().wait_stream(s)
        self.graph, self.graph_budgets = _Graph(h.value), self._limits

    def results(self, info_local: LocalSolveInfo | None = None):
        """``(U_final, F_final, ADMMInfo, status, zncc, bad_last)`` on the host, as the reference
        loop leaves them (core/pipeline.py:359-380)."""
        n = int(self.steps.item())
        hist = self.hist.cpu().numpy()
        iters = self.pcg_iters.cpu().numpy()
        infos = [info_local] if info_local is not None else []
        st = self.status.cpu().numpy()
        zn = self.zncc.cpu().numpy()
        ni = self.n_iter.cpu().numpy()
        bad = self.bad.cpu().numpy().astype(bool)
        per_step = self.time_s / max(1, n)
        for k in range(2, n + 1):
            s = st[k]
            n_bad = int(np.sum(bad[k] & (s != STATUS_INVALID_SUBSET) & (s != STATUS_SKIPPED)))
            infos.append(LocalSolveInfo(n_iter=ni[k], status=s, zncc=zn[k], solve_time=per_step, n_bad=n_bad))
        info = ADMMInfo(
            beta=self.system.beta, mu=self.system.mu, n_steps=n,
            update_global=tuple(float(hist[k, 0]) for k in range(2, n + 1)),
            update_local=tuple(float(hist[k, 1]) for k in range(2, n + 1)),
            primal_residual_u=tuple(float(hist[k, 2]) for k in range(1, n + 1)),
            primal_residual_f=tuple(float(hist[k, 3]) for k in range(1, n + 1)),
            local_info=tuple(infos), subpb2_time=float("nan"), pcg_iterations=tuple(int(i) for i in iters[1 : n + 1]),
        )  # fmt: skip
        U = self.U2.cpu().numpy()
        F = self.F2.cpu().numpy().reshape(self.N, 3, 3)
        if n >= 2:
            status, zncc, bad_last = st[n], zn[n], bad[n]
        elif info_local is not None:
            status, zncc, bad_last = info_local.status, info_local.zncc, None
        else:
            status = zncc = bad_last = None
        return U, F, info, status, zncc, bad_last
