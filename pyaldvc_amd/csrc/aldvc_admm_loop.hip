// The ADMM iterations of a frame (core/pipeline.py:286-366) as one call of the C ABI, and that
// call recorded into a HIP graph, for callers that have no graph API of their own.
//
// The loop is the fixed sequence of the four per-iteration entries (3-DOF local step, its cleanup,
// the global step, the dual update with the stopping rule), each reading the device stop flag, so
// enqueueing max_steps iterations up front gives the reference loop's result: the iterations after
// the flag is set leave every buffer alone. With sync_every = 0 no call reads anything back, which
// is what makes the sequence capturable.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/aldvc_hip.h"
#include "aldvc_abi_util.h"

using namespace aldvc;

namespace {

constexpr int REUSE_MOMENTS = 16;  // aldvc_icgn3_tuned: the reference moments of the first 3-DOF pass

__global__ void k_copy_f64(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int check_loop(const AldvcAdmmLoop* L) {
    if (!L) return fail(-2, "null loop descriptor");
    if (L->struct_bytes != sizeof(AldvcAdmmLoop))
        return fail(-2, "AldvcAdmmLoop layout differs from the library's (struct_bytes)");
    if (L->nx < 1 || L->ny < 1 || L->nz < 1 || (int64_t)L->nx * L->ny * L->nz != L->n_nodes)
        return fail(-2, "n_nodes must equal nx * ny * nz");
    if (L->max_steps < 1) return fail(-2, "max_steps must be at least 1");
    if (L->method != 0 && L->method != 1) return fail(-2, "method must be 0 (FEM) or 1 (FD)");
    if (L->pcg_budget < 1 || L->pcg_budget > L->pcg_limit || L->fill_budget < 1 || L->fill_budget > L->fill_limit)
        return fail(-2, "iteration budgets must lie in [1, limit]");
    if (!L->U1 || !L->U2 || !L->v || !L->U1raw || !L->U1n || !L->U2n || !L->F2 || !L->W || !L->F2n)
        return fail(-2, "null state buffer");
    if (!L->status || !L->zncc || !L->n_iter || !L->bad || !L->history || !L->pcg_iterations || !L->fill_iterations ||
        !L->stop || !L->steps || !L->capped)
        return fail(-2, "null record buffer");
    return 0;
}

int global_step(const AldvcAdmmLoop* L, int step, int sync, hipStream_t st) {
    int* capped = L->pcg_budget < L->pcg_limit ? L->capped : nullptr;
    if (L->method == 1)
        return aldvc_admm_global_step_fd(L->codes, L->nx, L->ny, L->nz, L->spacing_x, L->spacing_y, L->spacing_z,
                                         L->beta, L->mu, L->alpha, L->U1n, L->F2, L->W, L->v, L->pcg_tol,
                                         L->pcg_budget, L->U2n, L->F2n, L->pcg_iterations + step, L->stop, sync, capped,
                                         L->ws_global, L->ws_global_bytes, st);
    return aldvc_admm_global_step(L->mats, L->Ae, L->nx, L->ny, L->nz, L->elem_active, L->node_active, L->beta, L->mu,
                                  L->alpha, L->U1n, L->F2, L->W, L->v, L->pcg_tol, L->pcg_budget, L->U2n, L->F2n,
                                  L->pcg_iterations + step, L->stop, sync, capped, L->ws_global, L->ws_global_bytes, st);
}

int dual_update(const AldvcAdmmLoop* L, int step, bool check_stop, hipStream_t st) {
    return aldvc_dual_update(L->n_nodes, L->U1n, L->U2n, L->F2n, L->U1, L->U2, L->F2, L->W, L->v, L->accumulate,
                             check_stop ? 1 : 0, L->admm_tol, step, L->history, L->stop, L->steps, L->ws_dual,
                             L->ws_dual_bytes, st);
}

int local_step(const AldvcAdmmLoop* L, int step, int sync, hipStream_t st) {
    const int64_t N = L->n_nodes;
    const int flags = L->flags | (step > 2 ? REUSE_MOMENTS : 0);
    if (int rc = aldvc_icgn3_tuned(L->coords, N, L->U2, L->U1raw, L->F2, L->v, L->hx, L->hy, L->hz, L->stride, L->f,
                                   L->gx, L->gy, L->gz, L->mask, L->g, L->vol_nz, L->vol_ny, L->vol_nx, L->mode,
                                   L->H_all, L->meanf, L->bottomf, L->valid, L->mu, L->icgn_tol, L->icgn_dp_tol,
                                   L->icgn_max_iter, L->icgn_patience, L->noise_gain, L->predictive, L->split_index,
                                   L->split_keep, L->keep_bytes, L->n_iter + step * N, L->status + step * N,
                                   L->zncc + step * N, L->stop, L->variant, L->lds_budget, flags, L->ws_icgn3,
                                   L->ws_icgn3_bytes, st))
        return rc;
    int* capped = L->fill_budget < L->fill_limit ? L->capped : nullptr;
    return aldvc_subpb1_cleanup(L->nx, L->ny, L->nz, L->edge_ok, L->status + step * N, L->U1raw, L->U2,
                                L->outlier_threshold, L->median_eps, L->fill_tol, L->fill_budget, L->U1n,
                                L->bad + step * N, L->stop, sync, L->fill_iterations + step, capped, L->ws_cleanup,
                                L->ws_cleanup_bytes, st);
}

int enqueue(const AldvcAdmmLoop* L, int sync, hipStream_t st) {
    const int64_t n3 = 3 * L->n_nodes;
    hipLaunchKernelGGL(k_copy_f64, dim3((unsigned)((n3 + 255) / 256 < 4096 ? (n3 + 255) / 256 : 4096)), dim3(256), 0,
                       st, (const double*)L->U1, L->U1n, n3);
    if (int rc = launch_status()) return rc;
    if (int rc = global_step(L, 1, sync, st)) return rc;
    if (int rc = dual_update(L, 1, false, st)) return rc;
    for (int step = 2; step <= L->max_steps; ++step) {
        if (int rc = local_step(L, step, sync, st)) return rc;
        if (int rc = global_step(L, step, sync, st)) return rc;
        if (int rc = dual_update(L, step, true, st)) return rc;
    }
    return 0;
}

}  // namespace

extern "C" {

int aldvc_admm_loop_enqueue(const AldvcAdmmLoop* loop, int sync_every, hipStream_t stream) {
    clear_error();
    if (int rc = check_loop(loop)) return rc;
    if (sync_every < 0) return fail(-2, "sync_every must be 0 (all iterations enqueued) or positive");
    return enqueue(loop, sync_every, stream);
}

int aldvc_admm_loop_graph(const AldvcAdmmLoop* loop, hipStream_t capture_stream, void** graph) {
    clear_error();
    if (int rc = check_loop(loop)) return rc;
    if (!graph) return fail(-2, "null graph handle");
    if (!capture_stream) return fail(-2, "the null stream cannot be captured; pass a stream created by the caller");
    *graph = nullptr;
    if (hipStreamBeginCapture(capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(-5, "hipStreamBeginCapture failed");
    const int rc = enqueue(loop, 0, capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(capture_stream, &g);  // always ends the capture
    if (rc) {  // the enqueue's own message stays the error
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess || !g) return fail(-5, "hipStreamEndCapture failed");
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess || !exec) return fail(-5, "hipGraphInstantiate failed");
    *graph = static_cast<void*>(exec);
    return 0;
}

int aldvc_graph_launch(void* graph, hipStream_t stream) {
    clear_error();
    if (!graph) return fail(-2, "null graph");
    if (hipGraphLaunch(static_cast<hipGraphExec_t>(graph), stream) != hipSuccess) return fail(-5, "hipGraphLaunch failed");
    return 0;
}

int aldvc_graph_destroy(void* graph) {
    clear_error();
    if (!graph) return 0;
    if (hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph)) != hipSuccess)
        return fail(-5, "hipGraphExecDestroy failed");
    return 0;
}

}  // extern "C"
