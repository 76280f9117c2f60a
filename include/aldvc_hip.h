/*
 * aldvc_hip.h -- C ABI of libaldvc_hip.so, the gfx950 (MI355X) local DVC step.
 *
 * Every entry point takes DEVICE pointers (caller-owned memory, e.g. torch tensors or
 * hipMalloc'd buffers), plain sizes and a hipStream_t; it enqueues its kernels on that stream
 * and returns 0, or a negative code with a message for aldvc_last_error(): -2 bad argument,
 * -3 workspace too small, -4 a copy failed, -5 a launch failed. Nothing synchronises the
 * stream. Per-node failures are status codes in the outputs, never errors.
 *
 * Layouts (reference al_dvc 1.2.1, core/data_structures.py:1-11, solver/numba_kernels.py:9-16):
 *   volumes     (nz, ny, nx) C-order float32, all with the shape of f
 *   coords      (N, 3) int64 node centres [x, y, z]
 *   P           (N, 12) float64 = [F row-major (9), u, v, w]
 *   H, L        (N, 12, 12) float64; pattern (12, 12) float64
 *   status      int8 codes: 0 converged, 1 max_iter, 2 out_of_bounds, 3 invalid_subset,
 *               4 singular, 5 nan, 6 skipped, 7 stalled (numba_kernels.py:29-36)
 *   gx/gy/gz == NULL  gradients evaluated on the fly by the 7-point stencil
 *   mask == NULL      every voxel counts (the reference's 1x1x1 placeholder)
 *   split_index == NULL  whole subsets; else row r >= 0 of split_keep (keep_bytes per row,
 *               bit i = sampled voxel i in (dz, dy, dx) order) restricts node n's subset
 *   interp mode 0 linear, 1 Keys cubic (a = -0.5), 2 cubic B-spline on prefiltered coefficients
 */
#ifndef ALDVC_HIP_H
#define ALDVC_HIP_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ABI version of this header (2). */
int aldvc_abi_version(void);

/* Hash of the sources the library was compiled from (16 hex digits); no counterpart in the
 * reference: the Python loader refuses a library whose hash differs from the sources on disk. */
const char *aldvc_source_hash(void);

/* Message of the last failed call on this thread ("" after a successful one). */
const char *aldvc_last_error(void);

/* Measurement hook, no counterpart in the reference: with on != 0, the following aldvc_icgn12
 * (variants 1 and 2) and aldvc_icgn3 calls on this thread record HIP events around their main
 * iteration kernel on the call's stream. */
int aldvc_kernel_timing(int on);

/* Duration in ms of the last timed kernel on this thread (waits for it); -1 if there is none. */
float aldvc_last_kernel_ms(void);

/* Measurement hook, no counterpart in the reference: with buf (device, 2 x capacity unsigned 64-bit
 * values) the following fused 12-DOF and 3-DOF launches on this thread over at most capacity nodes
 * rounded down to a multiple of 1024 (capacity >= N rounded up to a multiple of 1024) write the
 * wall-clock tick at which the solve of the i-th entry of the launch's valid-node list starts,
 * buf[2 i], and ends, buf[2 i + 1] (entries beyond the valid nodes are left alone); NULL switches it
 * off. *clock_khz (host, nullable) receives the tick rate. For measuring how the launch fills the
 * machine: nodes in flight over time, the tail, per-node durations. */
int aldvc_workgroup_clock(void *buf, int64_t capacity, int *clock_khz);

/* 7-point central-difference gradients, zero within `border` voxels of any face.
 * Replaces io/volume_ops.py:124-197 (_gradient_stencil7 / compute_gradients). */
int aldvc_gradients(const float *f, int64_t nz, int64_t ny, int64_t nx, int border, float *gx, float *gy,
                    float *gz, hipStream_t stream);

/* Interpolate `vol` at n points (float64 coordinates, float32 result); NaN outside the
 * admissible domain 1 <= c <= n-3. Replaces solver/interp_kernels.py:139-152 (sample_points). */
int aldvc_sample_points(const float *vol, int64_t nz, int64_t ny, int64_t nx, const double *zs,
                        const double *ys, const double *xs, int64_t n, int mode, float *out,
                        hipStream_t stream);

/* Per-node Hessian, Cholesky factor, ZNSSD normalisation and validity.
 * Replaces solver/numba_kernels.py:499-509 (precompute_nodes) and
 * solver/cuda_kernels.py:1635-1710 (precompute_nodes_cuda). Outputs: H, L (N,12,12), meanf,
 * bottomf (N) float64, n_valid (N) int64, valid (N) uint8. */
int aldvc_precompute(const int64_t *coords, int64_t N, int hx, int hy, int hz, int stride, const float *f,
                     const float *gx, const float *gy, const float *gz, const uint8_t *mask, int64_t nz,
                     int64_t ny, int64_t nx, double min_valid_ratio, double cond_max,
                     const int64_t *split_index, const uint8_t *split_keep, int64_t keep_bytes, double *H,
                     double *L, double *meanf, double *bottomf, int64_t *n_valid, uint8_t *valid,
                     hipStream_t stream);

/* aldvc_precompute with a workspace: flags = 32 takes the sums of the nodes without keep rows
 * (stored gradients, stride 1) by a separable pass over the distinct node coordinates (double
 * sums, weighted by the mask) when the workspace holds aldvc_precompute_row_bytes for them (else,
 * or when the node set needs more, the per-node pass runs, as with flags 0). Same reference
 * function as aldvc_precompute. */
int aldvc_precompute_tuned(const int64_t *coords, int64_t N, int hx, int hy, int hz, int stride,
                           const float *f, const float *gx, const float *gy, const float *gz,
                           const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, double min_valid_ratio,
                           double cond_max, const int64_t *split_index, const uint8_t *split_keep,
                           int64_t keep_bytes, double *H, double *L, double *meanf, double *bottomf,
                           int64_t *n_valid, uint8_t *valid, int flags, void *workspace,
                           size_t workspace_bytes, hipStream_t stream);

/* Workspace of flag 32 of aldvc_precompute_tuned for nodes with ncx distinct x, ncy distinct y and
 * z within a range of nplanes planes including the subset half-widths, over those without keep rows
 * whose subsets lie inside the volume (an upper bound is fine); 0 when a count is not positive. No
 * counterpart in the reference (a workspace size). */
size_t aldvc_precompute_row_bytes(int64_t ncx, int64_t ncy, int64_t nplanes, int64_t ny, int64_t nx);

/* Workspace bytes aldvc_icgn12 needs for N nodes (variant 0 adds the 2048-node scratch rows). */
size_t aldvc_icgn12_workspace_bytes(int64_t N, int hx, int hy, int hz, int stride, int variant);

/* 12-DOF inverse-compositional Gauss-Newton on the ZNSSD criterion over all valid nodes; P is
 * updated in place, n_iter/status/zncc are written for every node.
 * Replaces solver/numba_kernels.py:1211-1278 (icgn_12dof_parallel) and
 * solver/cuda_kernels.py:1397-1515 (icgn_12dof_cuda). H and pattern (noise-corrected Hessian,
 * used when noise_gain > 0) may be NULL; meanf (N) is the accumulation pilot of variant 1 and
 * may be NULL. variant 0 = plain (one 256-lane workgroup per node, two passes through scratch
 * rows, the reference GPU kernel's design), 1 = fused one-pass kernel (deformed-volume stencils
 * staged in LDS bricks where the warped subset fits one; memory gathers elsewhere), 2 = the fused
 * kernel with memory gathers only. */
int aldvc_icgn12(const int64_t *coords, int64_t N, double *P, int hx, int hy, int hz, int stride,
                 const float *f, const float *gx, const float *gy, const float *gz, const uint8_t *mask,
                 const float *g, int64_t nz, int64_t ny, int64_t nx, int mode, const double *L,
                 const double *H, const double *pattern, double noise_gain, const double *meanf,
                 const uint8_t *valid, double tol, double dp_tol, int max_iter, int patience, int predictive,
                 const int64_t *split_index, const uint8_t *split_keep, int64_t keep_bytes, int32_t *n_iter,
                 int8_t *status, double *zncc, int variant, void *workspace, size_t workspace_bytes,
                 hipStream_t stream);

/* aldvc_icgn12 with the launch choices of the fused variants made explicit (the autotuner's search
 * space; results keep their meaning, float32 sums may round differently): lds_budget = LDS bytes
 * per workgroup for the brick (0: three workgroups per CU; at most 160 KiB), flags = 1: no
 * plane-pair lane mapping, 2: node list in index order instead of spatial tiles, 4: workgroup b
 * takes list slot b instead of the XCD-aware mapping, 8: no 8 x 8 lane patches (the default
 * mapping of the brick path), 32: reference moments by the separable pass over the distinct node
 * coordinates when the gradients are stored and stride is 1, for the nodes without keep rows
 * (double sums, weighted by the mask; the workspace then also holds
 * aldvc_icgn12_row_moments_bytes past aldvc_icgn12_workspace_bytes, else, or when the node set
 * needs more, the per-node pass runs).
 * Same reference function as aldvc_icgn12. */
int aldvc_icgn12_tuned(const int64_t *coords, int64_t N, double *P, int hx, int hy, int hz, int stride,
                       const float *f, const float *gx, const float *gy, const float *gz, const uint8_t *mask,
                       const float *g, int64_t nz, int64_t ny, int64_t nx, int mode, const double *L_all,
                       const double *H_all, const double *pattern, double noise_gain, const double *meanf,
                       const uint8_t *valid, double tol, double dp_tol, int max_iter, int patience, int predictive,
                       const int64_t *split_index, const uint8_t *split_keep, int64_t keep_bytes, int32_t *n_iter,
                       int8_t *status, double *zncc, int variant, int lds_budget, int flags, void *workspace,
                       size_t workspace_bytes, hipStream_t stream);

/* Extra workspace of flag 32 of aldvc_icgn12_tuned / aldvc_icgn3_tuned for nodes with ncx distinct x,
 * ncy distinct y and z within a range of nplanes planes including the subset half-widths
 * (max z - min z + 2 hz + 1), over those without keep rows whose subsets lie inside the volume (an
 * upper bound is fine). 0 when a count is not positive. No counterpart in the reference (a
 * workspace size). */
size_t aldvc_icgn12_row_moments_bytes(int64_t ncx, int64_t ncy, int64_t nplanes, int64_t ny, int64_t nx);

/* aldvc_icgn12 that allocates nothing. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The workspace is a fixed 64 MiB device array inside the library, one per GPU, held
 * until the process ends; its address is looked up on the first call on each device (the stream's
 * device must be current). Calls whose workspace (aldvc_icgn12_workspace_bytes) exceeds it return
 * -2: at most 254,199 nodes for the fused variants (1, 2); the plain variant (0) needs 2048 rows of
 * S floats, so it fits only small subsets. Nothing in the scratch must start at zero (every call
 * writes what it reads). Calls that overlap on one GPU share the array and race: serialise them. */
int aldvc_icgn12_noalloc(const int64_t *coords, int64_t N, double *P, int hx, int hy, int hz, int stride,
                         const float *f, const float *gx, const float *gy, const float *gz, const uint8_t *mask,
                         const float *g, int64_t nz, int64_t ny, int64_t nx, int mode, const double *L,
                         const double *H, const double *pattern, double noise_gain, const double *meanf,
                         const uint8_t *valid, double tol, double dp_tol, int max_iter, int patience, int predictive,
                         const int64_t *split_index, const uint8_t *split_keep, int64_t keep_bytes,
                         int32_t *n_iter, int8_t *status, double *zncc, int variant, hipStream_t stream);

/* Workspace bytes aldvc_icgn3 needs for N nodes. */
size_t aldvc_icgn3_workspace_bytes(int64_t N);

/* ADMM subproblem 1: translation-only IC-GN with the proximal term mu (u - u_hat - v), F held
 * fixed. U receives the result (U_old is copied into it first).
 * Replaces solver/numba_kernels.py:1281-1346 (icgn_3dof_parallel) and
 * solver/cuda_kernels.py:1518-1632 (icgn_3dof_cuda). U_old, U, vdual (N,3); F_fixed (N,3,3). */
int aldvc_icgn3(const int64_t *coords, int64_t N, const double *U_old, double *U, const double *F_fixed,
                const double *vdual, int hx, int hy, int hz, int stride, const float *f, const float *gx,
                const float *gy, const float *gz, const uint8_t *mask, const float *g, int64_t nz, int64_t ny,
                int64_t nx, int mode, const double *H, const double *meanf, const double *bottomf,
                const uint8_t *valid, double mu, double tol, double dp_tol, int max_iter, int patience,
                double noise_gain, int predictive, const int64_t *split_index, const uint8_t *split_keep,
                int64_t keep_bytes, int32_t *n_iter, int8_t *status, double *zncc, void *workspace,
                size_t workspace_bytes, hipStream_t stream);

/* aldvc_icgn3 with a device flag `gate` (int): when *gate is nonzero at execution time the call
 * writes nothing (U, n_iter, status, zncc keep their contents). For ADMM loops recorded into a
 * graph, whose iterations after the stopping rule fired must leave the state alone (the loop of
 * core/pipeline.py:320-358 breaks there). */
int aldvc_icgn3_gated(const int64_t *coords, int64_t N, const double *U_old, double *U, const double *F_fixed,
                      const double *vdual, int hx, int hy, int hz, int stride, const float *f, const float *gx,
                      const float *gy, const float *gz, const uint8_t *mask, const float *g, int64_t nz,
                      int64_t ny, int64_t nx, int mode, const double *H_all, const double *meanf,
                      const double *bottomf, const uint8_t *valid, double mu, double tol, double dp_tol, int max_iter,
                      int patience, double noise_gain, int predictive, const int64_t *split_index,
                      const uint8_t *split_keep, int64_t keep_bytes, int32_t *n_iter, int8_t *status, double *zncc,
                      const int *gate, void *workspace, size_t workspace_bytes, hipStream_t stream);

/* aldvc_icgn3_gated (gate may be NULL) with the launch choices made explicit, as aldvc_icgn12_tuned:
 * variant = 1: LDS bricks where a chunk's warped stencils fit one (subsets wider than a brick in
 * x/y boxes), 2: memory gathers only; lds_budget and flags as aldvc_icgn12_tuned (flag 32 with its
 * workspace past aldvc_icgn3_workspace_bytes), plus flag 16:
 * the workspace already holds these nodes' reference moments from a previous call with the same
 * nodes, reference volume, gradients, mask / keep rows and meanf (they are then not recomputed;
 * an ADMM loop's later 3-DOF passes). The other two 3-DOF entries run variant 1 with lds_budget 0
 * and flags 0. Same reference function as aldvc_icgn3. */
int aldvc_icgn3_tuned(const int64_t *coords, int64_t N, const double *U_old, double *U, const double *F_fixed,
                      const double *vdual, int hx, int hy, int hz, int stride, const float *f, const float *gx,
                      const float *gy, const float *gz, const uint8_t *mask, const float *g, int64_t nz,
                      int64_t ny, int64_t nx, int mode, const double *H_all, const double *meanf,
                      const double *bottomf, const uint8_t *valid, double mu, double tol, double dp_tol, int max_iter,
                      int patience, double noise_gain, int predictive, const int64_t *split_index,
                      const uint8_t *split_keep, int64_t keep_bytes, int32_t *n_iter, int8_t *status, double *zncc,
                      const int *gate, int variant, int lds_budget, int flags, void *workspace,
                      size_t workspace_bytes, hipStream_t stream);

/* ZNCC of each valid node's subset under the warp P_all (NaN where not sampled).
 * Replaces solver/numba_kernels.py:1354-1387 (evaluate_zncc_parallel). */
int aldvc_evaluate_zncc(const int64_t *coords, int64_t N, const double *P_all, int hx, int hy, int hz,
                        int stride, const float *f, const uint8_t *mask, const float *g, int64_t nz, int64_t ny,
                        int64_t nx, int mode, const uint8_t *valid, double *out, hipStream_t stream);

/* Workspace bytes aldvc_zncc_map needs (0 for the float32 engine). */
size_t aldvc_zncc_map_workspace_bytes(int64_t n, int hx, int hy, int hz, int rx, int ry, int rz, int precision);

/* ZNCC maps out (n, 2rz+1, 2ry+1, 2rx+1) float32 of the (2h+1)^3 template around each node
 * against the search window starting at (wx0, wy0, wz0); -2 for t_ok == 0 or a flat template.
 * Replaces solver/integer_search.py:44-151 (_zncc_direct) and solver/cuda_kernels.py:1713-1755
 * (zncc_direct_cuda). precision 0 = float32. */
int aldvc_zncc_map(const float *f, const float *g, int64_t nz, int64_t ny, int64_t nx, const int64_t *coords,
                   int64_t n, int hx, int hy, int hz, int rx, int ry, int rz, const int64_t *wx0,
                   const int64_t *wy0, const int64_t *wz0, const uint8_t *t_ok, float *out, int precision,
                   void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Peak statistics of ZNCC maps: first-index argmax (C order), peak value, PCE and the 27-point
 * quadratic sub-voxel offset (x, y, z) for interior peaks; pinv is the (10, 27) pseudo-inverse
 * of the quadratic design matrix. Replaces the per-node peak logic of
 * solver/integer_search.py:444-477 (with _subvoxel_peak, :297-321). */
int aldvc_ncc_peaks(const float *maps, int64_t n, int rx, int ry, int rz, const double *pinv, int64_t *peak_idx,
                    double *peak_val, double *pce, double *sub, hipStream_t stream);

/* Bytes of the summed-volume table aldvc_subset_valid_fraction needs as workspace. */
size_t aldvc_mask_table_bytes(int64_t nz, int64_t ny, int64_t nx);

/* In-mask fraction of each node's window [c - h, c + h] (clipped to the volume, over the nominal
 * (2h+1)^3 volume) into frac (N) float64. Replaces mesh/grid_mesh.py:135-188
 * (subset_valid_fraction); winsize w gives h = w / 2 as there. */
int aldvc_subset_valid_fraction(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx,
                                const int64_t *coords, int64_t N, int hx, int hy, int hz, double *frac,
                                void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Packed keep rows (M, keep_bytes) of the 6-connected in-mask component around the centre of each
 * candidate subset cand[m] (bit i of a row = sampled voxel i in the kernels' order), with the kept
 * and in-mask sampled-voxel counts. Voxels within `margin` of a volume face count as masked.
 * Replaces solver/numba_kernels.py:223-300 (build_split_rows) and utils/flood_fill.py:12-72.
 * Windows up to 64 voxels along x and 4096 (y, z) rows. */
int aldvc_split_rows(const int64_t *coords, const int64_t *cand, int64_t M, int hx, int hy, int hz,
                     int stride, const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, int margin,
                     int64_t keep_bytes, uint8_t *rows, int64_t *n_keep, int64_t *n_inmask,
                     hipStream_t stream);

/* Bytes of the workspace aldvc_voi_moments needs for a volume of nz planes of ny rows. */
size_t aldvc_voi_moments_workspace_bytes(int64_t nz, int64_t ny);

/* sums[0] = sum (v - pilot), sums[1] = sum (v - pilot)^2 over the half-open VOI [z0, z1) x
 * [y0, y1) x [x0, x1), float64, deterministic. dtype: 0 f32, 1 u16, 2 u8, 3 i16, 4 f64.
 * Replaces io/volume_ops.py:53-70 (_voi_moments) inside voi_mean_std (:83-102). */
int aldvc_voi_moments(const void *vol, int dtype, int64_t nz, int64_t ny, int64_t nx, int z0, int z1, int y0,
                      int y1, int x0, int x1, double pilot, double *sums, void *workspace,
                      size_t workspace_bytes, hipStream_t stream);

/* out[i] = float32((vol[i] - mean) * inv_std) in float64 arithmetic.
 * Replaces io/volume_ops.py:73-80 (_normalize_into) inside normalize_volume (:105-121). */
int aldvc_normalize(const void *vol, int dtype, int64_t n, double mean, double inv_std, float *out,
                    hipStream_t stream);

/* Cubic B-spline coefficients of vol (whole-sample mirror boundaries) into out; scratch is a third
 * volume-sized float32 buffer. Replaces io/volume_ops.py:215-222 (prefilter_bspline, i.e.
 * scipy.ndimage.spline_filter(order=3, mode="mirror") into float32). */
int aldvc_bspline_prefilter(const float *vol, int64_t nz, int64_t ny, int64_t nx, float *out, float *scratch,
                            hipStream_t stream);

/* vol[i] = NaN where mask[i] == 0. Replaces the masking step of io/volume_ops.py:225-244
 * (prepare_deformed). */
int aldvc_mask_nan(float *vol, const uint8_t *mask, int64_t n, hipStream_t stream);

/* Block mean of vol (nz, ny, nx) by an integer factor into out (nz/f, ny/f, nx/f), trailing voxels
 * dropped; each mean summed in float64 and rounded once to float32. Replaces
 * solver/integer_search.py:235-243 (block_downsample). */
int aldvc_block_downsample(const float *vol, int64_t nz, int64_t ny, int64_t nx, int factor, float *out,
                           hipStream_t stream);

/* ---- ADMM global step on the hex8 node grid (float64). A grid of nx*ny*nz nodes, node
 * n = (iz*ny + iy)*nx + ix; elem_active (nz-1)(ny-1)(nx-1) and node_active (nz)(ny)(nx) uint8;
 * element matrices 8x8 row-major in the grid's corner order; vectors (N, 3). */

/* y = A x on active nodes, 0 elsewhere, A the operator of the element matrix Ae. Replaces the
 * assembled products of solver/global_operators.py:58-70 (build_fem_operators). */
int aldvc_fem_apply(const double *Ae, int nx, int ny, int nz, const uint8_t *elem_active,
                    const uint8_t *node_active, const double *x, double *y, hipStream_t stream);

/* b = beta sum_j G_j (F - W)[:, :, j] + mu M (U - v) + alpha K U; mats = K, M, G_x, G_y, G_z
 * (5 x 64), FmW_j (3, N, 3) = (F - W)[:, i, j] per j. Replaces solver/subpb2_solver.py:108-127
 * (global_rhs). */
int aldvc_fem_rhs(const double *mats, int nx, int ny, int nz, const uint8_t *elem_active,
                  const uint8_t *node_active, double beta, double mu, double alpha, const double *FmW_j,
                  const double *Umv, const double *U, double *b, hipStream_t stream);

/* F (N, 3, 3) = (G_j^T U[:, i]) / lumped mass, NaN at inactive nodes. Replaces
 * solver/global_operators.py:147-170 (nodal_gradient, FEM). */
int aldvc_fem_nodal_gradient(const double *mats, int nx, int ny, int nz, const uint8_t *elem_active,
                             const uint8_t *node_active, const double *U, double *F, hipStream_t stream);

/* Workspace bytes of aldvc_pcg_fem. */
size_t aldvc_pcg_workspace_bytes(int nx, int ny, int nz);

/* Jacobi-preconditioned CG for A X = B (three columns together) on the active nodes; X holds the
 * start and receives the solution; B must be 0 at inactive nodes. Stops when every column's
 * |R| / |B| < tol or after max_iter iterations (count in *iterations, host memory). Replaces
 * solver/subpb2_solver.py:77-105 (pcg_multi) inside solve_subpb2 (:130-155). */
int aldvc_pcg_fem(const double *Ae, int nx, int ny, int nz, const uint8_t *elem_active,
                  const uint8_t *node_active, const double *B, double *X, double tol, int max_iter,
                  int *iterations, void *workspace, size_t workspace_bytes, hipStream_t stream);

/* aldvc_pcg_fem with the stopping rule left on the device: every one of the max_iter iterations
 * is enqueued (those after convergence return at once) and nothing is read back, so the call can
 * be recorded into a graph. *iterations_dev (device, nullable) receives the count; a nonzero
 * *gate (device, nullable) switches the solve off. Same workspace as aldvc_pcg_fem. */
int aldvc_pcg_fem_async(const double *Ae, int nx, int ny, int nz, const uint8_t *elem_active,
                        const uint8_t *node_active, const double *B, double *X, double tol, int max_iter,
                        int *iterations_dev, const int *gate, void *workspace, size_t workspace_bytes,
                        hipStream_t stream);

/* ---- One ADMM iteration on the device (core/pipeline.py:286-366). Every call takes `stop`, a
 * device int (nullable): when nonzero at execution time the call leaves its outputs alone. */

/* Workspace bytes of aldvc_admm_global_step. */
size_t aldvc_admm_global_workspace_bytes(int nx, int ny, int nz);

/* The global step from the per-node arrays: b = beta sum_j G_j (F - W)[:, :, j] + mu M (U - v) +
 * alpha K U (zero rows at inactive nodes), U2 = solution of A U2 = b from U2 = U by the
 * device-controlled PCG (inactive nodes keep U), F2 = nodal gradient of U2 (NaN at inactive
 * nodes). U, v, U2 (N, 3); F, W, F2 (N, 3, 3); Ae = (beta + alpha) K + mu M (8 x 8).
 * sync_every = 0: all max_iter CG slots are enqueued (graph capture); > 0: the host reads the stop
 * flag after sync_every iterations, then at intervals doubling up to 4 sync_every, and enqueues no
 * more after convergence (eager callers).
 * capped (device int, nullable) is set to 1 when the solve used all max_iter iterations, so a
 * graph recorded with a smaller iteration budget can be checked after replay.
 * Replaces solver/subpb2_solver.py:108-155 (global_rhs, solve_subpb2) followed by
 * solver/global_operators.py:147-170 (nodal_gradient). */
int aldvc_admm_global_step(const double *mats, const double *Ae, int nx, int ny, int nz, const uint8_t *elem_active,
                           const uint8_t *node_active, double beta, double mu, double alpha, const double *U,
                           const double *F, const double *W, const double *v, double tol, int max_iter, double *U2,
                           double *F2, int *iterations_dev, const int *gate, int sync_every, int *capped,
                           void *workspace, size_t workspace_bytes, hipStream_t stream);

/* ---- The finite-difference discretisation of the global step (subpb2_method "fd"). Node grid
 * (nx, ny, nz) with spacing (hx, hy, hz); the operators are described by one byte per node, its
 * code (aldvc_fd_codes), and applied matrix-free: D_j (central differences, one-sided where one
 * neighbour is missing, zero rows at invalid nodes), Kg = sum_j D_j^T D_j, M = diag(valid),
 * G_j = D_j^T; the active nodes are the valid ones. Grids of 1 node along an axis are allowed. */

/* codes (N) from node_valid (N, 0/1) and edge_ok ((N, 3) 0/1, the +x, +y, +z edge of every node of
 * a cut mesh, or NULL). Replaces solver/global_operators.py:74-137 (_difference_operator,
 * build_fd_operators). */
int aldvc_fd_codes(int nx, int ny, int nz, const uint8_t *node_valid, const uint8_t *edge_ok, uint8_t *codes,
                   hipStream_t stream);

/* y = cK Kg x + cM M x on the valid nodes, 0 elsewhere; x, y (N, 3). The operator of
 * build_global_system (solver/subpb2_solver.py:45-74) with cK = beta + alpha, cM = mu. */
int aldvc_fd_apply(const uint8_t *codes, int nx, int ny, int nz, double hx, double hy, double hz, double cK,
                   double cM, const double *x, double *y, hipStream_t stream);

/* b = mu M (U - v) + beta sum_j G_j (F - W)[:, :, j] + alpha Kg U (N, 3), zero rows at invalid
 * nodes; U, v (N, 3), F, W (N, 3, 3). Replaces solver/subpb2_solver.py:108-127 (global_rhs). */
int aldvc_fd_rhs(const uint8_t *codes, int nx, int ny, int nz, double hx, double hy, double hz, double beta,
                 double mu, double alpha, const double *U, const double *F, const double *W, const double *v,
                 double *b, hipStream_t stream);

/* F[a, i, j] = (D_j U[:, i])(a) (N, 3, 3), NaN at invalid nodes. Replaces the FD branch of
 * solver/global_operators.py:151-170 (nodal_gradient). */
int aldvc_fd_nodal_gradient(const uint8_t *codes, int nx, int ny, int nz, double hx, double hy, double hz,
                            const double *U, double *F, hipStream_t stream);

/* Jacobi-preconditioned CG for (cK Kg + cM M) X = B on the valid nodes, as aldvc_pcg_fem (same
 * workspace, aldvc_pcg_workspace_bytes). Replaces solver/subpb2_solver.py:77-105 (pcg_multi). */
int aldvc_pcg_fd(const uint8_t *codes, int nx, int ny, int nz, double hx, double hy, double hz, double cK, double cM,
                 const double *B, double *X, double tol, int max_iter, int *iterations, void *workspace,
                 size_t workspace_bytes, hipStream_t stream);

/* aldvc_admm_global_step with the FD operators (same arguments otherwise, same workspace,
 * aldvc_admm_global_workspace_bytes). Replaces solver/subpb2_solver.py:108-155 followed by
 * solver/global_operators.py:151-170 for subpb2_method "fd" (core/pipeline.py:216, :305). */
int aldvc_admm_global_step_fd(const uint8_t *codes, int nx, int ny, int nz, double hx, double hy, double hz,
                              double beta, double mu, double alpha, const double *U, const double *F,
                              const double *W, const double *v, double tol, int max_iter, double *U2, double *F2,
                              int *iterations_dev, const int *gate, int sync_every, int *capped, void *workspace,
                              size_t workspace_bytes, hipStream_t stream);

/* Workspace bytes of aldvc_subpb1_cleanup. */
size_t aldvc_subpb1_cleanup_workspace_bytes(int nx, int ny, int nz);

/* The local step's cleanup on the node grid (nx, ny, nz): bad = status != CONVERGED, or, when
 * threshold > 0 and more than 27 nodes converged, flagged by the universal median test of U
 * (eps: the test's noise level, 0.1 in the reference); U_out = bad ? U_hat : U; bad (N) 0/1.
 * edge_ok (N, 3) uint8, the +x, +y, +z edge of every node of a cut mesh, or NULL: without it the
 * unknown nodes are spring-filled first (tolerance fill_tol relative, at most fill_max_iter CG
 * iterations; sync_every and capped as aldvc_admm_global_step; fill_iterations (device int,
 * nullable) receives the fill's CG iteration count). Replaces solver/subpb1_solver.py:140-164 with
 * utils/outlier_detection.py:98-160 and utils/inpaint.py:35-121. */
int aldvc_subpb1_cleanup(int nx, int ny, int nz, const uint8_t *edge_ok, const int8_t *status, const double *U,
                         const double *U_hat, double threshold, double eps, double fill_tol, int fill_max_iter,
                         double *U_out, uint8_t *bad, const int *stop, int sync_every, int *fill_iterations,
                         int *capped, void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Workspace bytes of aldvc_spring_fill. */
size_t aldvc_spring_fill_workspace_bytes(int nx, int ny, int nz);

/* Spring-model fill of the NaN entries of vals (N, 3) on the node grid, column by column, in place:
 * the discrete Laplace equation over 6-neighbour springs (none across cut edges when edge_ok is
 * given), known nodes as data, a 1e-10 Tikhonov term. Jacobi-CG over the unknown nodes to tol
 * (relative), at most max_iter iterations: with at most 7,168 unknown nodes each column is its own
 * solve in one launch, else the three columns run together until all meet tol; *iterations (host)
 * receives the count (the largest over the columns). Replaces utils/inpaint.py:35-121
 * (fill_nan_grid, "spring"). */
int aldvc_spring_fill(int nx, int ny, int nz, const uint8_t *edge_ok, double *vals, double tol, int max_iter,
                      int *iterations, void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Workspace bytes of aldvc_dual_update. */
size_t aldvc_dual_update_workspace_bytes(int64_t N);

/* The end of an ADMM iteration: history[4 step + k] = RMS of U2n - U2, U1n - U1, U2n - U1n and
 * F2n' - F2 where F2n' = finite(F2n) ? F2n : F2; then W = W + (F2n' - F2) and v = v + (U2n - U1n)
 * (accumulate; else the differences alone), U1 = U1n, U2 = U2n, F2 = F2n'; *steps = step. With
 * check_stop, *stop = 1 when either of the first two RMS values is below admm_tol.
 * Replaces core/pipeline.py:300-358 (the dual update and the stopping rule). */
int aldvc_dual_update(int64_t N, const double *U1n, const double *U2n, const double *F2n, double *U1, double *U2,
                      double *F2, double *W, double *v, int accumulate, int check_stop, double admm_tol, int step,
                      double *history, int *stop, int *steps, void *workspace, size_t workspace_bytes,
                      hipStream_t stream);

/* ---- The whole ADMM loop of a frame in one call, and as one HIP graph (core/pipeline.py:286-366).
 * AldvcAdmmLoop names every device buffer and setting of the loop; its fields are the arguments of
 * the four entries above that one iteration calls. Sequence: U1n = U1; the first global pass
 * (step 1: global step, dual update without the stopping rule); then for step = 2 .. max_steps:
 * aldvc_icgn3_tuned (U_old = U2, F_fixed = F2, vdual = v, out U1raw; from step 3 on with flag 16,
 * the reference moments of step 2), aldvc_subpb1_cleanup (U = U1raw, U_hat = U2, out U1n),
 * the global step (U = U1n, F = F2, W, v, out U2n, F2n) and aldvc_dual_update with the stopping
 * rule. Every call of step k writes row k of the per-step records; after the stop flag is set the
 * later calls change nothing. */
typedef struct AldvcAdmmLoop {
    size_t struct_bytes; /* sizeof(AldvcAdmmLoop): checked, so a caller built against another layout is refused */
    /* local 3-DOF step (aldvc_icgn3_tuned) */
    const int64_t *coords;
    int64_t n_nodes; /* N = nx * ny * nz of the node grid */
    int hx, hy, hz, stride;
    const float *f, *gx, *gy, *gz; /* gx..gz NULL: gradients on the fly */
    const uint8_t *mask;
    const float *g;
    int64_t vol_nz, vol_ny, vol_nx;
    int mode;
    const double *H_all, *meanf, *bottomf;
    const uint8_t *valid;
    double icgn_tol, icgn_dp_tol;
    int icgn_max_iter, icgn_patience;
    double noise_gain;
    int predictive;
    const int64_t *split_index;
    const uint8_t *split_keep;
    int64_t keep_bytes;
    int variant, lds_budget, flags;
    /* node grid, cleanup (aldvc_subpb1_cleanup) */
    int nx, ny, nz;
    const uint8_t *edge_ok; /* (N, 3) or NULL */
    double outlier_threshold, median_eps, fill_tol;
    int fill_budget, fill_limit; /* CG iterations per fill; a budget below the limit sets *capped when used up */
    /* global step: method 0 = FEM (mats, Ae, elem_active, node_active), 1 = FD (codes, spacing) */
    int method;
    const double *mats, *Ae;
    const uint8_t *elem_active, *node_active;
    const uint8_t *codes;
    double spacing_x, spacing_y, spacing_z;
    double beta, mu, alpha, pcg_tol;
    int pcg_budget, pcg_limit;
    /* dual update */
    int accumulate;
    double admm_tol;
    int max_steps;
    /* state: (N, 3) and (N, 9) float64 */
    double *U1, *U2, *v, *U1raw, *U1n, *U2n;
    double *F2, *W, *F2n;
    /* per-step records, rows 0 .. max_steps: status, zncc, n_iter, bad (N per row); history (4 per
     * row); pcg_iterations, fill_iterations (1 per row); stop, steps, capped (one int each) */
    int8_t *status;
    double *zncc;
    int32_t *n_iter;
    uint8_t *bad;
    double *history;
    int *pcg_iterations, *fill_iterations, *stop, *steps, *capped;
    /* workspaces of the four entries (their *_workspace_bytes) */
    void *ws_icgn3, *ws_cleanup, *ws_global, *ws_dual;
    size_t ws_icgn3_bytes, ws_cleanup_bytes, ws_global_bytes, ws_dual_bytes;
} AldvcAdmmLoop;

/* Enqueue the whole loop on `stream`. sync_every > 0: the CG solves are polled by the host as
 * aldvc_admm_global_step polls them (eager callers; the call then waits on the stream); 0: nothing
 * is read back and every CG slot of the budgets is enqueued, so the sequence can be captured. */
int aldvc_admm_loop_enqueue(const AldvcAdmmLoop *loop, int sync_every, hipStream_t stream);

/* Record aldvc_admm_loop_enqueue(loop, 0, .) into a HIP graph by capturing `capture_stream` (a
 * stream the caller created, not the null stream, with no other work being captured) and
 * instantiate it: *graph receives an executable graph for aldvc_graph_launch. A replay runs the
 * loop on whatever the loop's buffers hold at launch time; reset stop, steps, capped, the records
 * and the state (U1, U2, F2 = the local result, W = v = 0) before each. For callers without a
 * graph API of their own (C, Fortran, Java over JNI). */
int aldvc_admm_loop_graph(const AldvcAdmmLoop *loop, hipStream_t capture_stream, void **graph);

/* Launch / free an executable graph made by aldvc_admm_loop_graph. */
int aldvc_graph_launch(void *graph, hipStream_t stream);
int aldvc_graph_destroy(void *graph);

/* ---- Differentiable warp residual (no counterpart in the reference's backward; the forward is
 * solver/numba_kernels.py:1354-1387, evaluate_zncc_parallel, with the subset's affine warp replaced
 * by the trilinear FE interpolation of mesh/hex8.py:58-61). Node displacements node_u
 * (gnz, gny, gnx, 3) float32 on nodes at (ox, oy, oz) + (i sx, j sy, k sz); subsets of half-widths
 * (hx, hy, hz) at centers (S, 3) int64 [x, y, z]. */

/* C[s] = 2 (1 - ZNCC) of subset s (NaN when it leaves the volume or the sampling domain) and
 * stats (S, 8) = n, m_f, b_f, m_g, b_g, ZNCC, valid, C for the backward pass. */
int aldvc_warp_znssd(const float *f, const float *g, int64_t nz, int64_t ny, int64_t nx, int mode,
                     const int64_t *centers, int64_t S, int hx, int hy, int hz, const float *node_u, int gnx,
                     int gny, int gnz, double ox, double oy, double oz, double sx, double sy, double sz, double *C,
                     double *stats, hipStream_t stream);

/* Workspace bytes of aldvc_warp_znssd_backward (no counterpart in the reference, which has no
 * backward): per-voxel fields of the volume (36 B per voxel) plus small per-subset and per-cell parts. */
size_t aldvc_warp_backward_workspace_bytes(int64_t nz, int64_t ny, int64_t nx, int64_t S, int hx, int hy, int hz,
                                           int gnx, int gny, int gnz);

/* Accumulates grad_u += dL/du (float64, node layout of node_u) and, when grad_g is not NULL,
 * grad_g += dL/dg (float32 volume) for dL/dC = grad_C (S). Deterministic: every sum runs in a fixed
 * order (no float atomics), so two calls on the same inputs give bit-identical gradients. */
int aldvc_warp_znssd_backward(const float *f, const float *g, int64_t nz, int64_t ny, int64_t nx, int mode,
                              const int64_t *centers, int64_t S, int hx, int hy, int hz, const float *node_u,
                              int gnx, int gny, int gnz, double ox, double oy, double oz, double sx, double sy,
                              double sz, const double *stats, const double *grad_C, double *grad_u, float *grad_g,
                              void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Bridging elements of a cut mesh: out[e] = 1 when the eight corners (voxels [x, y, z], int64,
 * corners[e * 24 + 3 * k + d]) of element e do not all lie in the 6-connected in-mask component
 * of its first corner inside the corners' bounding box (clamped to the volume), or when that
 * corner is masked; 0 otherwise; 2 when the box exceeds max_extent (1..64) voxels along an axis.
 * Replaces mesh/mesh_cut.py:44-106 (_bridging_jit, bridging_elements). */
int aldvc_bridging_elements(const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, const int64_t *corners, int64_t E,
                            int max_extent, uint8_t *out, hipStream_t stream);

/* ---- Post-step: per-node displacement uncertainty (float64). */

/* Standard deviation of u, v, w (out: n x 3, voxels) at every node from its IC-GN Hessian H
 * (n x 12 x 12), final zncc, status (null: all converged), valid, n_valid and bottomf; NaN where
 * no estimate exists (not valid, zncc not finite, n_valid <= 14, or status not converged).
 * moment_diag is a HOST array of 4 doubles: the diagonal of the sampled subset's moment matrix
 * (sums of X^2, Y^2, Z^2 and the voxel count); gain is the gradient stencil's noise gain and
 * noise_correction selects the two-term model (0: 2 s^2 H^-1 with the stored Hessian).
 * *singular (device int, cleared by the caller) is set to 1 when a node that gets a value has an
 * exactly singular corrected Hessian; the caller then recomputes the batch with the
 * pseudo-inverse. Replaces solver/uncertainty.py:94-133 (displacement_uncertainty). */
int aldvc_displacement_uncertainty(const double *H, const double *zncc, const int8_t *status, const uint8_t *valid,
                                   const int64_t *n_valid, const double *bottomf, int64_t n, const double *moment_diag,
                                   double gain, int noise_correction, double *out, int *singular, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* ALDVC_HIP_H */
