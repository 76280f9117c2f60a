// Local DVC step on gfx950: per-node precompute and 12-DOF / 3-DOF IC-GN subset solvers.
//
// Two 12-DOF kernels share the float64 control step:
//   * icgn12_plain  -- the drop-in design of the reference GPU kernel (cuda_kernels.py:491-835):
//                      one 256-lane workgroup per node, pass 1 samples the warped subset into a
//                      global scratch row and reduces the ZNSSD statistics, pass 2 re-reads the
//                      row for the residual and the 12-vector, lane 0 runs the control in float64.
//   * icgn12_fused  -- one pass per iteration: every lane accumulates the 44 raw moments
//                      (n, sums of f, f^2, g, g^2, fg and of SD, SD*f, SD*g) from which the
//                      centred ZNSSD gradient follows in float64 after the reduction, so no
//                      scratch row exists and each subset voxel is read once per iteration.
//                      Nodes are mapped to workgroups so that the nodes in flight on one XCD
//                      are neighbours in the node list (their subsets overlap in that XCD's L2).
// Both keep the reference's float32 sampling / float64 control split (cuda_kernels.py:17-21).
#include "aldvc_device.h"

#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>

#include "aldvc_abi_util.h"

using namespace aldvc;

namespace {

constexpr int BLOCK = 256;
constexpr int NWAVES = BLOCK / 64;
constexpr int64_t PLAIN_CHUNK = 2048;  // nodes per launch of the plain kernel (cuda_kernels.py:67)

struct IcgnArgs {
    Volumes V;
    Subset sub;
    const int64_t* coords;
    double* P;             // (N, 12) in/out
    const double* L_all;   // (N, 144)
    const double* H_all;   // (N, 144) or nullptr
    const double* pattern; // (144) or nullptr
    const double* meanf;   // (N) precomputed reference mean: the accumulation pilot of the fused kernel
    const int64_t* split_index;
    const uint8_t* split_keep;
    int64_t keep_bytes;
    double noise_gain, tol, dp_tol;
    int max_iter, patience, predictive;
    int32_t* n_iter;
    int8_t* status;
    double* zncc;
    int xcd;  // 1: XCD-aware workgroup -> node mapping; 0: workgroup b takes list slot b
    // translation-only solves (ADMM subproblem 1): fixed displacement gradient, previous and
    // output displacements, dual variable, proximal weight
    const double* F;      // (N, 9)
    const double* U_old;  // (N, 3)
    const double* vdual;  // (N, 3)
    const double* bottomf;  // (N) reference subset norms of the precompute
    double* U;            // (N, 3) out
    double mu;
    // diagnostics (null in normal use): wall-clock ticks at the start and end of every
    // workgroup that takes a node, [2 b] and [2 b + 1] for workgroup b
    unsigned long long* wg_clock;
};

// Launch choices of the fused 12-DOF kernels that leave the results' meaning unchanged (the
// autotuner's search space): LDS bytes per workgroup for the brick (0: three workgroups per CU),
// and flag bits.
enum : int { TUNE_NO_PAIRS = 1, TUNE_INDEX_ORDER = 2, TUNE_NO_XCD = 4, TUNE_NO_PATCHES = 8 };
// translation-only solves: the workspace already holds the reference moments of these nodes
// (a previous call with the same nodes, reference volume, gradients, selection and pilots)
enum : int { ICGN3_REUSE_MOMENTS = 16 };
enum : int { LANES_WALK = 0, LANES_PAIRS = 1, LANES_PATCHES = 2 };  // voxel -> lane mappings of the brick path

// Per-node iteration state of the control lane (numba_kernels.py:663-677). It lives in LDS and
// only lane 0 touches it: held in registers, the float64 control would set the register budget
// of every lane of the sampling loops.
struct Ctl12 {
    double P[12], P_best[12], b[12], dP[12], y[12], work[36];
    double H0[144], L0[144];
    double norm_init, best_zncc, best_dp, last_dp, dp_prev, zncc;
    int stall;
};

template <typename C>
__device__ inline void ctl12_reset(C& c) {
    c.norm_init = -1.0;
    c.best_zncc = -2.0;
    c.best_dp = 1e300;
    c.last_dp = 1e300;
    c.dp_prev = -1.0;
    c.zncc = NAN;
    c.stall = 0;
}

__device__ inline void ctl12_init(Ctl12& c, const double* P) {
#pragma unroll 1
    for (int k = 0; k < 12; ++k) c.P[k] = c.P_best[k] = P[k];
    ctl12_reset(c);
}

// L L^T x = b with every array in memory and no unrolling (numba_kernels.py:106-119).
__device__ inline void chol_solve12_mem(const double* L, const double* b, double* y, double* x) {
#pragma unroll 1
    for (int i = 0; i < 12; ++i) {
        double s = b[i];
#pragma unroll 1
        for (int k = 0; k < i; ++k) s -= L[i * 12 + k] * y[k];
        y[i] = s / L[i * 12 + i];
    }
#pragma unroll 1
    for (int i = 11; i >= 0; --i) {
        double s = y[i];
#pragma unroll 1
        for (int k = i + 1; k < 12; ++k) s -= L[k * 12 + i] * x[k];
        x[i] = s / L[i * 12 + i];
    }
}

// One IC-GN decision from the reduced gradient c.b (already scaled by bottomf_d) and the ZNCC:
// numba_kernels.py:726-790. Returns -1 to continue, else the final status; c.P is updated.
__device__ int ctl12_step(Ctl12& c, double zncc, double bottomf_d, double n_valid, const double* L, const double* H,
                          const double* pattern, double noise_gain, double tol, double dp_tol, int patience,
                          int predictive, double half_scale) {
    double norm_abs = 0.0;
#pragma unroll 1
    for (int k = 0; k < 12; ++k) norm_abs += c.b[k] * c.b[k];
    norm_abs = sqrt(norm_abs);
    if (!isfinite(norm_abs)) {
        c.zncc = NAN;
        return ST_NAN;
    }
    if (c.norm_init < 0.0) c.norm_init = norm_abs;
    const double norm_rel = c.norm_init > 1e-300 ? norm_abs / c.norm_init : 0.0;
    c.zncc = zncc;
    if (norm_rel < tol || norm_abs < ABS_TOL) return ST_CONVERGED;
#pragma unroll 1
    for (int k = 0; k < 12; ++k) c.b[k] = -c.b[k];
    bool use_corr = false;
    const double n_full = pattern ? pattern[9 * 12 + 9] : 0.0;
    if (noise_gain > 0.0 && n_full > 0.0 && c.last_dp < NOISE_CORR_STEP && H) {
        const double om = 1.0 - zncc;
        const double s2 = (om > 0.0 ? om : 0.0) * bottomf_d * bottomf_d / n_valid;
        const double corr = noise_gain * s2 * (n_valid / n_full);
        use_corr = corrected_cholesky12(H, pattern, corr, c.H0, c.L0);
    }
    chol_solve12_mem(use_corr ? c.L0 : L, c.b, c.y, c.dP);
    double dp_norm = 0.0;
#pragma unroll 1
    for (int k = 0; k < 12; ++k) {
        const double v = k < 9 ? c.dP[k] * half_scale : c.dP[k];
        dp_norm += v * v;
    }
    dp_norm = sqrt(dp_norm);
    if (!isfinite(dp_norm)) {
        c.zncc = NAN;
        return ST_NAN;
    }
    c.last_dp = dp_norm;
    if (dp_norm < dp_tol) return ST_CONVERGED;
    if (predictive && c.dp_prev > 0.0 && dp_norm < PREDICT_CONTRACTION * c.dp_prev &&
        dp_norm * dp_norm < dp_tol * c.dp_prev)
        return compose_warp(c.P, c.dP, c.work) ? ST_CONVERGED : ST_SINGULAR;
    c.dp_prev = dp_norm;
    bool improved = zncc > c.best_zncc + STALL_ZNCC_EPS;
    if (improved) {
        c.best_zncc = zncc;
#pragma unroll 1
        for (int k = 0; k < 12; ++k) c.P_best[k] = c.P[k];
    }
    if (dp_norm < c.best_dp * STALL_STEP_DECAY) {
        c.best_dp = dp_norm;
        improved = true;
    }
    if (improved) {
        c.stall = 0;
    } else {
        c.stall += 1;
        if (patience > 0 && c.stall >= patience) {
#pragma unroll 1
            for (int k = 0; k < 12; ++k) c.P[k] = c.P_best[k];
            c.zncc = c.best_zncc;
            return ST_STALLED;
        }
    }
    return compose_warp(c.P, c.dP, c.work) ? -1 : ST_SINGULAR;
}

// ---------------------------------------------------------------------------- wave-parallel 12-DOF control
// State of the wave-parallel control (the serial solver's b, y and dP scratch is not needed: those
// live in the lanes' registers; `work` holds the warp composition's 3x3 intermediates).
struct Ctl12w {
    double P[12], P_best[12], work[27];
    double H0[144], L0[144];
    double norm_init, best_zncc, best_dp, last_dp, dp_prev, zncc;
    int stall;
};

// The decisions of ctl12_step (numba_kernels.py:726-790) taken by the 64 lanes of wave 0 at once,
// lane k < 12 owning parameter k: the two triangular solves are column sweeps (the entry solved on
// its lane is broadcast, every later row subtracts its share), the norms are wave sums, and the
// warp composition is formed by every lane from broadcast values, lane k keeping entry k. The
// state stays in LDS (Ctl12), so none of it is live in the sampling loops' registers. The serial
// form takes one dependent LDS or memory access per multiply-add of the solve; this one a few
// dozen dependent steps per iteration.

// This lane's index in its wave, computed where it is used: index arithmetic derived from it stays in
// the control code instead of being hoisted out of the node loop and kept live (and spilled) through
// the sampling loops.
__device__ inline int lane_here() {
    int l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

// sum over lanes 0..11 in lane order (the order of the serial loops), the same value on every lane;
// broadcasts through scalar registers need no lane-index registers
__device__ inline double sum12(double v) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 12; ++k) s += lane_bcast(v, k);
    return s;
}

// W(P) <- W(P) W(dP)^-1 (numba_kernels.py:145-180) with dP_k on lane k. Lane q < 9 forms entry q
// of (I + dP)^-1 (its cofactor over the determinant, as inv3), then entry q of A (I + dP)^-1, and
// lanes 9..11 the new translation; the intermediates pass through `work` (LDS, 27 doubles). LDS
// operations of one wave complete in order, so a lane reads what another lane wrote before.
// Returns false when dP's 3x3 part is singular.
__device__ inline bool compose_warp_wave(double* P, double* work, double dpk) {
    const int lane = lane_here();
    double* dA = work;       // I + dP (3x3)
    double* dAi = work + 9;  // its inverse
    double* An = work + 18;  // (I + P) (I + dP)^-1
    if (lane < 9) dA[lane] = dpk + ((lane & 3) == 0 ? 1.0 : 0.0);  // diagonal entries 0, 4, 8
    __builtin_amdgcn_wave_barrier();
    const double det = dA[0] * (dA[4] * dA[8] - dA[5] * dA[7]) - dA[1] * (dA[3] * dA[8] - dA[5] * dA[6]) +
                       dA[2] * (dA[3] * dA[7] - dA[4] * dA[6]);
    if (fabs(det) < 1e-300) return false;
    const int q = lane < 9 ? lane : 0, r = q / 3, cl = q - 3 * r;
    if (lane < 9) {
        // adj[r][c] = a[c+1][r+1] a[c+2][r+2] - a[c+1][r+2] a[c+2][r+1] (indices mod 3)
        const int c1 = (cl + 1) % 3, c2 = (cl + 2) % 3, r1 = (r + 1) % 3, r2 = (r + 2) % 3;
        dAi[q] = (dA[3 * c1 + r1] * dA[3 * c2 + r2] - dA[3 * c1 + r2] * dA[3 * c2 + r1]) * (1.0 / det);
    }
    __builtin_amdgcn_wave_barrier();
    double mine = 0.0;
    if (lane < 9) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) acc += (P[3 * r + k] + (r == k ? 1.0 : 0.0)) * dAi[3 * k + cl];
        An[q] = acc;
        mine = acc - (r == cl ? 1.0 : 0.0);
    }
    __builtin_amdgcn_wave_barrier();
    const double t[3] = {lane_bcast(dpk, 9), lane_bcast(dpk, 10), lane_bcast(dpk, 11)};
    if (lane >= 9 && lane < 12) {
        const int i = lane - 9;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) acc += An[3 * i + k] * t[k];
        mine = P[lane] - acc;
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < 12) P[lane] = mine;
    return true;
}

// One IC-GN decision of wave 0 from the gradient entry bk (lane k, already scaled by bottomf_d)
// and the ZNCC. Returns -1 to continue, else the final status (the same value on every lane);
// c.P is updated.
__device__ int ctl12_step_wave(Ctl12w& c, double bk, double zncc, double bottomf_d, double n_valid, const double* L,
                               const double* H, const double* pattern, double noise_gain, double tol, double dp_tol,
                               int patience, int predictive, double half_scale) {
    const int lane = lane_here();
    const bool own = lane < 12;
    const double norm_abs = sqrt(sum12(bk * bk));
    if (!isfinite(norm_abs)) {
        if (lane == 0) c.zncc = NAN;
        return ST_NAN;
    }
    const double norm_init = c.norm_init < 0.0 ? norm_abs : c.norm_init;
    const double last_dp = c.last_dp, dp_prev = c.dp_prev, best_zncc = c.best_zncc, best_dp = c.best_dp;
    const int stall = c.stall;
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        c.norm_init = norm_init;
        c.zncc = zncc;
    }
    const double norm_rel = norm_init > 1e-300 ? norm_abs / norm_init : 0.0;
    if (norm_rel < tol || norm_abs < ABS_TOL) return ST_CONVERGED;
    bool use_corr = false;
    const double n_full = pattern ? pattern[9 * 12 + 9] : 0.0;
    if (noise_gain > 0.0 && n_full > 0.0 && last_dp < NOISE_CORR_STEP && H) {
        const double om = 1.0 - zncc;
        const double s2 = (om > 0.0 ? om : 0.0) * bottomf_d * bottomf_d / n_valid;
        const double corr = noise_gain * s2 * (n_valid / n_full);
        int ok = 0;
        if (lane == 0) ok = corrected_cholesky12(H, pattern, corr, c.H0, c.L0) ? 1 : 0;
        use_corr = __builtin_amdgcn_readfirstlane(ok) != 0;  // every lane is active again: lane 0's flag
        __builtin_amdgcn_wave_barrier();
    }
    const double* Lf = use_corr ? c.L0 : L;
    // forward sweep L y = -b, then back sweep L^T x = y; lane i reads row i (forward) or column i
    // (back) of the factor one entry per step (rows of lanes >= 12 are clamped and unused)
    const int row = own ? lane : 11;
    const double dii = Lf[row * 13];
    double s = own ? -bk : 0.0, y = 0.0;
#pragma unroll 1
    for (int k = 0; k < 12; ++k) {
        const double Lik = Lf[row * 12 + k];
        const double yk = lane_bcast(s / dii, k);
        if (lane == k) y = yk;
        if (lane > k) s -= Lik * yk;
    }
    double t = y, x = 0.0;
#pragma unroll 1
    for (int k = 11; k >= 0; --k) {
        const double Lki = Lf[k * 12 + row];
        const double xk = lane_bcast(t / dii, k);
        if (lane == k) x = xk;
        if (lane < k) t -= Lki * xk;
    }
    const double v = lane < 9 ? x * half_scale : x;
    const double dp_norm = sqrt(sum12(v * v));
    if (!isfinite(dp_norm)) {
        if (lane == 0) c.zncc = NAN;
        return ST_NAN;
    }
    if (lane == 0) c.last_dp = dp_norm;
    if (dp_norm < dp_tol) return ST_CONVERGED;
    if (predictive && dp_prev > 0.0 && dp_norm < PREDICT_CONTRACTION * dp_prev && dp_norm * dp_norm < dp_tol * dp_prev)
        return compose_warp_wave(c.P, c.work, x) ? ST_CONVERGED : ST_SINGULAR;
    if (lane == 0) c.dp_prev = dp_norm;
    bool improved = zncc > best_zncc + STALL_ZNCC_EPS;
    if (improved) {
        if (lane == 0) c.best_zncc = zncc;
        if (own) c.P_best[lane] = c.P[lane];
    }
    if (dp_norm < best_dp * STALL_STEP_DECAY) {
        if (lane == 0) c.best_dp = dp_norm;
        improved = true;
    }
    if (improved) {
        if (lane == 0) c.stall = 0;
    } else {
        if (lane == 0) c.stall = stall + 1;
        if (patience > 0 && stall + 1 >= patience) {
            if (own) c.P[lane] = c.P_best[lane];
            if (lane == 0) c.zncc = best_zncc;
            return ST_STALLED;
        }
    }
    return compose_warp_wave(c.P, c.work, x) ? -1 : ST_SINGULAR;
}

// Per-node state of a translation-only solve (ADMM subproblem 1, numba_kernels.py:931-1082):
// P[0..8] the fixed displacement gradient, P[9..11] the translation; lane 0 only, in LDS.
struct Ctl3 {
    double P[12], Hinv[9];
    double bf2, best_dn, last_dn, dn_prev, norm_init, zncc;
    int stall;
};

// Loads the node's state; returns -1, or the status that ends the node before any iteration
// (non-finite input, subset outside the volume, singular proximal Hessian: numba_kernels.py:962-982).
__device__ inline int ctl3_init(Ctl3& c, const IcgnArgs& a, int64_t n, bool in_volume) {
    bool finite = true;
#pragma unroll 1
    for (int k = 0; k < 9; ++k) {
        c.P[k] = a.F[9 * n + k];
        if (!isfinite(c.P[k])) finite = false;
    }
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {
        c.P[9 + k] = a.U_old[3 * n + k];
        if (!isfinite(c.P[9 + k])) finite = false;
    }
    c.bf2 = a.bottomf[n] * a.bottomf[n];
    if (c.bf2 < 1e-30) c.bf2 = 1e-30;
    c.best_dn = c.last_dn = 1e300;
    c.dn_prev = c.norm_init = -1.0;
    c.zncc = NAN;
    c.stall = 0;
    if (!finite) return ST_NAN;
    if (!in_volume) return ST_INVALID_SUBSET;
    if (!build_h3(a.H_all + 144 * n, c.bf2, a.mu, 0.0, c.Hinv)) return ST_SINGULAR;
    return -1;
}

// One Gauss-Newton step of the translation with the proximal term mu (u - u_old - v)
// (numba_kernels.py:1000-1075): gradient from the voxel-set moments (C = sum grad, A = sum grad f',
// B = sum grad g' per component; centred f', g' statistics), convergence on the relative or absolute
// gradient norm, noise-corrected Hessian once the steps are small, step-norm and stall rules, and
// the inverse-compositional update u -= (I + F) dt. Returns -1 to continue or the final status.
__device__ inline int ctl3_step(Ctl3& c, const IcgnArgs& a, int64_t n, double n_valid, double mf, double mg,
                                double bf, double bg, double sfg, const double* C, const double* A, const double* B) {
    const double ratio = bf / bg;
    double tb[3], norm_abs = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double bk = (A[k] - mf * C[k]) - ratio * (B[k] - mg * C[k]);
        tb[k] = bk * 2.0 / c.bf2 + a.mu * (c.P[9 + k] - a.U_old[3 * n + k] - a.vdual[3 * n + k]);
        norm_abs += tb[k] * tb[k];
    }
    norm_abs = sqrt(norm_abs);
    if (!isfinite(norm_abs)) {
        c.zncc = NAN;
        return ST_NAN;
    }
    if (c.norm_init < 0.0) c.norm_init = norm_abs;
    const double norm_rel = c.norm_init > 1e-300 ? norm_abs / c.norm_init : 0.0;
    c.zncc = (sfg - n_valid * mf * mg) / (bf * bg);
    if (norm_rel < a.tol || norm_abs < a.mu * 1e-4) return ST_CONVERGED;
    if (a.noise_gain > 0.0 && c.last_dn < NOISE_CORR_STEP) {
        const double om = 1.0 - c.zncc;
        const double s2 = (om > 0.0 ? om : 0.0) * bf * bf / n_valid;
        const double* H = a.H_all + 144 * n;
        if (!build_h3(H, c.bf2, a.mu, a.noise_gain * s2 * n_valid, c.Hinv)) build_h3(H, c.bf2, a.mu, 0.0, c.Hinv);
    }
    double dt[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) dt[i] = -(c.Hinv[3 * i] * tb[0] + c.Hinv[3 * i + 1] * tb[1] + c.Hinv[3 * i + 2] * tb[2]);
    const double dn = sqrt(dt[0] * dt[0] + dt[1] * dt[1] + dt[2] * dt[2]);
    if (!isfinite(dn)) {
        c.zncc = NAN;
        return ST_NAN;
    }
    c.last_dn = dn;
    int ctrl = -1;
    bool apply = true;
    if (dn < a.dp_tol) {
        ctrl = ST_CONVERGED;
        apply = false;
    } else if (a.predictive && c.dn_prev > 0.0 && dn < PREDICT_CONTRACTION * c.dn_prev && dn * dn < a.dp_tol * c.dn_prev) {
        ctrl = ST_CONVERGED;  // the predicted remaining steps are below dp_tol: take this one and stop
    } else {
        c.dn_prev = dn;
        if (dn < c.best_dn * STALL_STEP_DECAY) {
            c.best_dn = dn;
            c.stall = 0;
        } else {
            c.stall += 1;
            if (a.patience > 0 && c.stall >= a.patience) {
                ctrl = ST_STALLED;
                apply = false;
            }
        }
    }
    if (apply) {
        const double* P = c.P;
        const double d0 = dt[0], d1 = dt[1], d2 = dt[2];
        const double u0 = (1.0 + P[0]) * d0 + P[1] * d1 + P[2] * d2;
        const double u1 = P[3] * d0 + (1.0 + P[4]) * d1 + P[5] * d2;
        const double u2 = P[6] * d0 + P[7] * d1 + (1.0 + P[8]) * d2;
        c.P[9] -= u0;
        c.P[10] -= u1;
        c.P[11] -= u2;
    }
    return ctrl;
}

// Warp coefficients broadcast to the sampling lanes: A = I + F (row-major) and the translation,
// relative to the node centre (float32 stays accurate because the magnitudes are small).
__device__ inline void put_warp(float* w, const double* P) {
    w[0] = (float)(1.0 + P[0]);
    w[1] = (float)P[1];
    w[2] = (float)P[2];
    w[3] = (float)P[3];
    w[4] = (float)(1.0 + P[4]);
    w[5] = (float)P[5];
    w[6] = (float)P[6];
    w[7] = (float)P[7];
    w[8] = (float)(1.0 + P[8]);
    w[9] = (float)P[9];
    w[10] = (float)P[10];
    w[11] = (float)P[11];
}

// Mixed-radix walk of the subset voxels v = tid, tid + BLOCK, ... in sampling order.
struct VoxelWalk {
    int ix, iy, iz, v;
    int stx, sty, stz;
    __device__ inline void start(const Subset& s, int v0) {
        v = v0;
        iz = v0 / (s.sy * s.sx);
        const int rem = v0 - iz * s.sy * s.sx;
        iy = rem / s.sx;
        ix = rem - iy * s.sx;
        stz = BLOCK / (s.sy * s.sx);
        const int r2 = BLOCK - stz * s.sy * s.sx;
        sty = r2 / s.sx;
        stx = r2 - sty * s.sx;
    }
    __device__ inline void next(const Subset& s) {
        v += BLOCK;
        ix += stx;
        int cy = 0;
        if (ix >= s.sx) {
            ix -= s.sx;
            cy = 1;
        }
        iy += sty + cy;
        int cz = 0;
        if (iy >= s.sy) {
            iy -= s.sy;
            cz = 1;
        }
        iz += stz + cz;
    }
};

// One subset voxel: sampling-order index v, anchor-relative offsets and linear voxel offset.
struct Vox {
    int v, dx, dy, dz, o;
};

// VoxelWalk that also carries the anchor-relative offsets (dx, dy, dz) and the linear voxel
// offset o = dz * planep + dy * rowp + dx, updated by additions only (no per-voxel multiplies).
struct OffsetWalk {
    int v, ix, iy, dx, dy, dz, o;
    int stx, sty, sdx, sdy, sdz, odx, ody, odz;  // per-step increments (uniform)
    int wx, wy, ow_x, ow_y, oc_y, oc_z;          // wraps and carries (uniform)
    __device__ inline void start(const Subset& s, int v0, int rowp, int planep) {
        v = v0;
        const int iz = v0 / (s.sy * s.sx), rem = v0 - iz * s.sy * s.sx;
        iy = rem / s.sx;
        ix = rem - iy * s.sx;
        dx = -s.hx + s.stride * ix;
        dy = -s.hy + s.stride * iy;
        dz = -s.hz + s.stride * iz;
        o = dz * planep + dy * rowp + dx;
        const int stz = BLOCK / (s.sy * s.sx), r2 = BLOCK - stz * s.sy * s.sx;
        sty = r2 / s.sx;
        stx = r2 - sty * s.sx;
        sdx = s.stride * stx;
        sdy = s.stride * sty;
        sdz = s.stride * stz;
        odx = sdx;
        ody = sdy * rowp;
        odz = sdz * planep;
        wx = s.stride * s.sx;
        wy = s.stride * s.sy;
        ow_x = wx - s.stride * rowp;   // wrap in x carries one row
        ow_y = wy * rowp - s.stride * planep;  // wrap in y carries one plane
        oc_y = s.stride * rowp;
        oc_z = s.stride * planep;
    }
    __device__ inline void next(const Subset& s) {
        v += BLOCK;
        ix += stx;
        dx += sdx;
        o += odx;
        if (ix >= s.sx) {
            ix -= s.sx;
            dx -= wx;
            o -= ow_x;
            ++iy;
            dy += s.stride;
        }
        iy += sty;
        dy += sdy;
        o += ody;
        if (iy >= s.sy) {
            iy -= s.sy;
            dy -= wy;
            o -= ow_y;
            dz += s.stride;
        }
        dz += sdz;
        o += odz;
    }
};

__device__ inline bool subset_in_volume(const Volumes& V, const Subset& s, int64_t x0, int64_t y0, int64_t z0) {
    return x0 - s.hx >= 0 && x0 + s.hx < V.nx && y0 - s.hy >= 0 && y0 + s.hy < V.ny && z0 - s.hz >= 0 &&
           z0 + s.hz < V.nz;
}

__device__ inline const uint8_t* keep_row_of(const IcgnArgs& a, int64_t n) {
    if (!a.split_index || !a.split_keep) return nullptr;
    const int64_t r = a.split_index[n];
    return r >= 0 ? a.split_keep + r * a.keep_bytes : nullptr;
}

// ---------------------------------------------------------------------------- plain kernel
template <int MODE, bool STORED>
__global__ __launch_bounds__(BLOCK) void k_icgn12_plain(IcgnArgs a, const int64_t* __restrict__ list,
                                                        const int64_t* __restrict__ count, int64_t list_off,
                                                        float* __restrict__ gbuf) {
    const Volumes& V = a.V;
    const Subset& s = a.sub;
    const int tid = threadIdx.x;
    if (list_off + (int64_t)blockIdx.x >= *count) return;
    const int64_t n = list[list_off + blockIdx.x];
    float* gb = gbuf + (int64_t)blockIdx.x * s.S;
    const int64_t x0 = a.coords[3 * n], y0 = a.coords[3 * n + 1], z0 = a.coords[3 * n + 2];
    const uint8_t* keep = keep_row_of(a, n);

    __shared__ float sw[12];
    __shared__ double st[8];
    __shared__ int ctrl;
    __shared__ double red[NWAVES * 13];
    __shared__ double res[13];
    __shared__ Ctl12 c;

    if (tid == 0) {
        bool finite = true;
        ctl12_init(c, a.P + 12 * n);
#pragma unroll 1
        for (int k = 0; k < 12; ++k)
            if (!isfinite(c.P[k])) finite = false;
        ctrl = -1;
        if (!finite)
            ctrl = ST_NAN;
        else if (!keep && !subset_in_volume(V, s, x0, y0, z0))
            ctrl = ST_INVALID_SUBSET;  // a valid flag on a subset that leaves the volume: refuse it
    }
    __syncthreads();
    const double half_scale = (double)max(s.hx, max(s.hy, s.hz));
    const Domain dom = make_domain(V, (int)x0, (int)y0, (int)z0);
    const float* gbase = V.g + lin(V, z0, y0, x0);
    const int64_t obase = lin(V, z0, y0, x0);
    const int rowp = V.nx, planep = V.nx * V.ny;
    int it_done = 0;
    int status = ST_MAX_ITER;
    for (int it = 1; it <= a.max_iter; ++it) {
        if (ctrl >= 0) break;
        it_done = it;
        if (tid == 0) put_warp(sw, c.P);
        __syncthreads();
        // ---------------- pass 1: warp, sample, statistics (cuda_kernels.py:577-674)
        float p1[7] = {0, 0, 0, 0, 0, 0, 0};  // s1, s2, s1f, s2f, nv, nref, oob
        VoxelWalk w;
        w.start(s, tid);
        for (; w.v < s.S; w.next(s)) {
            const int dx = -s.hx + s.stride * w.ix, dy = -s.hy + s.stride * w.iy, dz = -s.hz + s.stride * w.iz;
            const int64_t o = obase + (int64_t)dz * planep + (int64_t)dy * rowp + dx;
            bool k;
            if (keep) {
                k = kept_bit(keep, w.v);
                const int64_t xx = x0 + dx, yy = y0 + dy, zz = z0 + dz;
                k = k && xx >= 0 && xx < V.nx && yy >= 0 && yy < V.ny && zz >= 0 && zz < V.nz;
            } else {
                k = !V.mask || V.mask[o] != 0;
            }
            if (!k) {
                gb[w.v] = NAN;
                continue;
            }
            p1[5] += 1.0f;
            const float X = (float)dx, Y = (float)dy, Z = (float)dz;
            const float xr = sw[0] * X + sw[1] * Y + sw[2] * Z + sw[9];
            const float yr = sw[3] * X + sw[4] * Y + sw[5] * Z + sw[10];
            const float zr = sw[6] * X + sw[7] * Y + sw[8] * Z + sw[11];
            if (!inside(dom, xr, yr, zr)) {
                p1[6] = 1.0f;
                gb[w.v] = NAN;
                continue;
            }
            const float val = sample_rel<MODE>(gbase, rowp, planep, xr, yr, zr, dom.kx, dom.ky, dom.kz);
            gb[w.v] = val;
            if (val != val) continue;
            const float fv = V.f[o];
            p1[4] += 1.0f;
            p1[0] += val;
            p1[1] += val * val;
            p1[2] += fv;
            p1[3] += fv * fv;
        }
        block_sum<7, NWAVES>(p1, res, red);
        const double* r1 = res;
        if (tid == 0) {
            const double n_valid = r1[4], n_ref = r1[5];
            if (r1[6] > 0.0) {
                ctrl = ST_OUT_OF_BOUNDS;
            } else if (n_valid < MIN_SUBSET_VOXELS || n_valid < MIN_VALID_FRACTION * n_ref) {
                ctrl = ST_INVALID_SUBSET;
            } else {
                const double mg = r1[0] / n_valid;
                double ssg = r1[1] - n_valid * mg * mg;
                ssg = ssg < 0.0 ? 0.0 : ssg;
                const double mf = r1[2] / n_valid;
                double ssf = r1[3] - n_valid * mf * mf;
                ssf = ssf < 0.0 ? 0.0 : ssf;
                const double bg = sqrt(ssg > 1e-30 ? ssg : 1e-30), bf = sqrt(ssf > 1e-30 ? ssf : 1e-30);
                st[0] = mf;
                st[1] = bf;
                st[2] = mg;
                st[3] = bg;
                st[4] = 1.0 / bf;
                st[5] = 1.0 / bg;
                st[6] = n_valid;
            }
        }
        __syncthreads();
        if (ctrl >= 0) {
            if (tid == 0) c.zncc = NAN;
            break;
        }
        // ---------------- pass 2: residual, 12-vector, ZNCC numerator (cuda_kernels.py:679-747)
        const float mf32 = (float)st[0], mg32 = (float)st[2], ibf32 = (float)st[4], ibg32 = (float)st[5];
        float p2[13] = {0};
        w.start(s, tid);
        for (; w.v < s.S; w.next(s)) {
            const float gv = gb[w.v];
            if (gv != gv) continue;
            const int dx = -s.hx + s.stride * w.ix, dy = -s.hy + s.stride * w.iy, dz = -s.hz + s.stride * w.iz;
            const int64_t o = obase + (int64_t)dz * planep + (int64_t)dy * rowp + dx;
            const float fd = V.f[o] - mf32, gd = gv - mg32;
            p2[12] += fd * gd;
            const float res = fd * ibf32 - gd * ibg32;
            float dfx, dfy, dfz;
            grad_at<STORED>(V, o, (int)z0 + dz, (int)y0 + dy, (int)x0 + dx, dfx, dfy, dfz);
            const float gxv = dfx * res, gyv = dfy * res, gzv = dfz * res;
            const float X = (float)dx, Y = (float)dy, Z = (float)dz;
            p2[0] += gxv * X;
            p2[1] += gxv * Y;
            p2[2] += gxv * Z;
            p2[3] += gyv * X;
            p2[4] += gyv * Y;
            p2[5] += gyv * Z;
            p2[6] += gzv * X;
            p2[7] += gzv * Y;
            p2[8] += gzv * Z;
            p2[9] += gxv;
            p2[10] += gyv;
            p2[11] += gzv;
        }
        block_sum<13, NWAVES>(p2, res, red);
        const double* r2 = res;
        if (tid == 0) {
#pragma unroll 1
            for (int k = 0; k < 12; ++k) c.b[k] = r2[k] * st[1];
            const double zncc = r2[12] * st[4] * st[5];
            const double* Hn = a.H_all ? a.H_all + 144 * n : nullptr;
            ctrl = ctl12_step(c, zncc, st[1], st[6], a.L_all + 144 * n, Hn, a.pattern, a.noise_gain, a.tol, a.dp_tol,
                              a.patience, a.predictive, half_scale);
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (ctrl >= 0) status = ctrl;
        for (int k = 0; k < 12; ++k) a.P[12 * n + k] = c.P[k];
        a.n_iter[n] = it_done;
        a.status[n] = (int8_t)status;
        // the reference GPU kernel reports a ZNCC only for these outcomes (cuda_kernels.py:833-835)
        a.zncc[n] = (status == ST_CONVERGED || status == ST_STALLED || status == ST_MAX_ITER) ? c.zncc : NAN;
    }
}

// ---------------------------------------------------------------------------- fused kernel
// The ZNSSD gradient of one iteration needs, over the iteration's valid voxel set V,
//   b_k = sum SD_k (f - m_f) - (b_f / b_g) sum SD_k (g - m_g)            (numba_kernels.py:705-729)
// with SD_k = grad_c(f) q_j, q = (X, Y, Z, 1). Every term is a raw moment: the reference-side ones
// (n, sum f', sum f'^2, sum grad_c q_j, sum grad_c q_j f') depend only on V, which is the node's
// kept set minus the voxels whose deformed sample is NaN (a sample outside the domain ends the
// node). So they are summed once per call over the kept set (k_ref_moments), and an iteration
// accumulates only the deformed-side moments; when some samples are NaN, their voxels are
// flagged in an LDS bitmask and a second pass sums their reference-side moments, which are
// subtracted. f' = f - pf and g' = g - pg are pilot-shifted (pf: precomputed reference mean,
// pg: the previous iteration's deformed mean) so that the float32 lane sums stay centred.
constexpr int NR = 27;  // 0 n, 1 sum f', 2 sum f'^2, 3 + 4c + j: sum grad_c q_j, 15 + 4c + j: sum grad_c q_j f'
constexpr int NG = 18;  // 0 n_valid, 1 oob, 2 n_nan, 3 sum g', 4 sum g'^2, 5 sum f'g', 6 + 4c + j: sum grad_c q_j g'
constexpr int NG3 = 9;  // translation-only: 0..5 as NG, 6 + c: sum grad_c g'
constexpr int MOM_STRIDE = 32;  // doubles per node in the moments workspace
constexpr int BRICK_XY = 40;    // LDS brick rows and columns: one box for subsets up to 35 voxels wide
// LDS budget of one fused workgroup, static and dynamic together: three per CU. gfx950 allocates a
// workgroup's LDS in 1280-byte granules (160 KB in 128), so three workgroups hold 42 granules each.
constexpr int LDS_PER_WG = 42 * 1280;

// XCD-aware slot of workgroup b over the M valid nodes: workgroup b runs on XCD b % 8, and the
// node list is dealt to the XCDs in chunks of XCD_CHUNK consecutive nodes (two 4 x 4 x 4 tiles of
// the tiled list at step 8), chunk q to XCD q % 8, so the nodes an XCD runs at once are neighbours
// that share subset voxels in its L2, while every XCD samples the whole volume and ends with the
// same share of the work (one contiguous eighth per XCD left the XCDs up to 12 % apart). Returns
// -1 for the workgroups beyond M (the grid, xcd_grid, covers every node).
constexpr int64_t XCD_CHUNK = 128;
__device__ inline int64_t xcd_slot(int64_t b, int64_t M, int xcd = 1) {
    if (!xcd) return b < M ? b : -1;
    const int64_t k = b & 7, j = b >> 3;
    const int64_t slot = ((j / XCD_CHUNK) * 8 + k) * XCD_CHUNK + (j % XCD_CHUNK);
    return slot < M ? slot : -1;
}

// list slot of the j-th node XCD k takes: its chunks are q = k, k + 8, k + 16, ...
__device__ inline int64_t chunk_slot(int k, int64_t j) {
    return ((j / XCD_CHUNK) * 8 + k) * XCD_CHUNK + (j % XCD_CHUNK);
}
constexpr int NODE_COUNTERS = 8;  // workspace header: count[0] valid nodes, count[8 + k] XCD k's next node

// workgroups of a launch over N nodes: N rounded up to whole rounds of 8 chunks
inline int64_t xcd_grid(int64_t N) { return (N + 8 * XCD_CHUNK - 1) / (8 * XCD_CHUNK) * (8 * XCD_CHUNK); }

// Kept-voxel test shared by every pass: the keep row (inside the volume by construction, checked
// anyway) or the reference mask.
__device__ inline bool voxel_kept(const Volumes& V, const uint8_t* keep, int v, int64_t o, int64_t xx, int64_t yy,
                                  int64_t zz) {
    if (keep) return kept_bit(keep, v) && xx >= 0 && xx < V.nx && yy >= 0 && yy < V.ny && zz >= 0 && zz < V.nz;
    return !V.mask || V.mask[o] != 0;
}

// Reference-side moments over the voxels selected by `pick(v)` (kept voxels, or the NaN-flagged
// ones of an iteration), reduced to `out` (NR doubles in LDS).
template <bool STORED, typename Pick>
__device__ inline void ref_moments(const IcgnArgs& a, const uint8_t* keep, int64_t x0, int64_t y0, int64_t z0,
                                   float pf, Pick pick, double* out, double* red) {
    const Volumes& V = a.V;
    const Subset& s = a.sub;
    const int64_t obase = lin(V, z0, y0, x0);
    const int rowp = V.nx, planep = V.nx * V.ny;
    float m[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) m[q] = 0.0f;
    VoxelWalk w;
    w.start(s, threadIdx.x);
    for (; w.v < s.S; w.next(s)) {
        const int dx = -s.hx + s.stride * w.ix, dy = -s.hy + s.stride * w.iy, dz = -s.hz + s.stride * w.iz;
        const int64_t o = obase + (int64_t)dz * planep + (int64_t)dy * rowp + dx;
        if (!pick(w.v, o, x0 + dx, y0 + dy, z0 + dz)) continue;
        const float fv = V.f[o] - pf;
        float gr[3];
        grad_at<STORED>(V, o, (int)z0 + dz, (int)y0 + dy, (int)x0 + dx, gr[0], gr[1], gr[2]);
        const float q[4] = {(float)dx, (float)dy, (float)dz, 1.0f};
        m[0] += 1.0f;
        m[1] += fv;
        m[2] += fv * fv;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            const float tf = gr[cc] * fv;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                m[3 + 4 * cc + j] += gr[cc] * q[j];
                m[15 + 4 * cc + j] += tf * q[j];
            }
        }
    }
    block_sum<NR, NWAVES>(m, out, red);
}

// The same moments over every kept voxel, summed along voxel columns in y: a thread walks half of
// one column (its X and Z fixed), keeping per-column sums of grad_c and grad_c f' that it folds into
// the X, Z and 1 moments once at the column's end; the Y moments and 1, f', f'^2 are summed per
// voxel. About two thirds of the per-voxel arithmetic of the per-voxel form and no index decode;
// neighbouring lanes take neighbouring x (coalesced loads), and a column stays inside its plane.
template <bool STORED, bool SEL>
__device__ inline void ref_moments_columns(const IcgnArgs& a, const uint8_t* keep, int64_t x0, int64_t y0, int64_t z0,
                                           float pf, double* out, double* red) {
    const Volumes& V = a.V;
    const Subset& s = a.sub;
    const int64_t obase = lin(V, z0, y0, x0);
    const int64_t rowp = V.nx, planep = (int64_t)V.nx * V.ny;
    float m[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) m[q] = 0.0f;
    const int ncol = s.sx * s.sz, yh = (s.sy + 1) >> 1;
    for (int t = threadIdx.x; t < 2 * ncol; t += BLOCK) {
        const int upper = t >= ncol ? 1 : 0, col = t - upper * ncol;  // first every lower half, then every upper
        const int iz = col / s.sx, ix = col - iz * s.sx;
        const int iy0 = upper ? yh : 0, iy1 = upper ? s.sy : yh;
        const int dx = -s.hx + s.stride * ix, dz = -s.hz + s.stride * iz;
        const int64_t ostep = s.stride * rowp;
        int64_t o = obase + (int64_t)dz * planep + (int64_t)(-s.hy + s.stride * iy0) * rowp + dx;
        // per-column sums of grad_c and grad_c f' (folded with X and Z below); the rest goes straight
        // into the thread's moments
        float A[3] = {0.0f, 0.0f, 0.0f}, B[3] = {0.0f, 0.0f, 0.0f};
        // without a selection the body is straight-line code, unrolled so that the next voxel's loads
        // are in flight while this one is summed
#pragma unroll(SEL ? 1 : 2)
        for (int iy = iy0; iy < iy1; ++iy, o += ostep) {
            const int dy = -s.hy + s.stride * iy;
            if constexpr (SEL)
                if (!voxel_kept(V, keep, (iz * s.sy + iy) * s.sx + ix, o, x0 + dx, y0 + dy, z0 + dz)) continue;
            const float fv = V.f[o] - pf, Y = (float)dy;
            float gr[3];
            grad_at<STORED>(V, o, (int)z0 + dz, (int)y0 + dy, (int)x0 + dx, gr[0], gr[1], gr[2]);
            m[0] += 1.0f;
            m[1] += fv;
            m[2] += fv * fv;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const float tf = gr[cc] * fv;
                A[cc] += gr[cc];
                B[cc] += tf;
                m[4 + 4 * cc] += gr[cc] * Y;
                m[16 + 4 * cc] += tf * Y;
            }
        }
        const float X = (float)dx, Z = (float)dz;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            m[3 + 4 * cc] += A[cc] * X;
            m[5 + 4 * cc] += A[cc] * Z;
            m[6 + 4 * cc] += A[cc];
            m[15 + 4 * cc] += B[cc] * X;
            m[17 + 4 * cc] += B[cc] * Z;
            m[18 + 4 * cc] += B[cc];
        }
    }
    block_sum<NR, NWAVES>(m, out, red);
}

// skip: non-null when the separable pass below may have taken the nodes without keep rows (its
// flag, set on the device, says whether it did)
template <bool STORED, bool SEL>
__global__ __launch_bounds__(BLOCK) void k_ref_moments(IcgnArgs a, const int64_t* __restrict__ list,
                                                       const int64_t* __restrict__ count, double* __restrict__ mom,
                                                       const int* __restrict__ skip) {
    const int64_t slot = xcd_slot(blockIdx.x, *count, a.xcd);
    if (slot < 0) return;
    const int64_t n = list[slot];
    const int64_t x0 = a.coords[3 * n], y0 = a.coords[3 * n + 1], z0 = a.coords[3 * n + 2];
    const uint8_t* keep = keep_row_of(a, n);
    if (skip && *skip && !keep) return;  // taken by the separable pass
    __shared__ double red[NWAVES * NR];
    __shared__ double res[NR];
    if (!keep && !subset_in_volume(a.V, a.sub, x0, y0, z0)) return;  // refused by the iteration kernel
    double pfd = a.meanf ? a.meanf[n] : 0.0;
    const float pf = isfinite(pfd) ? (float)pfd : 0.0f;
    ref_moments_columns<STORED, SEL>(a, keep, x0, y0, z0, pf, res, red);
    for (int q = threadIdx.x; q < NR; q += BLOCK) mom[n * MOM_STRIDE + q] = res[q];
}

// ---------------------------------------------------------------------------- separable reference moments
// For a node without keep rows (stride 1), the reference moments are box sums over its subset cube
// of per-voxel terms weighted by the reference mask w (1 everywhere without one): w, w f, w f^2,
// w grad_c, w grad_c f, and the grad terms times the voxel's offset X, Y or Z from the node, with
// the pilot pf folded in afterwards (sum w f' = sum w f - n pf, ...). Nodes on a lattice share
// most of these sums: the x-sums of one voxel row serve every node whose centre has that x and
// whose subset holds the row, and so on along y and z. So the sums are taken axis by axis at the
// distinct node coordinates only:
//   k_rs_xy: per plane z, distinct centre x (cx) and distinct centre y (cy), the 20 sums over the
//            sx x sy rectangle around (cx, cy): x-row sums in float (four lanes a row, a quarter
//            each, combined), added row by row in double with the row's offset Y from cy;
//   k_rs_z:  per node, its sz planes' rectangle sums, with the Z weights, in the layout of
//            k_ref_moments.
// A voxel is then read by about sx / (x spacing) lanes for each work item whose rows hold it (four
// consecutive windows per item: about 1.8 items at a lattice step of hy / 2) instead of once per
// node whose subset holds it. Every sum is formed from the node's own subset voxels in a fixed order
// (quarter rows, rows, planes), so a node's moments do not depend on the other nodes of the call:
// a volume cut into boxes gives the same bits as the whole. The distinct coordinates come from
// marks over the node list (k_rs_mark, k_rs_index). Rectangle sums:
// R[((p * RS_Q + q) * ncy + cyi) * ncx + cxi]
// for plane p = z - zbase, with q = 0 f, 1 f^2, 2 + c grad_c, 5 + c grad_c X, 8 + c grad_c f,
// 11 + c grad_c f X, 14 + c grad_c Y, 17 + c grad_c f Y, 20 the kept count (with a mask only), each
// weighted by w. Nodes with keep rows take the per-node pass.
constexpr int RS_Q = 21;
constexpr int RS_TILE = 16;  // centres x per work item (four lanes each)
constexpr int RS_KC = 8;     // voxels of a quarter row loaded at once
constexpr int RS_G = 4;      // consecutive centres y per work item: their windows share rows
enum : int { RS_NCX, RS_NCY, RS_ZLO, RS_ZHI, RS_FITS, RS_NPLANES, RS_ZBASE, RS_HDR = 64 };
enum : int { TUNE_ROW_MOMENTS = 32 };

// Device buffers of the separable pass, carved from the workspace after the node list and moments.
struct RowMoments {
    int* hdr;   // RS_HDR ints: the counts above; set on the device
    int* xidx;  // nx: mark of each x, then the index of the centre x it is (-1: none)
    int* yidx;  // ny
    int* cxs;   // the distinct centre x, ascending
    int* cys;   // the distinct centre y
    double* R;  // rectangle sums
    int64_t r_doubles;  // room for R
};

__host__ __device__ inline size_t row_moments_fixed_bytes(int64_t ny, int64_t nx) {
    return sizeof(int) * RS_HDR + (sizeof(int) * 2 * (size_t)(nx + ny) + 255) / 256 * 256;
}

__global__ void k_rs_clear(RowMoments r, int nx, int ny) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nx) r.xidx[i] = 0;
    if (i < ny) r.yidx[i] = 0;
    if (i < RS_HDR) r.hdr[i] = i == RS_ZLO ? INT_MAX : (i == RS_ZHI ? -1 : 0);
}

// Marks the centres of the listed nodes (every node 0..n_all-1 when there is no list) without keep
// rows whose subsets lie inside the volume (the others take the per-node pass, or are refused by
// the iteration kernel and need no moments).
__global__ void k_rs_mark(IcgnArgs a, const int64_t* __restrict__ list, const int64_t* __restrict__ count,
                          int64_t n_all, RowMoments r) {
    const int64_t M = list ? *count : n_all;
    int zlo = INT_MAX, zhi = -1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = list ? list[i] : i;
        const int64_t x0 = a.coords[3 * n], y0 = a.coords[3 * n + 1], z0 = a.coords[3 * n + 2];
        if (keep_row_of(a, n) || !subset_in_volume(a.V, a.sub, x0, y0, z0)) continue;
        r.xidx[x0] = 1;
        r.yidx[y0] = 1;
        zlo = min(zlo, (int)z0);
        zhi = max(zhi, (int)z0);
    }
    // the z range: one atomic pair per wave
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        zlo = min(zlo, __shfl_xor(zlo, off, 64));
        zhi = max(zhi, __shfl_xor(zhi, off, 64));
    }
    if ((threadIdx.x & 63) == 0 && zhi >= 0) {
        atomicMin(&r.hdr[RS_ZLO], zlo);
        atomicMax(&r.hdr[RS_ZHI], zhi);
    }
}

// marks -> indices and the ascending list of marked positions; returns the count (one block of
// 1024 threads, every thread calls it)
__device__ inline int rs_index_axis(int* idx, int* out, int n, int* wsum) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int base = 0;
    for (int c0 = 0; c0 < n; c0 += 1024) {
        const int i = c0 + threadIdx.x;
        const bool m = i < n && idx[i] != 0;
        const unsigned long long b = __ballot(m);
        if (lane == 0) wsum[wave] = __popcll(b);
        __syncthreads();
        int before = base;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        int total = base;
        for (int w = 0; w < 16; ++w) total += wsum[w];
        const int at = before + (int)__popcll(b & ((1ull << lane) - 1));
        if (i < n) {
            idx[i] = m ? at : -1;
            if (m) out[at] = i;
        }
        base = total;
        __syncthreads();  // wsum is rewritten by the next chunk
    }
    return base;
}

// Distinct centres and whether the rectangle sums (nq per plane and centre pair) fit the room given
// (else every node takes the per-node pass).
__global__ __launch_bounds__(1024) void k_rs_index(RowMoments r, int nx, int ny, int hz, int nq) {
    __shared__ int wsum[16];
    const int ncx = rs_index_axis(r.xidx, r.cxs, nx, wsum);
    const int ncy = rs_index_axis(r.yidx, r.cys, ny, wsum);
    if (threadIdx.x == 0) {
        const int zlo = r.hdr[RS_ZLO], zhi = r.hdr[RS_ZHI];
        const int64_t planes = zhi >= zlo ? (int64_t)(zhi - zlo) + 2 * hz + 1 : 0;
        r.hdr[RS_NCX] = ncx;
        r.hdr[RS_NCY] = ncy;
        r.hdr[RS_NPLANES] = (int)planes;
        r.hdr[RS_ZBASE] = zlo - hz;
        r.hdr[RS_FITS] = planes > 0 && ncx > 0 && ncy > 0 && planes * nq * ncy * (int64_t)ncx <= r.r_doubles;
    }
}

// sum over the four lanes of a quad (DPP quad permutations: lanes xor 1, then xor 2)
__device__ inline float quad_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
    return v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
}

// Rectangle sums a quad lane keeps, five per lane and the kept count on part 3 with a mask:
// RS_DST[part][i] is the sum q of slot i, taken from the row sum RS_SRC[part][i] (times y for slots
// 1 and 3 of parts 0-2: the y-weighted grad sums sit beside their unweighted ones, 0 and 2, which
// the window's y offsets need); -1: no slot.
constexpr int RS_SLOTS = 6;
__constant__ const int8_t RS_DST[4][RS_SLOTS] = {
    {2, 14, 8, 17, 0, -1}, {3, 15, 9, 18, 1, -1}, {4, 16, 10, 19, 5, -1}, {6, 7, 11, 12, 13, 20}};
constexpr int8_t RS_SRC[4][RS_SLOTS] = {
    {2, 2, 8, 8, 0, 14}, {3, 3, 9, 9, 1, 14}, {4, 4, 10, 10, 5, 14}, {6, 7, 11, 12, 13, 14}};

// One wave per work item (plane, centre y, RS_TILE centres x); lane 4 j + k takes centre x j of the
// tile and the k-th quarter of each of the window's sy rows. The quad's row sums are combined every
// row and each lane adds its slots of the rectangle sums. MASKED: every term weighted by the
// reference mask, and the kept count summed. Stored gradients only.
template <bool MASKED>
__global__ __launch_bounds__(BLOCK, 3) void k_rs_xy(Volumes V, Subset s, RowMoments r) {
    if (!r.hdr[RS_FITS]) return;
    constexpr int NM = MASKED ? 15 : 14;  // row sums: the 14 terms, and the kept count
    const int ncx = r.hdr[RS_NCX], ncy = r.hdr[RS_NCY], zbase = r.hdr[RS_ZBASE];
    const int ntile = (ncx + RS_TILE - 1) / RS_TILE;
    const int lane = threadIdx.x & 63, j = lane >> 2, part = lane & 3;
    // a quad's lanes split a row: lane part takes base = sx / 4 voxels from k0, and one more (at kx)
    // when part < sx % 4; a lane without one reads a valid voxel there and counts it zero
    const int base = s.sx >> 2, extra = s.sx & 3, k0 = part * base + min(part, extra);
    const int kx = min(k0 + base, s.sx - 1);
    const bool has_extra = part < extra;
    const int64_t plane = (int64_t)V.nx * V.ny, cstride = (int64_t)ncy * ncx;
    const bool ywt = part < 3;  // slots 1 and 3 are y-weighted
    // slot sources picked by 0/1 weights (a product with 0 and three exact additions of 0: the
    // lane's own row sum, bit for bit, without per-use compares)
    const float w0 = part == 0 ? 1.0f : 0.0f, w1 = part == 1 ? 1.0f : 0.0f, w2 = part == 2 ? 1.0f : 0.0f,
                w3 = part == 3 ? 1.0f : 0.0f;
    constexpr int NSL = MASKED ? RS_SLOTS : RS_SLOTS - 1;
    // the work item is the same on every lane of a wave: in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // XCD-aware items: workgroup b runs on XCD b % 8 (the grid is a multiple of 8) and XCD k takes
    // the k-th eighth of the (plane, group of RS_G centres y) pairs in order, so the windows an XCD
    // sums at once are neighbours whose rows its L2 holds
    const int ngroup = (ncy + RS_G - 1) / RS_G;
    const int64_t pairs = (int64_t)r.hdr[RS_NPLANES] * ngroup, xcd = blockIdx.x & 7;
    const int64_t lo = pairs * xcd / 8, items = (pairs * (xcd + 1) / 8 - lo) * ntile;
    for (int64_t it = (blockIdx.x >> 3) * (int64_t)NWAVES + wave; it < items;
         it += (int64_t)(gridDim.x >> 3) * NWAVES) {
        const int64_t pair = lo + it / ntile;
        const int tile = (int)(it - (it / ntile) * ntile), p = (int)(pair / ngroup);
        const int c0 = (int)(pair - (int64_t)p * ngroup) * RS_G, gn = min(RS_G, ncy - c0);
        const int cxi = tile * RS_TILE + j;
        const bool act = cxi < ncx;
        const int cx = act ? r.cxs[cxi] : r.cxs[0];
        int cyg[RS_G];
#pragma unroll
        for (int g = 0; g < RS_G; ++g) cyg[g] = r.cys[c0 + min(g, gn - 1)];
        const int64_t pbase = (int64_t)(zbase + p) * plane;
        const unsigned xoff = (unsigned)(cx - s.hx);  // the subset row's first voxel
        double P[RS_G][NSL];
#pragma unroll
        for (int g = 0; g < RS_G; ++g)
#pragma unroll
            for (int i = 0; i < NSL; ++i) P[g][i] = 0.0;
        for (int y = cyg[0] - s.hy; y <= cyg[gn - 1] + s.hy; ++y) {
            // the row's base is the same on every lane; lanes differ by 32-bit byte offsets
            const int64_t orow = pbase + (int64_t)y * V.nx;
            const char *fr = (const char*)(V.f + orow), *g0r = (const char*)(V.gx + orow),
                       *g1r = (const char*)(V.gy + orow), *g2r = (const char*)(V.gz + orow);
            float m[NM];
#pragma unroll
            for (int q = 0; q < NM; ++q) m[q] = 0.0f;
            auto add = [&](bool in, float fv, float g0, float g1, float g2, int k) {
                const float f1 = in ? fv : 0.0f, X = (float)(k - s.hx);
                const float g[3] = {in ? g0 : 0.0f, in ? g1 : 0.0f, in ? g2 : 0.0f};
                m[0] += f1;
                m[1] = __fmaf_rn(f1, f1, m[1]);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float tf = g[c] * f1;
                    m[2 + c] += g[c];
                    m[5 + c] = __fmaf_rn(g[c], X, m[5 + c]);
                    m[8 + c] += tf;
                    m[11 + c] = __fmaf_rn(tf, X, m[11 + c]);
                }
                if constexpr (MASKED) m[NM - 1] += in ? 1.0f : 0.0f;
            };
            // RS_KC voxels at a time, every load issued before the first is used; the loads of the
            // last, partial chunk stay on the lane's last voxel and count zero
            const uint8_t* mr = MASKED ? V.mask + orow : nullptr;
            for (int kc = 0; kc < max(base, 1); kc += RS_KC) {
                float fv[RS_KC + 1], gv[3][RS_KC + 1];
                uint8_t mk[RS_KC + 1];
#pragma unroll
                for (int t = 0; t <= RS_KC; ++t) {  // t = RS_KC: the extra voxel, fetched with the first chunk
                    const int k = t < RS_KC ? min(k0 + max(min(kc + t, base - 1), 0), s.sx - 1) : kx;
                    const unsigned e = xoff + (unsigned)k, b = 4u * e;
                    fv[t] = *(const float*)(fr + b);
                    gv[0][t] = *(const float*)(g0r + b);
                    gv[1][t] = *(const float*)(g1r + b);
                    gv[2][t] = *(const float*)(g2r + b);
                    if constexpr (MASKED) mk[t] = mr[e];
                }
                auto chunk = [&](auto full) {  // full: every voxel of the chunk is the lane's
#pragma unroll
                    for (int t = 0; t < RS_KC; ++t) {
                        bool in = decltype(full)::value || kc + t < base;
                        if constexpr (MASKED) in = in && mk[t] != 0;
                        add(in, fv[t], gv[0][t], gv[1][t], gv[2][t], k0 + kc + t);
                    }
                };
                if (kc + RS_KC <= base) chunk(std::true_type{});  // uniform
                else chunk(std::false_type{});
                if (kc == 0) {
                    bool in = has_extra;
                    if constexpr (MASKED) in = in && mk[RS_KC] != 0;
                    add(in, fv[RS_KC], gv[0][RS_KC], gv[1][RS_KC], gv[2][RS_KC], kx);
                }
            }
#pragma unroll
            for (int q = 0; q < NM; ++q) m[q] = quad_sum(m[q]);
            double v[NSL];
#pragma unroll
            for (int i = 0; i < NSL; ++i)
                v[i] = (double)__fmaf_rn(m[RS_SRC[3][i]], w3, __fmaf_rn(m[RS_SRC[2][i]], w2,
                                         __fmaf_rn(m[RS_SRC[1][i]], w1, m[RS_SRC[0][i]] * w0)));
            // the row into every window of the group that holds it (each window's rows in order)
#pragma unroll
            for (int g = 0; g < RS_G; ++g) {
                const int dy = y - cyg[g];
                if (g < gn && dy >= -s.hy && dy <= s.hy) {  // uniform
                    const double Y = (double)dy;
#pragma unroll
                    for (int i = 0; i < NSL; ++i) P[g][i] = fma((i == 1 || i == 3) && ywt ? Y : 1.0, v[i], P[g][i]);
                }
            }
        }
        if (act) {
#pragma unroll
            for (int g = 0; g < RS_G; ++g) {
                if (g >= gn) break;
                double* Rc = r.R + (int64_t)p * RS_Q * cstride + (int64_t)(c0 + g) * ncx + cxi;
#pragma unroll
                for (int i = 0; i < NSL; ++i)
                    if (RS_DST[part][i] >= 0) Rc[(int64_t)RS_DST[part][i] * cstride] = P[g][i];
            }
        }
    }
}

// Per node without keep rows: the sz planes' rectangle sums with the Z weights, the pilot folded
// in, written where k_ref_moments writes (same layout, double sums throughout).
__global__ __launch_bounds__(BLOCK) void k_rs_z(IcgnArgs a, const int64_t* __restrict__ list,
                                                const int64_t* __restrict__ count, RowMoments r,
                                                double* __restrict__ mom) {
    if (!r.hdr[RS_FITS]) return;
    const int64_t M = *count;
    const int ncx = r.hdr[RS_NCX], ncy = r.hdr[RS_NCY], zbase = r.hdr[RS_ZBASE];
    const int64_t cstride = (int64_t)ncy * ncx;
    const Subset& s = a.sub;
    const bool masked = a.V.mask != nullptr;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = list[i];
        const int64_t x0 = a.coords[3 * n], y0 = a.coords[3 * n + 1], z0 = a.coords[3 * n + 2];
        if (keep_row_of(a, n) || !subset_in_volume(a.V, s, x0, y0, z0)) continue;
        const double* Rn = r.R + (int64_t)r.yidx[y0] * ncx + r.xidx[x0];
        double S[RS_Q], SZ[6];
#pragma unroll
        for (int q = 0; q < RS_Q; ++q) S[q] = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) SZ[q] = 0.0;
        for (int dz = -s.hz; dz <= s.hz; ++dz) {
            const double* Rz = Rn + (int64_t)(z0 + dz - zbase) * RS_Q * cstride;
            const double Z = (double)dz;
#pragma unroll
            for (int q = 0; q < RS_Q - 1; ++q) {
                const double v = Rz[q * cstride];
                S[q] += v;
                if (q >= 2 && q < 5) SZ[q - 2] = fma(Z, v, SZ[q - 2]);
                if (q >= 8 && q < 11) SZ[q - 5] = fma(Z, v, SZ[q - 5]);
            }
            if (masked) S[RS_Q - 1] += Rz[(RS_Q - 1) * cstride];
        }
        const double pfd = a.meanf ? a.meanf[n] : 0.0;
        const double pf = isfinite(pfd) ? (double)(float)pfd : 0.0;  // the per-node pass's float pilot
        const double nv = masked ? S[RS_Q - 1] : (double)s.S;
        double* out = mom + n * MOM_STRIDE;
        out[0] = nv;
        out[1] = S[0] - nv * pf;
        out[2] = S[1] - 2.0 * pf * S[0] + nv * pf * pf;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double gq[4] = {S[5 + c], S[14 + c], SZ[c], S[2 + c]};
            const double fq[4] = {S[11 + c], S[17 + c], SZ[3 + c], S[8 + c]};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                out[3 + 4 * c + q] = gq[q];
                out[15 + 4 * c + q] = fq[q] - pf * gq[q];
            }
        }
    }
}

// Chunk plan of the LDS-brick path. The box of subset voxels [xc, xc + kx) x [yc, yc + ky) x
// [zc, zc + kz) (sampling indices) warps into the box spanned by the warped positions of its 8
// corners (warp_coord is monotone in X, Y and Z). When every corner is inside the sampling domain,
// so is every sample of the chunk, and the brick holds the stencils of all of them. Returns 0 when
// the chunk fits (origin and extents set), 1 when only its z extent is too large (retry with
// fewer planes), 2 when it cannot use a brick.
template <int MODE, int BRK>
__device__ inline int plan_chunk(const float* sw, const Subset& s, const Domain& dom, int xc, int kx, int yc, int ky,
                                 int zc, int kz, int planes, int rows, int& bx, int& by, int& bz, int& ex, int& ey,
                                 int& ez) {
    const float Xl = (float)(-s.hx + s.stride * xc), Xh = (float)(-s.hx + s.stride * (xc + kx - 1));
    const float Yl = (float)(-s.hy + s.stride * yc), Yh = (float)(-s.hy + s.stride * (yc + ky - 1));
    const float Zl = (float)(-s.hz + s.stride * zc), Zh = (float)(-s.hz + s.stride * (zc + kz - 1));
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float X = (c & 1) ? Xh : Xl, Y = (c & 2) ? Yh : Yl, Z = (c & 4) ? Zh : Zl;
        const float r[3] = {warp_coord(sw[0], sw[1], sw[2], sw[9], X, Y, Z),
                            warp_coord(sw[3], sw[4], sw[5], sw[10], X, Y, Z),
                            warp_coord(sw[6], sw[7], sw[8], sw[11], X, Y, Z)};
        if (!inside(dom, r[0], r[1], r[2])) return 2;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fminf(lo[d], r[d]);
            hi[d] = fmaxf(hi[d], r[d]);
        }
    }
    constexpr int LO = MODE == INTERP_LINEAR ? 0 : -1, HI = MODE == INTERP_LINEAR ? 1 : 2;
    bx = (int)floorf(lo[0]) + LO;
    by = (int)floorf(lo[1]) + LO;
    bz = (int)floorf(lo[2]) + LO;
    ex = (int)floorf(hi[0]) + HI - bx + 1;
    ey = (int)floorf(hi[1]) + HI - by + 1;
    ez = (int)floorf(hi[2]) + HI - bz + 1;
    if (ex > BRK || ey > rows) return 2;
    return ez > planes ? 1 : 0;
}

// Stage planes [0, ez) of a brick of `rows` rows of BRK floats (row stride BRK, plane stride
// `bplane` floats) whose origin is the anchor-relative voxel (bx, by, bz): 16-byte buffer loads of
// the deformed volume (out-of-range offsets read zero; brick voxels outside the volume are never
// read by an in-domain sample). A thread owns at most two fixed (row, 16-byte group) positions
// (rows x groups <= 2 x BLOCK) and walks them through the planes, four planes at a time: the loads
// of both positions are in flight before any is stored. Planes 16-byte aligned in LDS take one
// 16-byte store per group, others four 4-byte stores.
template <int BRK>
__device__ inline void stage_brick(float* brick, int bplane, int cols, int rows, const BufVol& gb, int anchor,
                                   int rowp, int planep, int bx, int by, int bz, int ez) {
    static_assert(BRK % 4 == 0, "brick rows are whole 16-byte groups");
    static_assert(BRK * (BRK / 4) <= 2 * BLOCK, "a brick plane has at most two 16-byte groups per thread");
    const int Q = (cols + 3) >> 2;  // 16-byte groups per row: the columns the chunk's stencils reach
    const int RQ = rows * Q;        // 16-byte groups per plane
    const int org = anchor + bz * planep + by * rowp + bx;
    const bool wide = (bplane & 3) == 0;
    // p / Q as a float product: p < 2 x BLOCK, so (p + 0.5) / Q is at least 0.05 from an integer and
    // the float rounding (< 1e-4) cannot change the floor
    const float inv_q = 1.0f / (float)Q;
    const int pa = threadIdx.x, pb = threadIdx.x + BLOCK;
    const bool va = pa < RQ, vb = pb < RQ;
    const int ya = (int)(((float)pa + 0.5f) * inv_q), yb = (int)(((float)pb + 0.5f) * inv_q);
    const int qa = pa - ya * Q, qb = pb - yb * Q;
    // a lane without a second position loads the volume's first voxels of each plane (in range) and
    // stores nothing: the loads stay outside divergent branches, so they are all issued back to back
    uint32_t oa = va ? 4u * (uint32_t)(org + ya * rowp + 4 * qa) : 0u;
    uint32_t ob = vb ? 4u * (uint32_t)(org + yb * rowp + 4 * qb) : 0u;
    float* da = brick + ya * BRK + 4 * qa;
    float* db = brick + yb * BRK + 4 * qb;
    const uint32_t pstep = 4u * (uint32_t)planep;
    constexpr int U = 4;
    auto put = [&](float* d, const float4& v) {
        if (wide) {
            *reinterpret_cast<float4*>(d) = v;
        } else {
            d[0] = v.x;
            d[1] = v.y;
            d[2] = v.z;
            d[3] = v.w;
        }
    };
    for (int z0 = 0; z0 < ez; z0 += U) {
        float4 ra[U], rb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (z0 + u < ez) {
                ra[u] = buf_load4(gb, oa, 0);
                rb[u] = buf_load4(gb, ob, 0);
            }
            oa += pstep;
            ob += pstep;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (z0 + u < ez) {
                if (va) put(da + (z0 + u) * bplane, ra[u]);
                if (vb) put(db + (z0 + u) * bplane, rb[u]);
            }
        }
    }
}

// SEL: a reference mask or split keep rows select the subset voxels (false: every voxel counts)
// T3: translation-only solve of ADMM subproblem 1 (fixed F, proximal term; numba_kernels.py:931-1082)
// on the same sampling paths, with the deformed-side moments restricted to q = 1.
template <int MODE, bool STORED, bool BUF, int BRK, bool SEL, bool T3 = false>
__global__ __launch_bounds__(BLOCK, 3) void k_icgn12_fused(IcgnArgs a, const int64_t* __restrict__ list,
                                                        const int64_t* count,  // also the node counters
                                                        const double* __restrict__ mom, int brick_planes,
                                                        int brick_plane, int brick_rows, int lane_map, int tile_x,
                                                        int tile_y) {
    static_assert(BRK == 0 || BUF, "the brick path stages through the buffer descriptor");
    const Volumes& V = a.V;
    const Subset& s = a.sub;
    const int tid = threadIdx.x;
    // persistent workgroups: node slots come from per-XCD counters in the workspace header (zeroed
    // by the compaction), own XCD first, then the others' remaining chunks
    __shared__ int64_t s_slot;
    __shared__ int s_victim;
    int64_t* const ctr = const_cast<int64_t*>(count) + NODE_COUNTERS;
    const int64_t M = *count;
    const int home = a.xcd ? (int)(blockIdx.x & 7) : 0;
    if (tid == 0) s_victim = 0;
    for (;;) {
        if (tid == 0) {
            int64_t sl = -1;
            for (int v = s_victim; v < (a.xcd ? 8 : 1); ++v) {
                const int k = (home + v) & 7;
                const int64_t j = (int64_t)atomicAdd(reinterpret_cast<unsigned long long*>(ctr + k), 1ull);
                sl = a.xcd ? chunk_slot(k, j) : j;
                if (sl < M) {
                    s_victim = v;
                    break;
                }
                sl = -1;
            }
            s_slot = sl;
        }
        __syncthreads();
        const int64_t slot = s_slot;
        if (slot < 0) break;
        if (a.wg_clock && tid == 0) a.wg_clock[2 * slot] = wall_clock64();
        const int64_t n = list[slot];
        const int64_t x0 = a.coords[3 * n], y0 = a.coords[3 * n + 1], z0 = a.coords[3 * n + 2];
        const uint8_t* keep = keep_row_of(a, n);

        // dynamic LDS: one NaN bit per sampled voxel, then (brick path) the brick, 16-byte aligned
        extern __shared__ __align__(16) unsigned nanbits[];
        const int nwords = (s.S + 31) >> 5;
        float* brick = reinterpret_cast<float*>(nanbits + ((nwords + 3) & ~3));
        __shared__ float sw[14];  // warp (12), pilots pf, pg
        __shared__ int ctrl;
        __shared__ double red[NWAVES * NR];
        __shared__ double res[NG];
        __shared__ double exc[NR];
        __shared__ std::conditional_t<T3, Ctl3, Ctl12w> c;
        constexpr int NM = T3 ? NG3 : NG;  // deformed-side moments accumulated per lane

        double pg_next = 0.0;
        if (tid < 64) {
            int init = -1;
            if constexpr (T3) {
                if (tid == 0) init = ctl3_init(c, a, n, keep || subset_in_volume(V, s, x0, y0, z0));
            } else {
                // lane k of wave 0 loads parameter k
                bool finite = true;
                if (tid < 12) {
                    const double pk = a.P[12 * n + tid];
                    c.P[tid] = c.P_best[tid] = pk;
                    finite = isfinite(pk);
                }
                const bool all_finite = __ballot(!finite) == 0;
                if (tid == 0) {
                    ctl12_reset(c);
                    if (!all_finite)
                        init = ST_NAN;
                    else if (!(keep || subset_in_volume(V, s, x0, y0, z0)))
                        init = ST_INVALID_SUBSET;
                }
            }
            if (tid == 0) {
                ctrl = init;
                double pf = a.meanf ? a.meanf[n] : 0.0;
                if (!isfinite(pf)) pf = 0.0;
                sw[12] = (float)pf;
                pg_next = pf;
            }
        }
        __syncthreads();
        // the node's reference-side moments, read by the control every iteration
        __shared__ double s_mom[NR];
        for (int i = tid; i < NR; i += BLOCK) s_mom[i] = mom[n * MOM_STRIDE + i];
        // without noise correction the node's Cholesky factor is read every iteration by the control:
        // a copy in LDS (the corrected factor's slot, unused then) replaces its global loads
        const double* Lctl = nullptr;
        if constexpr (!T3) {
            if (!(a.noise_gain > 0.0 && a.H_all && a.pattern)) {
                for (int i = tid; i < 144; i += BLOCK) c.L0[i] = a.L_all[144 * n + i];
                Lctl = c.L0;
            } else {
                Lctl = a.L_all + 144 * n;
            }
        }
        __syncthreads();
        const double* mn = s_mom;
        const double half_scale = (double)max(s.hx, max(s.hy, s.hz));
        const Domain dom = make_domain(V, (int)x0, (int)y0, (int)z0);
        const float* gbase = V.g + lin(V, z0, y0, x0);
        const int64_t obase = lin(V, z0, y0, x0);
        const int rowp = V.nx, planep = V.nx * V.ny;
        const BufVol gb = make_bufvol(V.g, (int64_t)V.nz * planep, rowp, planep);
        int it_done = 0;
        int status = ST_MAX_ITER;
        for (int it = 1; it <= a.max_iter; ++it) {
            if (ctrl >= 0) break;
            it_done = it;
            if (tid == 0) {
                put_warp(sw, c.P);
                sw[13] = (float)pg_next;
            }
            for (int i = tid; i < nwords; i += BLOCK) nanbits[i] = 0u;
            __syncthreads();
            const float pf = sw[12], pg = sw[13];
            float m[NM];
    #pragma unroll
            for (int q = 0; q < NM; ++q) m[q] = 0.0f;
            // one subset voxel: reference-side values first (they do not depend on the warp, so their
            // loads overlap the gather of the deformed sample), then the warp, the sample and the
            // deformed-side moments. `FROM_BRICK`: every sample of the chunk is known to be inside.
            auto voxel = [&](const auto& w, auto from_brick, int bx, int by, int bz) {
                constexpr bool FROM_BRICK = decltype(from_brick)::value;
                const int dx = w.dx, dy = w.dy, dz = w.dz;
                const int64_t o = obase + w.o;
                if constexpr (SEL)
                    if (!voxel_kept(V, keep, w.v, o, x0 + dx, y0 + dy, z0 + dz)) return;
                float fraw, gr[3];
                if constexpr (BUF && STORED) {
                    // a volume under 4 GB: one 32-bit byte offset serves all four reference arrays
                    const uint32_t ob = 4u * (uint32_t)((int)obase + w.o);
                    auto at = [ob](const float* vol) {
                        return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(vol) + ob);
                    };
                    fraw = at(V.f);
                    gr[0] = at(V.gx);
                    gr[1] = at(V.gy);
                    gr[2] = at(V.gz);
                } else {
                    fraw = V.f[o];
                    grad_at<STORED>(V, o, (int)z0 + dz, (int)y0 + dy, (int)x0 + dx, gr[0], gr[1], gr[2]);
                }
                const float X = (float)dx, Y = (float)dy, Z = (float)dz;
                const float xr = warp_coord(sw[0], sw[1], sw[2], sw[9], X, Y, Z);
                const float yr = warp_coord(sw[3], sw[4], sw[5], sw[10], X, Y, Z);
                const float zr = warp_coord(sw[6], sw[7], sw[8], sw[11], X, Y, Z);
                float val;
                if constexpr (FROM_BRICK) {
                    val = sample_brick<MODE, BRK>(brick, brick_plane, bx, by, bz, xr, yr, zr);
                } else {
                    if (!inside(dom, xr, yr, zr)) {
                        m[1] = 1.0f;
                        return;
                    }
                    if constexpr (BUF)
                        val = sample_buf<MODE>(gb, (int)obase, rowp, planep, xr, yr, zr, dom.kx, dom.ky, dom.kz);
                    else
                        val = sample_rel<MODE>(gbase, rowp, planep, xr, yr, zr, dom.kx, dom.ky, dom.kz);
                }
                // a NaN sample (masked in the deformed frame) drops the voxel from this iteration; the
                // accumulation is branch-free (gv = 0 adds nothing), so the reference-side loads above
                // stay ahead of the sample instead of being sunk into a branch
                const bool ok = val == val;
                const float fv = fraw - pf, gv = ok ? val - pg : 0.0f;
                const float q[4] = {X, Y, Z, 1.0f};
                m[3] += gv;
                m[4] += gv * gv;
                m[5] += fv * gv;
                if constexpr (T3) {
    #pragma unroll
                    for (int cc = 0; cc < 3; ++cc) m[6 + cc] += gr[cc] * gv;
                } else {
    #pragma unroll
                    for (int cc = 0; cc < 3; ++cc) {
                        const float tg = gr[cc] * gv;
    #pragma unroll
                        for (int j = 0; j < 4; ++j) m[6 + 4 * cc + j] += tg * q[j];
                    }
                }
                if (!ok) {  // rare: the valid count is the kept count less these
                    m[2] += 1.0f;
                    atomicOr(&nanbits[w.v >> 5], 1u << (w.v & 31));
                }
            };
            using Gather = std::integral_constant<bool, false>;
            if constexpr (BRK > 0) {
                // The subset in boxes of tile_x x tile_y voxel columns (one box when its rows fit the
                // brick), each box in chunks of planes whose warped stencils fit the brick; a chunk
                // that does not fit (large rotation or shear, or a sample outside the domain) gathers
                // from memory.
                constexpr int OVER = MODE == INTERP_LINEAR ? 1 : 3;
                // chunk plans (kz, fit, brick origin, brick extents), double-buffered by chunk parity: wave 0
                // writes chunk c + 1's slot after the barrier that starts chunk c, when every wave has
                // read chunk c - 1's
                __shared__ int s_plan[2][8];
                int chunk = 0;
                const int sx = s.sx, sy = s.sy, plane_vox = sx * sy, lane = tid & 63;
                // the voxels of box [xc, xc + kx) x [yc, yc + ky) x [zc, zc + kz), t = tid, tid + BLOCK,
                // ... with x fastest, decoded per voxel (box chunks that gather, boxes of other lane maps)
                auto box_voxels = [&](auto from_brick, int xc, int kx, int yc, int ky, int zc, int kz, int bx, int by,
                                      int bz) {
                    const int per_plane = kx * ky;
                    for (int t = tid; t < kz * per_plane; t += BLOCK) {
                        const int pz = t / per_plane, r = t - pz * per_plane, py = r / kx;
                        const int ix = xc + r - py * kx, iy = yc + py, iz = zc + pz;
                        Vox q;
                        q.dx = -s.hx + s.stride * ix;
                        q.dy = -s.hy + s.stride * iy;
                        q.dz = -s.hz + s.stride * iz;
                        q.v = (iz * sy + iy) * sx + ix;
                        q.o = q.dz * planep + q.dy * rowp + q.dx;
                        voxel(q, from_brick, bx, by, bz);
                    }
                };
                for (int yc = 0; yc < sy; yc += tile_y) {
                    for (int xc = 0; xc < sx; xc += tile_x) {
                        const int kx = min(tile_x, sx - xc), ky = min(tile_y, sy - yc);
                        const bool whole = kx == sx && ky == sy;
                        for (int zc = 0; zc < s.sz;) {
                            // wave 0 plans the chunk (the plan is uniform; the other waves do not repeat
                            // it) and publishes it through LDS at the barrier that also ends the previous
                            // chunk's use of the brick
                            if (tid < 64) {
                                int kz = min(brick_planes - OVER, s.sz - zc);
                                int bx = 0, by = 0, bz = 0, ex = 0, ey = 0, ez = 0;
                                int fit = plan_chunk<MODE, BRK>(sw, s, dom, xc, kx, yc, ky, zc, kz, brick_planes,
                                                                brick_rows, bx, by, bz, ex, ey, ez);
                                while (fit == 1 && kz > 1)
                                    fit = plan_chunk<MODE, BRK>(sw, s, dom, xc, kx, yc, ky, zc, --kz, brick_planes,
                                                                brick_rows, bx, by, bz, ex, ey, ez);
                                if (tid == 0) {
                                    int* pl = s_plan[chunk & 1];
                                    pl[0] = kz;
                                    pl[1] = fit;
                                    pl[2] = bx;
                                    pl[3] = by;
                                    pl[4] = bz;
                                    pl[5] = ex;
                                    pl[6] = ey;
                                    pl[7] = ez;
                                }
                            }
                            __syncthreads();
                            const int* pl = s_plan[chunk & 1];
                            const int kz = pl[0], fit = pl[1], bx = pl[2], by = pl[3], bz = pl[4];
                            const int ex = pl[5], ey = pl[6], ez = pl[7];
                            ++chunk;
                            using Brick = std::integral_constant<bool, true>;
                            if (fit == 0) {
                                stage_brick<BRK>(brick, brick_plane, ex, ey, gb, (int)obase, rowp, planep, bx, by, bz,
                                                 ez);
                                __syncthreads();
                                if (lane_map == LANES_PATCHES) {
                                    // 8 x 8 patches of one plane per wave (lane l: x + (l & 7), y + (l >> 3)).
                                    // With a brick row stride of 40 floats the 8 rows of a patch start
                                    // 40 y mod 64 = 8 different bank groups apart, so the 64 lanes read 64
                                    // different banks, and a small rotation or shear only shifts whole patch
                                    // rows by a bank or two. The patch counters are wave-uniform and advance
                                    // by NWAVES patches with carries (scalar work, no division); a lane adds
                                    // its fixed offsets within the patch. The box's columns and rows beyond
                                    // its last whole patch follow in a short tail.
                                    const int pxn = kx >> 3, pyn = ky >> 3;
                                    if (pxn > 0 && pyn > 0) {
                                        // A wave takes the chunk's patch columns (pxi, pyi) in turn and walks
                                        // each through the chunk's kz planes, so a lane's X and Y stay fixed
                                        // along the walk: the warp's X and Y terms, the column's offsets and
                                        // the X- and Y-weighted moments are formed once per column (the
                                        // column's sums of grad_c g' are folded in at its end), and a sample
                                        // adds only its Z terms.
                                        const int lx = lane & 7, ly = lane >> 3;
                                        const int npc = pxn * pyn, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
                                        int pxi = wave, pyi = 0;
                                        while (pxi >= pxn) {
                                            pxi -= pxn;
                                            ++pyi;
                                        }
                                        for (int pc = wave; pc < npc; pc += NWAVES) {
                                            const int ix = xc + 8 * pxi + lx, iy = yc + 8 * pyi + ly;
                                            const int dx = -s.hx + ix, dy = -s.hy + iy, dz0 = -s.hz + zc;
                                            const float X = (float)dx, Y = (float)dy;
                                            const float wx = warp_xy(sw[0], sw[1], sw[9], X, Y);
                                            const float wy = warp_xy(sw[3], sw[4], sw[10], X, Y);
                                            const float wz = warp_xy(sw[6], sw[7], sw[11], X, Y);
                                            const int o0 = dz0 * planep + dy * rowp + dx;
                                            const int v0 = (zc * sy + iy) * sx + ix;
                                            float A[3] = {0.0f, 0.0f, 0.0f};  // the column's sums of grad_c g'
                                            for (int pz = 0; pz < kz; ++pz) {
                                                const int dz = dz0 + pz, o = o0 + pz * planep;
                                                if constexpr (SEL)
                                                    if (!voxel_kept(V, keep, v0 + pz * plane_vox, obase + o, x0 + dx,
                                                                    y0 + dy, z0 + dz))
                                                        continue;
                                                float fraw, gr[3];
                                                if constexpr (STORED) {
                                                    const uint32_t ob = 4u * (uint32_t)((int)obase + o);
                                                    auto at = [ob](const float* vol) {
                                                        return *reinterpret_cast<const float*>(
                                                            reinterpret_cast<const char*>(vol) + ob);
                                                    };
                                                    fraw = at(V.f);
                                                    gr[0] = at(V.gx);
                                                    gr[1] = at(V.gy);
                                                    gr[2] = at(V.gz);
                                                } else {
                                                    fraw = V.f[obase + o];
                                                    grad_at<false>(V, obase + o, (int)z0 + dz, (int)y0 + dy, (int)x0 + dx,
                                                                   gr[0], gr[1], gr[2]);
                                                }
                                                const float Z = (float)dz;
                                                const float val = sample_brick<MODE, BRK>(
                                                    brick, brick_plane, bx, by, bz, warp_z(sw[2], Z, wx), warp_z(sw[5], Z, wy),
                                                    warp_z(sw[8], Z, wz));
                                                const bool ok = val == val;
                                                const float fv = fraw - pf, gv = ok ? val - pg : 0.0f;
                                                m[3] += gv;
                                                m[4] += gv * gv;
                                                m[5] += fv * gv;
                                                if constexpr (T3) {
    #pragma unroll
                                                    for (int cc = 0; cc < 3; ++cc) A[cc] = __fmaf_rn(gr[cc], gv, A[cc]);
                                                } else {
                                                    const float gz = gv * Z;
    #pragma unroll
                                                    for (int cc = 0; cc < 3; ++cc) {
                                                        A[cc] = __fmaf_rn(gr[cc], gv, A[cc]);
                                                        m[8 + 4 * cc] = __fmaf_rn(gr[cc], gz, m[8 + 4 * cc]);
                                                    }
                                                }
                                                if (!ok) {
                                                    const int v = v0 + pz * plane_vox;
                                                    m[2] += 1.0f;
                                                    atomicOr(&nanbits[v >> 5], 1u << (v & 31));
                                                }
                                            }
    #pragma unroll
                                            for (int cc = 0; cc < 3; ++cc) {
                                                if constexpr (T3) {
                                                    m[6 + cc] += A[cc];
                                                } else {
                                                    m[6 + 4 * cc] = __fmaf_rn(X, A[cc], m[6 + 4 * cc]);
                                                    m[7 + 4 * cc] = __fmaf_rn(Y, A[cc], m[7 + 4 * cc]);
                                                    m[9 + 4 * cc] += A[cc];
                                                }
                                            }
                                            pxi += NWAVES;
                                            while (pxi >= pxn) {
                                                pxi -= pxn;
                                                ++pyi;
                                            }
                                        }
                                    }
                                    const int tx = kx - 8 * pxn, wxp = 8 * pxn;
                                    const int per_plane = tx * ky + (ky - 8 * pyn) * wxp;
                                    // integer quotients as float products: (n + 0.5) / d lies at least 0.5 / d
                                    // from an integer, and for n below 2^22 its float rounding (under
                                    // n 2^-23 / d) stays inside that margin, so the floor is exact
                                    const float inv_pp = 1.0f / (float)max(per_plane, 1);
                                    const float inv_tx = 1.0f / (float)max(tx, 1), inv_wx = 1.0f / (float)max(wxp, 1);
                                    for (int t = tid; t < kz * per_plane; t += BLOCK) {
                                        const int pz = (int)(((float)t + 0.5f) * inv_pp), r = t - pz * per_plane;
                                        int x, y;
                                        if (r < tx * ky) {
                                            y = (int)(((float)r + 0.5f) * inv_tx);
                                            x = wxp + (r - y * tx);
                                        } else {
                                            const int r2 = r - tx * ky, yy = (int)(((float)r2 + 0.5f) * inv_wx);
                                            y = 8 * pyn + yy;
                                            x = r2 - yy * wxp;
                                        }
                                        const int ix = xc + x, iy = yc + y, iz = zc + pz;
                                        Vox q;
                                        q.dx = -s.hx + ix;
                                        q.dy = -s.hy + iy;
                                        q.dz = -s.hz + iz;
                                        q.v = (iz * sy + iy) * sx + ix;
                                        q.o = q.dz * planep + q.dy * rowp + q.dx;
                                        voxel(q, Brick{}, bx, by, bz);
                                    }
                                } else if (whole && lane_map == LANES_PAIRS) {
                                    // Planes in pairs: lane l < sx takes voxel x = l of row y in the first plane,
                                    // lane l >= sx voxel x = l - sx of row y in the second. The brick's plane
                                    // stride is sx modulo 64 banks, so every lane of a wave reads its own LDS bank.
                                    // The second plane's last 2 sx - 64 voxels of each row, and a last unpaired
                                    // plane, follow.
                                                            const int inB = lane >= sx ? 1 : 0, xl = lane - inB * sx;
                                    const int npair = kz >> 1;
                                    int p = 0, y = tid >> 6;
                                    while (y >= sy) {
                                        y -= sy;
                                        ++p;
                                    }
                                    Vox q;
                                    q.dx = -s.hx + xl;
                                    q.dy = -s.hy + y;
                                    q.dz = -s.hz + zc + 2 * p + inB;
                                    q.v = ((zc + 2 * p + inB) * sy + y) * sx + xl;
                                    q.o = q.dz * planep + q.dy * rowp + q.dx;
                                    const int wrap_v = sy * sx, wrap_o = 2 * planep - sy * rowp;
                                    while (p < npair) {
                                        voxel(q, Brick{}, bx, by, bz);
                                        y += NWAVES;
                                        q.dy += NWAVES;
                                        q.v += NWAVES * sx;
                                        q.o += NWAVES * rowp;
                                        if (y >= sy) {
                                            y -= sy;
                                            ++p;
                                            q.dy -= sy;
                                            q.dz += 2;
                                            q.v += wrap_v;
                                            q.o += wrap_o;
                                        }
                                    }
                                    const int nt = 2 * sx - 64, ntail = npair * sy * nt;
                                    for (int t = tid; t < ntail; t += BLOCK) {
                                        const int r = t / nt, k = t - r * nt, pp = r / sy, yy = r - pp * sy;
                                        const int iz = zc + 2 * pp + 1, xx = 64 - sx + k;
                                        Vox e;
                                        e.dx = -s.hx + xx;
                                        e.dy = -s.hy + yy;
                                        e.dz = -s.hz + iz;
                                        e.v = (iz * sy + yy) * sx + xx;
                                        e.o = e.dz * planep + e.dy * rowp + e.dx;
                                        voxel(e, Brick{}, bx, by, bz);
                                    }
                                    if (kz & 1) {
                                        const int iz = zc + kz - 1;
                                        for (int u = tid; u < sy * sx; u += BLOCK) {
                                            const int yy = u / sx, xx = u - yy * sx;
                                            Vox e;
                                            e.dx = -s.hx + xx;
                                            e.dy = -s.hy + yy;
                                            e.dz = -s.hz + iz;
                                            e.v = (iz * sy + yy) * sx + xx;
                                            e.o = e.dz * planep + e.dy * rowp + e.dx;
                                            voxel(e, Brick{}, bx, by, bz);
                                        }
                                    }
                                } else if (whole) {
                                    OffsetWalk w;
                                    w.start(s, zc * plane_vox + tid, rowp, planep);
                                    const int vend = (zc + kz) * plane_vox;
                                    for (; w.v < vend; w.next(s)) voxel(w, Brick{}, bx, by, bz);
                                } else {
                                    box_voxels(Brick{}, xc, kx, yc, ky, zc, kz, bx, by, bz);
                                }
                            } else if (whole) {
                                OffsetWalk w;
                                w.start(s, zc * plane_vox + tid, rowp, planep);
                                const int vend = (zc + kz) * plane_vox;
                                for (; w.v < vend; w.next(s)) voxel(w, Gather{}, 0, 0, 0);
                            } else {
                                box_voxels(Gather{}, xc, kx, yc, ky, zc, kz, 0, 0, 0);
                            }
                            zc += kz;
                        }
                    }
                }
            } else {
                OffsetWalk w;
                w.start(s, tid, rowp, planep);
                for (; w.v < s.S; w.next(s)) voxel(w, Gather{}, 0, 0, 0);
            }
            block_sum<NM, NWAVES>(m, res, red);
            const bool correct = res[2] > 0.0 && !(res[1] > 0.0);  // block-uniform
            if (correct) {
                ref_moments<STORED>(
                    a, keep, x0, y0, z0, pf,
                    [&](int v, int64_t, int64_t, int64_t, int64_t) {
                        return ((nanbits[v >> 5] >> (v & 31)) & 1u) != 0;
                    },
                    exc, red);
            }
            // the control: lane 0 for the translation-only solve, wave 0 for the 12-DOF one
            if (T3 ? tid == 0 : tid < 64) {
                const int lane = tid;
                const double n_ref = mn[0], n_valid = n_ref - res[2];  // every kept voxel was sampled once
                int st = -1;
                if (res[1] > 0.0) {
                    st = ST_OUT_OF_BOUNDS;
                } else if (n_valid < MIN_SUBSET_VOXELS || n_valid < MIN_VALID_FRACTION * n_ref) {
                    st = ST_INVALID_SUBSET;
                }
                if (st >= 0) {
                    if (lane == 0) {
                        c.zncc = NAN;
                        ctrl = st;
                    }
                } else {
                    // reference-side moments of this iteration's voxel set
                    auto R = [&](int q) { return correct ? mn[q] - exc[q] : mn[q]; };
                    // centred statistics on the dynamic voxel set (numba_kernels.py:586-598)
                    const double mf = R(1) / n_valid, mg = res[3] / n_valid;
                    double ssf = R(2) - n_valid * mf * mf, ssg = res[4] - n_valid * mg * mg;
                    ssf = ssf < 0.0 ? 0.0 : ssf;
                    ssg = ssg < 0.0 ? 0.0 : ssg;
                    const double bf = sqrt(ssf > 1e-30 ? ssf : 1e-30), bg = sqrt(ssg > 1e-30 ? ssg : 1e-30);
                    if constexpr (T3) {
                        // q = 1 columns of the reference moments: C_c = R(6 + 4c), A_c = R(18 + 4c)
                        const double C[3] = {R(6), R(10), R(14)}, A[3] = {R(18), R(22), R(26)};
                        ctrl = ctl3_step(c, a, n, n_valid, mf, mg, bf, bg, res[5], C, A, res + 6);
                    } else {
                        const double zncc = (res[5] - n_valid * mf * mg) / (bf * bg);
                        const double ratio = bf / bg;
                        // lane k: gradient entry k (k < 9: component k / 3 against q_(k mod 3); k >= 9:
                        // component k - 9 against q = 1), from moment q = 4 c + j
                        double bk = 0.0;
                        const int lane = lane_here();
                        if (lane < 12) {
                            const int q = lane < 9 ? 4 * (lane / 3) + lane % 3 : 4 * (lane - 9) + 3;
                            const double C = R(3 + q), A = R(15 + q), B = res[6 + q];
                            bk = (A - mf * C) - ratio * (B - mg * C);
                        }
                        const double* Hn = a.H_all ? a.H_all + 144 * n : nullptr;
                        const int r = ctl12_step_wave(c, bk, zncc, bf, n_valid, Lctl, Hn, a.pattern, a.noise_gain,
                                                      a.tol, a.dp_tol, a.patience, a.predictive, half_scale);
                        if (lane == 0) ctrl = r;
                    }
                    pg_next = (double)pg + mg;
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            if (ctrl >= 0) status = ctrl;
            if constexpr (T3) {
                for (int k = 0; k < 3; ++k) a.U[3 * n + k] = c.P[9 + k];
            } else {
                for (int k = 0; k < 12; ++k) a.P[12 * n + k] = c.P[k];
            }
            a.n_iter[n] = it_done;
            a.status[n] = (int8_t)status;
            a.zncc[n] = (status == ST_CONVERGED || status == ST_STALLED || status == ST_MAX_ITER) ? c.zncc : NAN;
            if (a.wg_clock) a.wg_clock[2 * slot + 1] = wall_clock64();
        }
        __syncthreads();  // the next node's setup reuses the LDS state
    }
}

// ---------------------------------------------------------------------------- precompute
// Per node: H = sum SD SD^T over the kept voxels, sum f, sum f^2, n_valid (float32 lane partials,
// float64 cross-lane sums), then meanf/bottomf, the float64 Cholesky factor and the validity
// rules of numba_kernels.py:391-417. Subsets that leave the volume without a keep row are
// rejected before any read (numba_kernels.py:342-343).
// Shared tail of the precompute: the 81 sums r (the Hessian's upper triangle row by row, sum f,
// sum f^2, the kept count) -> H, the Cholesky factor L, mean, norm, count and validity of one node
// (thread 0 decides and factors in LDS, Hs and Ls: 144 doubles each; every thread of the block
// calls it and copies H and L out).
__device__ inline void precompute_finish(const double* r, int S, double min_valid_ratio, double cond_max, double* Hs,
                                         double* Ls, double* H, double* L, double* meanf, double* bottomf,
                                         int64_t* nvalid, uint8_t* valid) {
    if (threadIdx.x == 0) {
        int q = 0;
        for (int i = 0; i < 12; ++i)
            for (int j = i; j < 12; ++j) {
                Hs[i * 12 + j] = r[q];
                Hs[j * 12 + i] = r[q];
                ++q;
            }
        const double nv = r[80];
        const int64_t nvi = (int64_t)(nv + 0.5);
        double mf = 0.0, bf = 1.0;
        bool ok = true;
        if (nvi < 27 || (double)nvi < min_valid_ratio * (double)S) {
            ok = false;
        } else {
            mf = r[78] / nv;
            double ssf = r[79] - nv * mf * mf;
            ssf = ssf < 0.0 ? 0.0 : ssf;
            bf = sqrt(ssf > 1e-30 ? ssf : 1e-30);
            if (bf < 1e-10) ok = false;
        }
        *nvalid = nvi;
        *meanf = mf;
        *bottomf = bf;
        if (ok) ok = cholesky12(Hs, Ls);
        else
            for (int i = 0; i < 144; ++i) Ls[i] = 0.0;
        if (ok) {
            double dmin = Ls[0], dmax = Ls[0];
            for (int i = 1; i < 12; ++i) {
                dmin = Ls[i * 13] < dmin ? Ls[i * 13] : dmin;
                dmax = Ls[i * 13] > dmax ? Ls[i * 13] : dmax;
            }
            if (dmin <= 0.0 || (dmax / dmin) * (dmax / dmin) > cond_max) ok = false;
        }
        *valid = ok ? 1 : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 144; i += blockDim.x) {
        H[i] = Hs[i];
        L[i] = Ls[i];
    }
}

// skip: non-null when the separable precompute may have taken the nodes without keep rows whose
// subsets lie inside the volume (its flag, set on the device, says whether it did)
template <bool STORED>
__global__ __launch_bounds__(BLOCK) void k_precompute(Volumes V, Subset s, const int64_t* __restrict__ coords,
                                                      int64_t N, double min_valid_ratio, double cond_max,
                                                      const int64_t* split_index, const uint8_t* split_keep,
                                                      int64_t keep_bytes, double* H_all, double* L_all,
                                                      double* meanf, double* bottomf, int64_t* nvalid,
                                                      uint8_t* valid, const int* __restrict__ skip) {
    const int64_t n = blockIdx.x;
    if (n >= N) return;
    const int tid = threadIdx.x;
    const int64_t x0 = coords[3 * n], y0 = coords[3 * n + 1], z0 = coords[3 * n + 2];
    const uint8_t* keep =
        (split_index && split_keep && split_index[n] >= 0) ? split_keep + split_index[n] * keep_bytes : nullptr;
    if (skip && *skip && !keep && subset_in_volume(V, s, x0, y0, z0)) return;
    double* H = H_all + 144 * n;
    double* L = L_all + 144 * n;
    if (!keep && !subset_in_volume(V, s, x0, y0, z0)) {
        for (int i = tid; i < 144; i += BLOCK) {
            H[i] = 0.0;
            L[i] = 0.0;
        }
        if (tid == 0) {
            meanf[n] = 0.0;
            bottomf[n] = 1.0;
            nvalid[n] = 0;
            valid[n] = 0;
        }
        return;
    }
    constexpr int NQ = 81;
    __shared__ double red[NWAVES * NQ];
    __shared__ double res[NQ];
    __shared__ double Hs[144];
    __shared__ double Ls[144];
    float acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0f;
    const int64_t obase = lin(V, z0, y0, x0);
    const int rowp = V.nx, planep = V.nx * V.ny;
    VoxelWalk w;
    w.start(s, tid);
    for (; w.v < s.S; w.next(s)) {
        const int dx = -s.hx + s.stride * w.ix, dy = -s.hy + s.stride * w.iy, dz = -s.hz + s.stride * w.iz;
        const int64_t o = obase + (int64_t)dz * planep + (int64_t)dy * rowp + dx;
        bool k;
        if (keep) {
            k = kept_bit(keep, w.v);
            const int64_t xx = x0 + dx, yy = y0 + dy, zz = z0 + dz;
            k = k && xx >= 0 && xx < V.nx && yy >= 0 && yy < V.ny && zz >= 0 && zz < V.nz;
        } else {
            k = !V.mask || V.mask[o] != 0;
        }
        if (!k) continue;
        const float fv = V.f[o];
        float gxv, gyv, gzv;
        grad_at<STORED>(V, o, (int)z0 + dz, (int)y0 + dy, (int)x0 + dx, gxv, gyv, gzv);
        const float X = (float)dx, Y = (float)dy, Z = (float)dz;
        const float sd[12] = {gxv * X, gxv * Y, gxv * Z, gyv * X, gyv * Y, gyv * Z,
                              gzv * X, gzv * Y, gzv * Z, gxv, gyv, gzv};
        int q = 0;
#pragma unroll
        for (int i = 0; i < 12; ++i)
#pragma unroll
            for (int j = i; j < 12; ++j) acc[q++] += sd[i] * sd[j];
        acc[78] += fv;
        acc[79] += fv * fv;
        acc[80] += 1.0f;
    }
    block_sum<NQ, NWAVES>(acc, res, red);
    precompute_finish(res, s.S, min_valid_ratio, cond_max, Hs, Ls, H, L, meanf + n, bottomf + n, nvalid + n,
                      valid + n);
}

// ---------------------------------------------------------------------------- separable precompute
// The precompute's sums are box sums too: H[(c,i),(d,j)] = sum w g_c g_d q_i q_j with q = (X, Y, Z, 1)
// is, per gradient product p = (c, d) (six of them), a sum of g_c g_d times a monomial of degree
// <= 2 in the voxel's offsets; with sum w f, sum w f^2 and the kept count that is 63 box sums per
// node, taken axis by axis at the distinct node coordinates as the reference moments are (same
// marks, index and buffers):
//   k_pc_xy:   per plane and centre pair, per product p: g g {1, Y, Y^2, X, X Y, X^2} (x-row sums of
//              g g {1, X, X^2} in float, added row by row in double times {1, Y, Y^2}), and f, f^2
//              and the kept count -- PC_Q = 39 rectangle sums;
//   k_pc_zfin: per node, one wave: the sz planes' rectangle sums times {1, Z, Z^2}, gathered into
//              the 81 sums of the per-node kernel, and its tail (validity, Cholesky).
// Rectangle sums R[((p * PC_Q + q) * ncy + cyi) * ncx + cxi], q = 6 p + k for product p and
// k = 0 g g, 1 g g Y, 2 g g Y^2, 3 g g X, 4 g g X Y, 5 g g X^2; 36 f, 37 f^2, 38 the kept count.
constexpr int PC_Q = 39;
constexpr int PC_G = 2;  // consecutive centres y per work item of k_pc_xy
constexpr int PC_SLOTS = 12;  // rectangle sums per quad lane: products {0, 1}, {2, 3}, {4} + f, f^2, n, {5}
// x-row sums: 3 p + {0, 1, 2} = g g {1, X, X^2} of product p, 18 f, 19 f^2, 20 the kept count
__constant__ const int8_t PC_DST[4][PC_SLOTS] = {{0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11},
                                                 {12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23},
                                                 {24, 25, 26, 27, 28, 29, 36, 37, 38, -1, -1, -1},
                                                 {30, 31, 32, 33, 34, 35, -1, -1, -1, -1, -1, -1}};
constexpr int8_t PC_SRC[4][PC_SLOTS] = {{0, 0, 0, 1, 1, 2, 3, 3, 3, 4, 4, 5},
                                        {6, 6, 6, 7, 7, 8, 9, 9, 9, 10, 10, 11},
                                        {12, 12, 12, 13, 13, 14, 18, 19, 20, 20, 20, 20},
                                        {15, 15, 15, 16, 16, 17, 17, 17, 17, 17, 17, 17}};
constexpr int8_t PC_YPOW[4][PC_SLOTS] = {{0, 1, 2, 0, 1, 0, 0, 1, 2, 0, 1, 0},
                                         {0, 1, 2, 0, 1, 0, 0, 1, 2, 0, 1, 0},
                                         {0, 1, 2, 0, 1, 0, 0, 0, 0, 0, 0, 0},
                                         {0, 1, 2, 0, 1, 0, 0, 0, 0, 0, 0, 0}};
constexpr int8_t PC_VALID[4] = {12, 12, 9, 6};  // slots in use per part (with the kept count)

// One wave per work item, as k_rs_xy (PC_G centres y per item); lane 4 j + k adds its PC_SLOTS
// rectangle sums of every window of the item.
template <bool MASKED>
__global__ __launch_bounds__(BLOCK, 2) void k_pc_xy(Volumes V, Subset s, RowMoments r) {
    if (!r.hdr[RS_FITS]) return;
    constexpr int NM = MASKED ? 21 : 20;
    const int ncx = r.hdr[RS_NCX], ncy = r.hdr[RS_NCY], zbase = r.hdr[RS_ZBASE];
    const int ntile = (ncx + RS_TILE - 1) / RS_TILE;
    const int lane = threadIdx.x & 63, j = lane >> 2, part = lane & 3;
    const int base = s.sx >> 2, extra = s.sx & 3, k0 = part * base + min(part, extra);
    const int kx = min(k0 + base, s.sx - 1);
    const bool has_extra = part < extra;
    const int64_t plane = (int64_t)V.nx * V.ny, cstride = (int64_t)ncy * ncx;
    const int nslot = PC_VALID[part] - (MASKED || part != 2 ? 0 : 1);
    const float w0 = part == 0 ? 1.0f : 0.0f, w1 = part == 1 ? 1.0f : 0.0f, w2 = part == 2 ? 1.0f : 0.0f,
                w3 = part == 3 ? 1.0f : 0.0f;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ngroup = (ncy + PC_G - 1) / PC_G;
    const int64_t pairs = (int64_t)r.hdr[RS_NPLANES] * ngroup, xcd = blockIdx.x & 7;
    const int64_t lo = pairs * xcd / 8, items = (pairs * (xcd + 1) / 8 - lo) * ntile;
    for (int64_t it = (blockIdx.x >> 3) * (int64_t)NWAVES + wave; it < items;
         it += (int64_t)(gridDim.x >> 3) * NWAVES) {
        const int64_t pair = lo + it / ntile;
        const int tile = (int)(it - (it / ntile) * ntile), p = (int)(pair / ngroup);
        const int c0 = (int)(pair - (int64_t)p * ngroup) * PC_G, gn = min(PC_G, ncy - c0);
        const int cxi = tile * RS_TILE + j;
        const bool act = cxi < ncx;
        const int cx = act ? r.cxs[cxi] : r.cxs[0];
        int cyg[PC_G];
#pragma unroll
        for (int g = 0; g < PC_G; ++g) cyg[g] = r.cys[c0 + min(g, gn - 1)];
        const int64_t pbase = (int64_t)(zbase + p) * plane;
        const unsigned xoff = (unsigned)(cx - s.hx);
        double P[PC_G][PC_SLOTS];
#pragma unroll
        for (int g = 0; g < PC_G; ++g)
#pragma unroll
            for (int i = 0; i < PC_SLOTS; ++i) P[g][i] = 0.0;
        for (int y = cyg[0] - s.hy; y <= cyg[gn - 1] + s.hy; ++y) {
            const int64_t orow = pbase + (int64_t)y * V.nx;
            const char *fr = (const char*)(V.f + orow), *g0r = (const char*)(V.gx + orow),
                       *g1r = (const char*)(V.gy + orow), *g2r = (const char*)(V.gz + orow);
            const uint8_t* mr = MASKED ? V.mask + orow : nullptr;
            float m[NM];
#pragma unroll
            for (int q = 0; q < NM; ++q) m[q] = 0.0f;
            auto add = [&](bool in, float fv, float g0, float g1, float g2, int k) {
                const float f1 = in ? fv : 0.0f, X = (float)(k - s.hx);
                const float gc[3] = {in ? g0 : 0.0f, in ? g1 : 0.0f, in ? g2 : 0.0f};
                const float gg[6] = {gc[0] * gc[0], gc[0] * gc[1], gc[0] * gc[2],
                                     gc[1] * gc[1], gc[1] * gc[2], gc[2] * gc[2]};
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const float gx = gg[q] * X;
                    m[3 * q] += gg[q];
                    m[3 * q + 1] += gx;
                    m[3 * q + 2] = __fmaf_rn(gx, X, m[3 * q + 2]);
                }
                m[18] += f1;
                m[19] = __fmaf_rn(f1, f1, m[19]);
                if constexpr (MASKED) m[20] += in ? 1.0f : 0.0f;
            };
            for (int kc = 0; kc < max(base, 1); kc += RS_KC) {
                float fv[RS_KC + 1], gv[3][RS_KC + 1];
                uint8_t mk[RS_KC + 1];
#pragma unroll
                for (int t = 0; t <= RS_KC; ++t) {  // t = RS_KC: the extra voxel, fetched with the first chunk
                    const int k = t < RS_KC ? min(k0 + max(min(kc + t, base - 1), 0), s.sx - 1) : kx;
                    const unsigned e = xoff + (unsigned)k, b = 4u * e;
                    fv[t] = *(const float*)(fr + b);
                    gv[0][t] = *(const float*)(g0r + b);
                    gv[1][t] = *(const float*)(g1r + b);
                    gv[2][t] = *(const float*)(g2r + b);
                    if constexpr (MASKED) mk[t] = mr[e];
                }
                auto chunk = [&](auto full) {
#pragma unroll
                    for (int t = 0; t < RS_KC; ++t) {
                        bool in = decltype(full)::value || kc + t < base;
                        if constexpr (MASKED) in = in && mk[t] != 0;
                        add(in, fv[t], gv[0][t], gv[1][t], gv[2][t], k0 + kc + t);
                    }
                };
                if (kc + RS_KC <= base) chunk(std::true_type{});  // uniform
                else chunk(std::false_type{});
                if (kc == 0) {
                    bool in = has_extra;
                    if constexpr (MASKED) in = in && mk[RS_KC] != 0;
                    add(in, fv[RS_KC], gv[0][RS_KC], gv[1][RS_KC], gv[2][RS_KC], kx);
                }
            }
#pragma unroll
            for (int q = 0; q < NM; ++q) m[q] = quad_sum(m[q]);
            float v[PC_SLOTS];
#pragma unroll
            for (int i = 0; i < PC_SLOTS; ++i) {
                const int a0 = min((int)PC_SRC[0][i], NM - 1), a1 = min((int)PC_SRC[1][i], NM - 1);
                const int a2 = min((int)PC_SRC[2][i], NM - 1), a3 = min((int)PC_SRC[3][i], NM - 1);
                v[i] = __fmaf_rn(m[a3], w3, __fmaf_rn(m[a2], w2, __fmaf_rn(m[a1], w1, m[a0] * w0)));
            }
#pragma unroll
            for (int g = 0; g < PC_G; ++g) {
                const int dy = y - cyg[g];
                if (g < gn && dy >= -s.hy && dy <= s.hy) {  // uniform
                    const double Y = (double)dy, Y2 = Y * Y;
#pragma unroll
                    for (int i = 0; i < PC_SLOTS; ++i) {
                        const int yp = part < 2 ? PC_YPOW[0][i] : PC_YPOW[2][i];
                        P[g][i] = fma(yp == 0 ? 1.0 : (yp == 1 ? Y : Y2), (double)v[i], P[g][i]);
                    }
                }
            }
        }
        if (act) {
#pragma unroll
            for (int g = 0; g < PC_G; ++g) {
                if (g >= gn) break;
                double* Rc = r.R + (int64_t)p * PC_Q * cstride + (int64_t)(c0 + g) * ncx + cxi;
#pragma unroll
                for (int i = 0; i < PC_SLOTS; ++i)
                    if (i < nslot) Rc[(int64_t)PC_DST[part][i] * cstride] = P[g][i];
            }
        }
    }
}

// Index into the per-node z sums (3 q + z power) of each of the 81 precompute sums: the Hessian's
// upper triangle row by row over SD = (g_x X, g_x Y, g_x Z, g_y X, ..., g_z Z, g_x, g_y, g_z), then
// f (q 36), f^2 (37) and the kept count (38).
struct PcGather {
    uint8_t at[81];
};
constexpr PcGather make_pc_gather() {
    PcGather g{};
    // q = (X, Y, Z, 1) pairs -> (k, z power): 1 (0,0), Z (0,1), ZZ (0,2), Y (1,0), YZ (1,1), YY (2,0),
    // X (3,0), XZ (3,1), XY (4,0), XX (5,0)
    const int mk[4][4] = {{5, 4, 3, 3}, {4, 2, 1, 1}, {3, 1, 0, 0}, {3, 1, 0, 0}};
    const int mz[4][4] = {{0, 0, 1, 0}, {0, 0, 1, 0}, {1, 1, 2, 1}, {0, 0, 1, 0}};
    const int pp[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    int t = 0;
    for (int i = 0; i < 12; ++i)
        for (int j = i; j < 12; ++j) {
            const int ci = i < 9 ? i / 3 : i - 9, qi = i < 9 ? i % 3 : 3;
            const int cj = j < 9 ? j / 3 : j - 9, qj = j < 9 ? j % 3 : 3;
            g.at[t++] = (uint8_t)(3 * (6 * pp[ci][cj] + mk[qi][qj]) + mz[qi][qj]);
        }
    g.at[78] = 3 * 36;
    g.at[79] = 3 * 37;
    g.at[80] = 3 * 38;
    return g;
}
__constant__ const PcGather PC_GATHER = make_pc_gather();

// One 64-lane workgroup per node without keep rows inside the volume: lane q < PC_Q sums rectangle
// sum q over the node's sz planes times {1, Z, Z^2}; the 81 sums are gathered and finished as by the
// per-node kernel.
__global__ __launch_bounds__(64) void k_pc_zfin(Volumes V, Subset s, const int64_t* __restrict__ coords, int64_t N,
                                                const int64_t* __restrict__ split_index,
                                                const uint8_t* __restrict__ split_keep, RowMoments r,
                                                double min_valid_ratio, double cond_max, double* H_all,
                                                double* L_all, double* meanf, double* bottomf, int64_t* nvalid,
                                                uint8_t* valid) {
    if (!r.hdr[RS_FITS]) return;
    const int64_t n = blockIdx.x;
    if (n >= N || (split_index && split_keep && split_index[n] >= 0)) return;  // keep rows: per node
    const int64_t x0 = coords[3 * n], y0 = coords[3 * n + 1], z0 = coords[3 * n + 2];
    if (!subset_in_volume(V, s, x0, y0, z0)) return;
    __shared__ double zs[3 * PC_Q];
    __shared__ double res[81];
    __shared__ double Hs[144];
    __shared__ double Ls[144];
    const int ncx = r.hdr[RS_NCX], ncy = r.hdr[RS_NCY], zbase = r.hdr[RS_ZBASE];
    const int64_t cstride = (int64_t)ncy * ncx;
    const int q = threadIdx.x;
    if (q < PC_Q) {
        const double* Rq = r.R + (int64_t)q * cstride + (int64_t)r.yidx[y0] * ncx + r.xidx[x0];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int dz = -s.hz; dz <= s.hz; ++dz) {
            const double v = Rq[(int64_t)(z0 + dz - zbase) * PC_Q * cstride], Z = (double)dz;
            a0 += v;
            a1 = fma(Z, v, a1);
            a2 = fma(Z * Z, v, a2);
        }
        zs[3 * q] = a0;
        zs[3 * q + 1] = a1;
        zs[3 * q + 2] = a2;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 81; t += 64) res[t] = t == 80 && !V.mask ? (double)s.S : zs[PC_GATHER.at[t]];
    __syncthreads();
    precompute_finish(res, s.S, min_valid_ratio, cond_max, Hs, Ls, H_all + 144 * n, L_all + 144 * n, meanf + n,
                      bottomf + n, nvalid + n, valid + n);
}

// ---------------------------------------------------------------------------- ZNCC of given warps
// numba_kernels.py:1354-1387: ZNCC of each valid node's subset under the warp P (NaN when the
// warped subset leaves the domain or keeps too few voxels).
template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_evaluate_zncc(Volumes V, Subset s, const int64_t* __restrict__ coords,
                                                         int64_t N, const double* __restrict__ P_all,
                                                         const uint8_t* __restrict__ valid, double* out) {
    const int64_t n = blockIdx.x;
    if (n >= N) return;
    const int tid = threadIdx.x;
    const int64_t x0 = coords[3 * n], y0 = coords[3 * n + 1], z0 = coords[3 * n + 2];
    if (!valid[n] || !subset_in_volume(V, s, x0, y0, z0)) {
        if (tid == 0) out[n] = NAN;
        return;
    }
    constexpr int NQ = 8;  // n, n_ref, oob, f', f'^2, g', g'^2, f'g'
    __shared__ double red[NWAVES * NQ];
    __shared__ double res[NQ];
    float sw[12];
    put_warp(sw, P_all + 12 * n);
    const Domain dom = make_domain(V, (int)x0, (int)y0, (int)z0);
    const float* gbase = V.g + lin(V, z0, y0, x0);
    const int64_t obase = lin(V, z0, y0, x0);
    const int rowp = V.nx, planep = V.nx * V.ny;
    const float pf = V.f[obase];
    float m[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) m[q] = 0.0f;
    VoxelWalk w;
    w.start(s, tid);
    for (; w.v < s.S; w.next(s)) {
        const int dx = -s.hx + s.stride * w.ix, dy = -s.hy + s.stride * w.iy, dz = -s.hz + s.stride * w.iz;
        const int64_t o = obase + (int64_t)dz * planep + (int64_t)dy * rowp + dx;
        if (V.mask && V.mask[o] == 0) continue;
        m[1] += 1.0f;
        const float X = (float)dx, Y = (float)dy, Z = (float)dz;
        const float xr = sw[0] * X + sw[1] * Y + sw[2] * Z + sw[9];
        const float yr = sw[3] * X + sw[4] * Y + sw[5] * Z + sw[10];
        const float zr = sw[6] * X + sw[7] * Y + sw[8] * Z + sw[11];
        if (!inside(dom, xr, yr, zr)) {
            m[2] = 1.0f;
            continue;
        }
        const float val = sample_rel<MODE>(gbase, rowp, planep, xr, yr, zr, dom.kx, dom.ky, dom.kz);
        if (val != val) continue;
        const float fv = V.f[o] - pf, gv = val - pf;
        m[0] += 1.0f;
        m[3] += fv;
        m[4] += fv * fv;
        m[5] += gv;
        m[6] += gv * gv;
        m[7] += fv * gv;
    }
    block_sum<NQ, NWAVES>(m, res, red);
    if (tid == 0) {
        const double nv = res[0];
        double z = NAN;
        if (!(res[2] > 0.0) && nv >= MIN_SUBSET_VOXELS && nv >= MIN_VALID_FRACTION * res[1]) {
            const double mf = res[3] / nv, mg = res[5] / nv;
            double ssf = res[4] - nv * mf * mf, ssg = res[6] - nv * mg * mg;
            ssf = ssf < 0.0 ? 0.0 : ssf;
            ssg = ssg < 0.0 ? 0.0 : ssg;
            const double bf = sqrt(ssf > 1e-30 ? ssf : 1e-30), bg = sqrt(ssg > 1e-30 ? ssg : 1e-30);
            z = (res[7] - nv * mf * mg) / (bf * bg);
        }
        out[n] = z;
    }
}

// ---------------------------------------------------------------------------- small kernels
// 7-point central differences, float32 arithmetic, zero within `border` of a face
// (io/volume_ops.py:124-155).
__global__ void k_gradients(const float* __restrict__ f, int nz, int ny, int nx, int border, float* __restrict__ gx,
                            float* __restrict__ gy, float* __restrict__ gz) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int z = blockIdx.z;
    if (x >= nx) return;
    const int64_t o = ((int64_t)z * ny + y) * nx + x;
    const int64_t r = nx, p = (int64_t)nx * ny;
    if (z >= border && z < nz - border && y >= border && y < ny - border && x >= border && x < nx - border) {
        gx[o] = C1 * (f[o + 1] - f[o - 1]) + C2 * (f[o + 2] - f[o - 2]) + C3 * (f[o + 3] - f[o - 3]);
        gy[o] = C1 * (f[o + r] - f[o - r]) + C2 * (f[o + 2 * r] - f[o - 2 * r]) + C3 * (f[o + 3 * r] - f[o - 3 * r]);
        gz[o] = C1 * (f[o + p] - f[o - p]) + C2 * (f[o + 2 * p] - f[o - 2 * p]) + C3 * (f[o + 3 * p] - f[o - 3 * p]);
    } else {
        gx[o] = 0.0f;
        gy[o] = 0.0f;
        gz[o] = 0.0f;
    }
}

// interp_kernels.py:139-152: NaN outside the admissible domain
template <int MODE>
__global__ void k_sample_points(const float* __restrict__ vol, int nz, int ny, int nx, const double* zs,
                                const double* ys, const double* xs, int64_t n, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Volumes V{};
    V.nz = nz;
    V.ny = ny;
    V.nx = nx;
    // anchor at the voxel below the point so the relative coordinates stay small
    const double z = zs[i], y = ys[i], x = xs[i];
    if (!(x >= 1.0 && x <= nx - 3.0 && y >= 1.0 && y <= ny - 3.0 && z >= 1.0 && z <= nz - 3.0)) {
        out[i] = NAN;
        return;
    }
    const int ax = (int)floor(x), ay = (int)floor(y), az = (int)floor(z);
    const Domain d = make_domain(V, ax, ay, az);
    const float* base = vol + (((int64_t)az * ny + ay) * nx + ax);
    out[i] = sample_rel<MODE>(base, nx, nx * ny, (float)(x - ax), (float)(y - ay), (float)(z - az), d.kx, d.ky, d.kz);
}

// Indices of the valid nodes in order (single workgroup scan; N is at most a few million).
__global__ __launch_bounds__(1024) void k_compact(const uint8_t* __restrict__ valid, int64_t N, int64_t* __restrict__ list,
                                                  int64_t* __restrict__ count) {
    __shared__ int64_t wsum[16];
    __shared__ int64_t base;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t start = 0; start < N; start += 1024) {
        const int64_t i = start + threadIdx.x;
        const int flag = (i < N && valid[i]) ? 1 : 0;
        // inclusive scan within the wave
        int64_t v = flag;
        for (int off = 1; off < 64; off <<= 1) {
            const int64_t t = __shfl_up(v, off, 64);
            if (lane >= off) v += t;
        }
        if (lane == 63) wsum[wave] = v;
        __syncthreads();
        int64_t before = 0;
        for (int k = 0; k < wave; ++k) before += wsum[k];
        if (flag) list[base + before + v - 1] = i;
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t tot = 0;
            for (int k = 0; k < 16; ++k) tot += wsum[k];
            base += tot;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = base;
    if (threadIdx.x < 8) count[NODE_COUNTERS + threadIdx.x] = 0;
}

// Valid nodes grouped by spatial tile (a counting sort on the tile of each node centre, tiles in
// (z, y, x) order). Consecutive list entries are then neighbours whose subsets overlap, which is
// what the XCD mapping of the fused kernels relies on for L2 reuse. The order inside a tile is
// whatever the atomics give: every node is solved independently, so results do not depend on it.
constexpr int TILE_HIST_MAX = 12288;

__global__ __launch_bounds__(1024) void k_compact_tiled(const uint8_t* __restrict__ valid, const int64_t* __restrict__ coords,
                                                        int64_t N, int tx, int ty, int tz, int ntx, int nty, int ntz,
                                                        int64_t* __restrict__ list, int64_t* __restrict__ count) {
    __shared__ unsigned hist[TILE_HIST_MAX];
    __shared__ unsigned wtot[16];
    const int T = ntx * nty * ntz;
    for (int t = threadIdx.x; t < T; t += 1024) hist[t] = 0u;
    __syncthreads();
    auto tile_of = [&](int64_t n) {  // coordinates clamped to int first: 32-bit quotients
        auto c32 = [](int64_t v) { return (int)(v < INT_MIN ? INT_MIN : (v > INT_MAX ? INT_MAX : v)); };
        int cx = c32(coords[3 * n]) / tx, cy = c32(coords[3 * n + 1]) / ty, cz = c32(coords[3 * n + 2]) / tz;
        cx = cx < 0 ? 0 : (cx >= ntx ? ntx - 1 : cx);
        cy = cy < 0 ? 0 : (cy >= nty ? nty - 1 : cy);
        cz = cz < 0 ? 0 : (cz >= ntz ? ntz - 1 : cz);
        return (cz * nty + cy) * ntx + cx;
    };
    for (int64_t i = threadIdx.x; i < N; i += 1024)
        if (valid[i]) atomicAdd(&hist[tile_of(i)], 1u);
    __syncthreads();
    // exclusive scan of the histogram: each thread scans a contiguous run of tiles
    const int per = (T + 1023) / 1024;
    const int t0 = threadIdx.x * per, t1 = min(T, t0 + per);
    unsigned run = 0;
    for (int t = t0; t < t1; ++t) run += hist[t];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned incl = run;
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    unsigned before = incl - run;
    for (int k = 0; k < wave; ++k) before += wtot[k];
    for (int t = t0; t < t1; ++t) {
        const unsigned h = hist[t];
        hist[t] = before;
        before += h;
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < N; i += 1024)
        if (valid[i]) list[atomicAdd(&hist[tile_of(i)], 1u)] = i;
    if (threadIdx.x == 0) {
        unsigned tot = 0;
        for (int k = 0; k < 16; ++k) tot += wtot[k];
        *count = tot;
    }
    if (threadIdx.x < 8) count[NODE_COUNTERS + threadIdx.x] = 0;
}

// A launch switched off by a device flag (a stopped ADMM loop under graph replay): no node runs.
__global__ void k_gate_count(int64_t* __restrict__ count, const int* __restrict__ gate) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && gate && *gate) *count = 0;
}

__global__ void k_copy_gated(const double* __restrict__ src, double* __restrict__ dst, int64_t n,
                             const int* __restrict__ gate) {
    if (gate && *gate) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Status/iteration/ZNCC defaults of the reference batch wrappers (numba_kernels.py:832-840).
__global__ void k_init_outputs(const uint8_t* __restrict__ valid, int64_t N, int32_t* n_iter, int8_t* status,
                               double* zncc, const int* __restrict__ gate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || (gate && *gate)) return;
    n_iter[i] = 0;
    status[i] = valid[i] ? ST_SKIPPED : ST_INVALID_SUBSET;
    zncc[i] = NAN;
}

// Brick tiles along one axis: a subset whose span plus the interpolation stencil and a voxel of
// slack fits `room` brick floats is one tile; a wider one is cut into equal tiles of at most
// (room - 5) / stride + 1 voxels, rounded up to whole 8-voxel patches where that still fits.
int brick_tile(int n, int stride, int room) {
    const int kmax = (room - 5) / stride + 1;
    if (n <= kmax) return n;
    const int nt = (n + kmax - 1) / kmax;
    const int t = (n + nt - 1) / nt, t8 = (t + 7) & ~7;
    return t8 <= kmax ? t8 : t;
}

// Workgroups of a persistent launch of `kern` with `lds` bytes of dynamic LDS: as many as are
// resident on the device at once (a multiple of 8, so workgroup b stays on XCD b % 8), at most nb;
// nb when the occupancy cannot be read (the node counters make any grid correct).
int64_t persistent_grid(const void* kern, size_t lds, int64_t nb) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, BLOCK, lds) != hipSuccess || per < 1 || cus < 1)
        return nb;
    const int64_t g = (int64_t)per * cus / 8 * 8;
    return std::max<int64_t>(8, std::min<int64_t>(g, nb));
}

// The fused iteration kernel: LDS bricks (variant 1, when the volume is addressable with 32-bit
// byte offsets), else memory gathers.
template <int MODE, bool STORED, bool SEL, bool T3 = false>
void launch_fused(int variant, const IcgnArgs& a, const int64_t* list, const int64_t* count, int64_t nb, size_t bits,
                  bool buf, double* mom, int lds_budget, int flags, hipStream_t st) {
    // LDS brick: BRICK_XY^2 floats per plane, as many planes as the LDS budget holds (by default
    // three workgroups per CU); subsets wider than the brick in x/y boxes of tile_x x tile_y voxel
    // columns. Pair mapping (stride 1, one box, rows of 32..64 voxels): 39 rows per plane and a
    // plane stride of sx modulo 64 floats; otherwise 40 rows and a plane of 40 x 40.
    const Subset& sb = a.sub;
    const int tile_x = brick_tile(sb.sx, sb.stride, BRICK_XY), tile_y = brick_tile(sb.sy, sb.stride, BRICK_XY);
    const bool whole = tile_x == sb.sx && tile_y == sb.sy;
    const bool patches = !(flags & TUNE_NO_PATCHES) && sb.stride == 1 && sb.sx >= 8 && sb.sy >= 8;
    const bool pair = !patches && whole && !(flags & TUNE_NO_PAIRS) && sb.stride == 1 && sb.sx >= 32 &&
                      sb.sx <= 64 && sb.sy >= NWAVES && 2 * sb.hy + 5 <= BRICK_XY - 1;
    const int rows = pair ? BRICK_XY - 1 : BRICK_XY;
    const int plane = pair ? BRICK_XY * rows + ((sb.sx - BRICK_XY * rows) % 64 + 64) % 64 : BRICK_XY * rows;
    int planes = 0;
    // whether the warped stencils of a chunk fit is decided per chunk on the device
    if (variant == 1 && buf) {
        hipFuncAttributes fa{};
        if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_icgn12_fused<MODE, STORED, true, BRICK_XY, SEL, T3>)) ==
            hipSuccess) {
            const int64_t budget = lds_budget > 0 ? lds_budget : LDS_PER_WG;
            const int64_t room = budget - (int64_t)fa.sharedSizeBytes - (int64_t)bits;
            planes = (int)std::min<int64_t>(sb.sz + 3, room / (int64_t)(4 * plane));
        }
        if (planes < 6) planes = 0;  // too few planes per chunk: the staging would dominate
    }
    const size_t lds = bits + (size_t)planes * 4 * plane;
    auto kern = &k_icgn12_fused<MODE, STORED, true, BRICK_XY, SEL, T3>;
    IcgnArgs ka = a;
    const KernelTimer& tm = timer();
    ka.wg_clock = (tm.wg_clock && tm.wg_capacity >= nb) ? tm.wg_clock : nullptr;
    if (planes > 0 && hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        planes = 0;
    auto kgb = &k_icgn12_fused<MODE, STORED, true, 0, SEL, T3>;
    auto kgp = &k_icgn12_fused<MODE, STORED, false, 0, SEL, T3>;
    const int64_t grid = planes > 0 ? persistent_grid(reinterpret_cast<const void*>(kern), lds, nb)
                                    : persistent_grid(reinterpret_cast<const void*>(buf ? kgb : kgp), bits, nb);
    timer_begin(st);
    if (planes > 0) {
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BLOCK), lds, st, ka, list, count, (const double*)mom, planes,
                           plane, rows, patches ? LANES_PATCHES : (pair ? LANES_PAIRS : LANES_WALK), tile_x, tile_y);
    } else if (buf) {
        hipLaunchKernelGGL(kgb, dim3((unsigned)grid), dim3(BLOCK), bits, st, ka, list, count, (const double*)mom, 0, 0, 0, 0,
                           0, 0);
    } else {
        hipLaunchKernelGGL(kgp, dim3((unsigned)grid), dim3(BLOCK), bits, st, ka, list, count, (const double*)mom, 0, 0, 0, 0,
                           0, 0);
    }
}

// Reference moments of every listed node; `sel`: a mask or keep rows select the subset voxels.
// `rm`: buffers of the separable pass (null: per node only); it runs for stored gradients at stride
// 1 and takes the nodes without keep rows, and the per-node pass then takes only those with keep
// rows, or every node when the rectangle sums did not fit.
template <bool STORED>
void launch_ref_moments(bool sel, const IcgnArgs& a, const int64_t* list, const int64_t* count, int64_t N, int64_t nb,
                        double* mom, const RowMoments* rm, hipStream_t st) {
    const int* skip = nullptr;
    if (STORED && rm && a.sub.stride == 1) {
        const int nx = a.V.nx, ny = a.V.ny;
        const unsigned nodes = (unsigned)std::min<int64_t>((N + 255) / 256, 4096);
        hipLaunchKernelGGL(k_rs_clear, dim3((unsigned)((std::max(std::max(nx, ny), (int)RS_HDR) + 255) / 256)),
                           dim3(256), 0, st, *rm, nx, ny);
        hipLaunchKernelGGL(k_rs_mark, dim3(nodes), dim3(256), 0, st, a, list, count, (int64_t)0, *rm);
        hipLaunchKernelGGL(k_rs_index, dim3(1), dim3(1024), 0, st, *rm, nx, ny, a.sub.hz, RS_Q);
        if (a.V.mask) hipLaunchKernelGGL(k_rs_xy<true>, dim3(2048), dim3(BLOCK), 0, st, a.V, a.sub, *rm);
        else hipLaunchKernelGGL(k_rs_xy<false>, dim3(2048), dim3(BLOCK), 0, st, a.V, a.sub, *rm);
        hipLaunchKernelGGL(k_rs_z, dim3(nodes), dim3(BLOCK), 0, st, a, list, count, *rm, mom);
        skip = rm->hdr + RS_FITS;
    }
    if (sel)
        hipLaunchKernelGGL((k_ref_moments<STORED, true>), dim3((unsigned)nb), dim3(BLOCK), 0, st, a, list, count, mom,
                           skip);
    else
        hipLaunchKernelGGL((k_ref_moments<STORED, false>), dim3((unsigned)nb), dim3(BLOCK), 0, st, a, list, count, mom,
                           skip);
}

// The valid-node count lives on the device (written by k_compact); grids are sized for all N
// nodes and the workgroups past the count exit at once, so no host round trip is needed.
template <int MODE, bool STORED>
void launch_icgn12(int variant, const IcgnArgs& a, const int64_t* list, const int64_t* count, int64_t N, float* gbuf,
                   int lds_budget, int flags, const RowMoments* rm, hipStream_t st) {
    if (variant == 0) {
        for (int64_t off = 0; off < N; off += PLAIN_CHUNK) {
            const int64_t cnt = std::min<int64_t>(PLAIN_CHUNK, N - off);
            hipLaunchKernelGGL((k_icgn12_plain<MODE, STORED>), dim3((unsigned)cnt), dim3(BLOCK), 0, st, a, list, count,
                               off, gbuf);
        }
    } else {
        const int64_t nb = xcd_grid(N);
        double* mom = (double*)gbuf;  // variant 1 keeps the reference moments where variant 0 keeps scratch rows
        const bool sel = a.V.mask != nullptr || (a.split_index != nullptr && a.split_keep != nullptr);
        launch_ref_moments<STORED>(sel, a, list, count, N, nb, mom, rm, st);
        const size_t bits = sizeof(unsigned) * (size_t)(((a.sub.S + 31) / 32 + 3) & ~3);
        const bool buf = (int64_t)a.V.nx * a.V.ny * a.V.nz < ((int64_t)1 << 30);  // 32-bit byte offsets
        if (sel) launch_fused<MODE, STORED, true>(variant, a, list, count, nb, bits, buf, mom, lds_budget, flags, st);
        else launch_fused<MODE, STORED, false>(variant, a, list, count, nb, bits, buf, mom, lds_budget, flags, st);
        timer_end(st);
    }
}

// Translation-only solves: reference moments, then the fused kernel in its T3 form (LDS bricks when
// the subset's warped rows fit one, else buffer or pointer gathers).
template <int MODE, bool STORED>
void launch_icgn3(int variant, const IcgnArgs& a, const int64_t* list, const int64_t* count, int64_t N, double* mom,
                  int lds_budget, int flags, const RowMoments* rm, hipStream_t st) {
    const int64_t nb = xcd_grid(N);
    const bool sel = a.V.mask != nullptr || (a.split_index != nullptr && a.split_keep != nullptr);
    if (!(flags & ICGN3_REUSE_MOMENTS)) launch_ref_moments<STORED>(sel, a, list, count, N, nb, mom, rm, st);
    const size_t bits = sizeof(unsigned) * (size_t)(((a.sub.S + 31) / 32 + 3) & ~3);
    const bool buf = (int64_t)a.V.nx * a.V.ny * a.V.nz < ((int64_t)1 << 30);
    if (sel) launch_fused<MODE, STORED, true, true>(variant, a, list, count, nb, bits, buf, mom, lds_budget, flags, st);
    else launch_fused<MODE, STORED, false, true>(variant, a, list, count, nb, bits, buf, mom, lds_budget, flags, st);
    timer_end(st);
}

// Valid-node list: in index order (the reference's order, used by the plain kernel), or grouped
// by tiles one subset wide (the fused kernels). Tiles grow until the histogram fits in LDS.
void launch_compact(bool tiled, const uint8_t* valid, const int64_t* coords, int64_t N, int hx, int hy, int hz,
                    const Volumes& V, int64_t* list, int64_t* count, hipStream_t stream) {
    if (!tiled || N > UINT32_MAX) {
        hipLaunchKernelGGL(k_compact, dim3(1), dim3(1024), 0, stream, valid, N, list, count);
        return;
    }
    int tx = std::max(8, 2 * hx), ty = std::max(8, 2 * hy), tz = std::max(8, 2 * hz);
    auto tiles = [&](int t, int n) { return (n + t - 1) / t; };
    while ((int64_t)tiles(tx, V.nx) * tiles(ty, V.ny) * tiles(tz, V.nz) > TILE_HIST_MAX) {
        tx *= 2;
        ty *= 2;
        tz *= 2;
    }
    hipLaunchKernelGGL(k_compact_tiled, dim3(1), dim3(1024), 0, stream, valid, coords, N, tx, ty, tz,
                       tiles(tx, V.nx), tiles(ty, V.ny), tiles(tz, V.nz), list, count);
}

int check_common(const Volumes& V, const Subset& s, int mode) {
    if (V.nz < 4 || V.ny < 4 || V.nx < 4) return fail(-2, "volumes must be at least 4 voxels along every axis");
    if ((int64_t)V.nx * V.ny * V.nz > INT32_MAX)
        return fail(-2, "volumes beyond 2^31 voxels must be tiled (32-bit sample offsets)");
    if (s.hx < 0 || s.hy < 0 || s.hz < 0 || s.stride < 1) return fail(-2, "bad subset half-widths or stride");
    if (mode < 0 || mode > 2) return fail(-2, "interp mode must be 0 (linear), 1 (cubic) or 2 (bspline)");
    if (!V.f || !V.g) return fail(-2, "f and g are required");
    if ((V.gx == nullptr) != (V.gy == nullptr) || (V.gx == nullptr) != (V.gz == nullptr))
        return fail(-2, "gx, gy, gz must be all given or all null");
    return 0;
}

// ---------------------------------------------------------------------------- no-alloc scratch
// Workspace of aldvc_icgn12_noalloc: a fixed device array, one per GPU the library runs on, held
// until the process ends. In host code `scratch` names a host-side stand-in, never device memory:
// its device address is looked up once per device (with that device current) and only that
// address is handed to kernels.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_dev[NOALLOC_MAX_DEVICES] = {};

// Device address of the current device's scratch array (looked up on the first call per device).
int scratch_address(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return fail(-5, "hipGetDevice failed");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(-2, "device index beyond the no-alloc scratch table");
    std::lock_guard<std::mutex> guard(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(-5, (std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e)).c_str());
        scratch_dev[dev] = p;
    }
    *out = scratch_dev[dev];
    return 0;
}

}  // namespace

// ============================================================================ C ABI
extern "C" {

int aldvc_abi_version(void) { return ALDVC_ABI_VERSION; }

const char* aldvc_last_error(void) { return last_error_cstr(); }

int aldvc_kernel_timing(int on) {
    clear_error();
    KernelTimer& t = timer();
    if (on && !t.a) {
        if (hipEventCreate(&t.a) != hipSuccess || hipEventCreate(&t.b) != hipSuccess)
            return fail(-5, "hipEventCreate failed");
    }
    t.on = on != 0;
    t.have = false;
    return 0;
}

int aldvc_workgroup_clock(void* buf, int64_t capacity, int* clock_khz) {
    clear_error();
    KernelTimer& t = timer();
    if (buf && capacity < 1) return fail(-2, "capacity must be positive");
    t.wg_clock = static_cast<unsigned long long*>(buf);
    t.wg_capacity = buf ? capacity : 0;
    if (clock_khz) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(clock_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess)
            return fail(-5, "the wall-clock rate could not be read");
    }
    return 0;
}

float aldvc_last_kernel_ms(void) {
    KernelTimer& t = timer();
    float ms = -1.0f;
    if (!t.have || hipEventSynchronize(t.b) != hipSuccess || hipEventElapsedTime(&ms, t.a, t.b) != hipSuccess)
        return -1.0f;
    return ms;
}

int aldvc_gradients(const float* f, int64_t nz, int64_t ny, int64_t nx, int border, float* gx, float* gy, float* gz,
                    hipStream_t stream) {
    clear_error();
    if (!f || !gx || !gy || !gz) return fail(-2, "null pointer");
    if (nz < 1 || ny < 1 || nx < 1 || ny > 65535 || nz > 65535) return fail(-2, "bad volume shape");
    dim3 grid((unsigned)((nx + 255) / 256), (unsigned)ny, (unsigned)nz);
    hipLaunchKernelGGL(k_gradients, grid, dim3(256), 0, stream, f, (int)nz, (int)ny, (int)nx, border, gx, gy, gz);
    return launch_status();
}

int aldvc_sample_points(const float* vol, int64_t nz, int64_t ny, int64_t nx, const double* zs, const double* ys,
                        const double* xs, int64_t n, int mode, float* out, hipStream_t stream) {
    clear_error();
    if (!vol || !zs || !ys || !xs || !out) return fail(-2, "null pointer");
    if (nz < 4 || ny < 4 || nx < 4) return fail(-2, "volume too small");
    if (n <= 0) return 0;
    const unsigned nb = (unsigned)((n + 255) / 256);
    switch (mode) {
        case 0: hipLaunchKernelGGL(k_sample_points<0>, dim3(nb), dim3(256), 0, stream, vol, (int)nz, (int)ny, (int)nx, zs, ys, xs, n, out); break;
        case 1: hipLaunchKernelGGL(k_sample_points<1>, dim3(nb), dim3(256), 0, stream, vol, (int)nz, (int)ny, (int)nx, zs, ys, xs, n, out); break;
        case 2: hipLaunchKernelGGL(k_sample_points<2>, dim3(nb), dim3(256), 0, stream, vol, (int)nz, (int)ny, (int)nx, zs, ys, xs, n, out); break;
        default: return fail(-2, "bad interp mode");
    }
    return launch_status();
}

size_t aldvc_precompute_row_bytes(int64_t ncx, int64_t ncy, int64_t nplanes, int64_t ny, int64_t nx) {
    if (ncx <= 0 || ncy <= 0 || nplanes <= 0) return 0;
    return row_moments_fixed_bytes(ny, nx) + sizeof(double) * PC_Q * (size_t)nplanes * (size_t)ncy * (size_t)ncx;
}

int aldvc_precompute_tuned(const int64_t* coords, int64_t N, int hx, int hy, int hz, int stride, const float* f,
                           const float* gx, const float* gy, const float* gz, const uint8_t* mask, int64_t nz,
                           int64_t ny, int64_t nx, double min_valid_ratio, double cond_max, const int64_t* split_index,
                           const uint8_t* split_keep, int64_t keep_bytes, double* H_all, double* L_all, double* meanf,
                           double* bottomf, int64_t* n_valid, uint8_t* valid, int flags, void* workspace,
                           size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    Volumes V{f, gx, gy, gz, f, mask, (int)nz, (int)ny, (int)nx};
    const Subset s = make_subset(hx, hy, hz, stride);
    if (int e = check_common(V, s, 1)) return e;
    if (flags & ~TUNE_ROW_MOMENTS) return fail(-2, "unknown tuning flags");
    if (N < 0) return fail(-2, "negative node count");
    if (N == 0) return 0;
    if (!coords || !H_all || !L_all || !meanf || !bottomf || !n_valid || !valid) return fail(-2, "null pointer");
    if (split_keep && keep_bytes < (s.S + 7) / 8) return fail(-2, "keep rows shorter than ceil(S/8) bytes");
    if (N > INT32_MAX) return fail(-2, "too many nodes for one launch");
    const int* skip = nullptr;
    const size_t fixed = row_moments_fixed_bytes(ny, nx);
    if ((flags & TUNE_ROW_MOMENTS) && gx && stride == 1 && workspace && workspace_bytes >= fixed + PC_Q * sizeof(double)) {
        // the separable pass for the nodes without keep rows inside the volume (the per-node kernel
        // then takes the others, or every node when the rectangle sums do not fit)
        RowMoments rm;
        char* p = (char*)workspace;
        rm.hdr = (int*)p;
        rm.xidx = rm.hdr + RS_HDR;
        rm.yidx = rm.xidx + nx;
        rm.cxs = rm.yidx + ny;
        rm.cys = rm.cxs + nx;
        rm.R = (double*)(p + fixed);
        rm.r_doubles = (int64_t)((workspace_bytes - fixed) / sizeof(double));
        IcgnArgs a{};
        a.V = V;
        a.sub = s;
        a.coords = coords;
        a.split_index = split_index;
        a.split_keep = split_keep;
        a.keep_bytes = keep_bytes;
        const unsigned nodes = (unsigned)std::min<int64_t>((N + 255) / 256, 4096);
        hipLaunchKernelGGL(k_rs_clear, dim3((unsigned)((std::max(std::max<int64_t>(nx, ny), (int64_t)RS_HDR) + 255) / 256)),
                           dim3(256), 0, stream, rm, (int)nx, (int)ny);
        hipLaunchKernelGGL(k_rs_mark, dim3(nodes), dim3(256), 0, stream, a, (const int64_t*)nullptr,
                           (const int64_t*)nullptr, N, rm);
        hipLaunchKernelGGL(k_rs_index, dim3(1), dim3(1024), 0, stream, rm, (int)nx, (int)ny, hz, PC_Q);
        if (mask) hipLaunchKernelGGL(k_pc_xy<true>, dim3(2048), dim3(BLOCK), 0, stream, V, s, rm);
        else hipLaunchKernelGGL(k_pc_xy<false>, dim3(2048), dim3(BLOCK), 0, stream, V, s, rm);
        hipLaunchKernelGGL(k_pc_zfin, dim3((unsigned)N), dim3(64), 0, stream, V, s, coords, N, split_index, split_keep,
                           rm, min_valid_ratio, cond_max, H_all, L_all, meanf, bottomf, n_valid, valid);
        skip = rm.hdr + RS_FITS;
    }
    if (gx)
        hipLaunchKernelGGL(k_precompute<true>, dim3((unsigned)N), dim3(BLOCK), 0, stream, V, s, coords, N,
                           min_valid_ratio, cond_max, split_index, split_keep, keep_bytes, H_all, L_all, meanf,
                           bottomf, n_valid, valid, skip);
    else
        hipLaunchKernelGGL(k_precompute<false>, dim3((unsigned)N), dim3(BLOCK), 0, stream, V, s, coords, N,
                           min_valid_ratio, cond_max, split_index, split_keep, keep_bytes, H_all, L_all, meanf,
                           bottomf, n_valid, valid, skip);
    return launch_status();
}

int aldvc_precompute(const int64_t* coords, int64_t N, int hx, int hy, int hz, int stride, const float* f,
                     const float* gx, const float* gy, const float* gz, const uint8_t* mask, int64_t nz, int64_t ny,
                     int64_t nx, double min_valid_ratio, double cond_max, const int64_t* split_index,
                     const uint8_t* split_keep, int64_t keep_bytes, double* H_all, double* L_all, double* meanf,
                     double* bottomf, int64_t* n_valid, uint8_t* valid, hipStream_t stream) {
    return aldvc_precompute_tuned(coords, N, hx, hy, hz, stride, f, gx, gy, gz, mask, nz, ny, nx, min_valid_ratio,
                                  cond_max, split_index, split_keep, keep_bytes, H_all, L_all, meanf, bottomf,
                                  n_valid, valid, 0, nullptr, 0, stream);
}

}  // extern "C"

namespace {
// Buffers of the separable reference-moment pass in the workspace past the `base` bytes every call
// needs, when the call asks for it (TUNE_ROW_MOMENTS), the fused path runs on stored gradients at
// stride 1, and the workspace holds more than the fixed part.
const RowMoments* row_moments_at(RowMoments& rm, int flags, int variant, const Volumes& V, const Subset& s,
                                 void* ws, size_t ws_bytes, size_t base) {
    if (!(flags & TUNE_ROW_MOMENTS) || variant == 0 || !V.gx || s.stride != 1) return nullptr;
    const size_t fixed = row_moments_fixed_bytes(V.ny, V.nx);
    if (ws_bytes < base + fixed + RS_Q * sizeof(double)) return nullptr;
    char* p = (char*)ws + base;
    rm.hdr = (int*)p;
    rm.xidx = rm.hdr + RS_HDR;
    rm.yidx = rm.xidx + V.nx;
    rm.cxs = rm.yidx + V.ny;
    rm.cys = rm.cxs + V.nx;
    rm.R = (double*)(p + fixed);
    rm.r_doubles = (int64_t)((ws_bytes - base - fixed) / sizeof(double));
    return &rm;
}
}  // namespace

extern "C" {

size_t aldvc_icgn12_row_moments_bytes(int64_t ncx, int64_t ncy, int64_t nplanes, int64_t ny, int64_t nx) {
    if (ncx <= 0 || ncy <= 0 || nplanes <= 0) return 0;
    return row_moments_fixed_bytes(ny, nx) + sizeof(double) * RS_Q * (size_t)nplanes * (size_t)ncy * (size_t)ncx;
}

size_t aldvc_icgn12_workspace_bytes(int64_t N, int hx, int hy, int hz, int stride, int variant) {
    const Subset s = make_subset(hx, hy, hz, stride);
    size_t bytes = 256 + sizeof(int64_t) * (size_t)(N > 0 ? N : 1);  // count + node list
    bytes = (bytes + 255) / 256 * 256;
    if (variant == 0) bytes += sizeof(float) * (size_t)std::min<int64_t>(N, PLAIN_CHUNK) * (size_t)s.S;
    else bytes += sizeof(double) * MOM_STRIDE * (size_t)(N > 0 ? N : 1);  // reference moments per node
    return bytes;
}

int aldvc_icgn12_tuned(const int64_t* coords, int64_t N, double* P, int hx, int hy, int hz, int stride,
                       const float* f, const float* gx, const float* gy, const float* gz, const uint8_t* mask,
                       const float* g, int64_t nz, int64_t ny, int64_t nx, int mode, const double* L_all,
                       const double* H_all, const double* pattern, double noise_gain, const double* meanf,
                       const uint8_t* valid, double tol, double dp_tol, int max_iter, int patience, int predictive,
                       const int64_t* split_index, const uint8_t* split_keep, int64_t keep_bytes, int32_t* n_iter,
                       int8_t* status, double* zncc, int variant, int lds_budget, int flags, void* workspace,
                       size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    if (lds_budget < 0 || lds_budget > 160 * 1024) return fail(-2, "lds_budget must be 0 or at most 160 KiB");
    if (flags & ~(TUNE_NO_PAIRS | TUNE_INDEX_ORDER | TUNE_NO_XCD | TUNE_NO_PATCHES | TUNE_ROW_MOMENTS))
        return fail(-2, "unknown tuning flags");
    Volumes V{f, gx, gy, gz, g, mask, (int)nz, (int)ny, (int)nx};
    const Subset s = make_subset(hx, hy, hz, stride);
    if (int e = check_common(V, s, mode)) return e;
    if (variant < 0 || variant > 2)
        return fail(-2, "variant must be 0 (plain), 1 (fused) or 2 (fused, gathers only)");
    if (N < 0) return fail(-2, "negative node count");
    if (N == 0) return 0;
    if (!coords || !P || !L_all || !valid || !n_iter || !status || !zncc) return fail(-2, "null pointer");
    if (N > INT32_MAX) return fail(-2, "too many nodes for one launch");
    if (split_keep && keep_bytes < (s.S + 7) / 8) return fail(-2, "keep rows shorter than ceil(S/8) bytes");
    if (!(H_all && pattern && noise_gain > 0.0)) {
        H_all = nullptr;
        pattern = nullptr;
        noise_gain = 0.0;
    }
    if (variant != 0 && sizeof(unsigned) * (size_t)((s.S + 31) / 32) > 48 * 1024)
        return fail(-2, "subset too large for the fused kernel's LDS voxel mask (use variant 0)");
    const size_t need = aldvc_icgn12_workspace_bytes(N, hx, hy, hz, stride, variant);
    if (!workspace || workspace_bytes < need) return fail(-3, "workspace too small (see aldvc_icgn12_workspace_bytes)");
    int64_t* count = (int64_t*)workspace;
    int64_t* list = (int64_t*)((char*)workspace + 256);
    size_t list_bytes = (256 + sizeof(int64_t) * (size_t)N + 255) / 256 * 256;
    float* gbuf = (float*)((char*)workspace + list_bytes);
    hipLaunchKernelGGL(k_init_outputs, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, stream, valid, N, n_iter,
                       status, zncc, (const int*)nullptr);
    launch_compact(variant != 0 && !(flags & TUNE_INDEX_ORDER), valid, coords, N, hx, hy, hz, V, list, count, stream);
    IcgnArgs a{V, s, coords, P, L_all, H_all, pattern, meanf, split_index, split_keep, keep_bytes,
               noise_gain, tol, dp_tol, max_iter, patience, predictive, n_iter, status, zncc,
               (flags & TUNE_NO_XCD) ? 0 : 1};
    RowMoments rm;
    const RowMoments* rmp = row_moments_at(rm, flags, variant, V, s, workspace, workspace_bytes, need);
    const bool stored = gx != nullptr;
#define ALDVC_DISPATCH(MD)                                                                                \
    if (stored) launch_icgn12<MD, true>(variant, a, list, count, N, gbuf, lds_budget, flags, rmp, stream);    \
    else launch_icgn12<MD, false>(variant, a, list, count, N, gbuf, lds_budget, flags, rmp, stream);
    switch (mode) {
        case 0: ALDVC_DISPATCH(0) break;
        case 1: ALDVC_DISPATCH(1) break;
        default: ALDVC_DISPATCH(2) break;
    }
#undef ALDVC_DISPATCH
    return launch_status();
}

int aldvc_icgn12(const int64_t* coords, int64_t N, double* P, int hx, int hy, int hz, int stride, const float* f,
                 const float* gx, const float* gy, const float* gz, const uint8_t* mask, const float* g, int64_t nz,
                 int64_t ny, int64_t nx, int mode, const double* L_all, const double* H_all, const double* pattern,
                 double noise_gain, const double* meanf, const uint8_t* valid, double tol, double dp_tol,
                 int max_iter, int patience, int predictive, const int64_t* split_index, const uint8_t* split_keep,
                 int64_t keep_bytes, int32_t* n_iter, int8_t* status, double* zncc, int variant, void* workspace,
                 size_t workspace_bytes, hipStream_t stream) {
    return aldvc_icgn12_tuned(coords, N, P, hx, hy, hz, stride, f, gx, gy, gz, mask, g, nz, ny, nx, mode, L_all, H_all,
                              pattern, noise_gain, meanf, valid, tol, dp_tol, max_iter, patience, predictive,
                              split_index, split_keep, keep_bytes, n_iter, status, zncc, variant, 0, 0, workspace,
                              workspace_bytes, stream);
}

int aldvc_icgn12_noalloc(const int64_t* coords, int64_t N, double* P, int hx, int hy, int hz, int stride,
                         const float* f, const float* gx, const float* gy, const float* gz, const uint8_t* mask,
                         const float* g, int64_t nz, int64_t ny, int64_t nx, int mode, const double* L_all,
                         const double* H_all, const double* pattern, double noise_gain, const double* meanf,
                         const uint8_t* valid, double tol, double dp_tol, int max_iter, int patience, int predictive,
                         const int64_t* split_index, const uint8_t* split_keep, int64_t keep_bytes,
                         int32_t* n_iter, int8_t* status, double* zncc, int variant, hipStream_t stream) {
    clear_error();
    const size_t need = aldvc_icgn12_workspace_bytes(N, hx, hy, hz, stride, variant);
    if (need > NOALLOC_SCRATCH_BYTES) {
        const std::string msg = "aldvc_icgn12_noalloc: " + std::to_string(N) + " nodes need " + std::to_string(need) +
                                " bytes of workspace, beyond its fixed " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                                "-byte scratch (at most 254,199 nodes for the fused variants); "
                                "call aldvc_icgn12 with a workspace";
        return fail(-2, msg.c_str());
    }
    if (N <= 0) return N < 0 ? fail(-2, "negative node count") : 0;
    void* ws = nullptr;
    if (int e = scratch_address(&ws)) return e;
    return aldvc_icgn12_tuned(coords, N, P, hx, hy, hz, stride, f, gx, gy, gz, mask, g, nz, ny, nx, mode, L_all, H_all,
                              pattern, noise_gain, meanf, valid, tol, dp_tol, max_iter, patience, predictive,
                              split_index, split_keep, keep_bytes, n_iter, status, zncc, variant, 0, 0, ws,
                              NOALLOC_SCRATCH_BYTES, stream);
}

size_t aldvc_icgn3_workspace_bytes(int64_t N) { return aldvc_icgn12_workspace_bytes(N, 0, 0, 0, 1, 1); }

int aldvc_icgn3_tuned(const int64_t* coords, int64_t N, const double* U_old, double* U, const double* F_fixed,
                      const double* vdual, int hx, int hy, int hz, int stride, const float* f, const float* gx,
                      const float* gy, const float* gz, const uint8_t* mask, const float* g, int64_t nz, int64_t ny,
                      int64_t nx, int mode, const double* H_all, const double* meanf, const double* bottomf,
                      const uint8_t* valid, double mu, double tol, double dp_tol, int max_iter, int patience,
                      double noise_gain, int predictive, const int64_t* split_index, const uint8_t* split_keep,
                      int64_t keep_bytes, int32_t* n_iter, int8_t* status, double* zncc, const int* gate,
                      int variant, int lds_budget, int flags, void* workspace, size_t workspace_bytes,
                      hipStream_t stream) {
    clear_error();
    if (variant != 1 && variant != 2) return fail(-2, "variant must be 1 (LDS bricks where they fit) or 2 (gathers only)");
    if (lds_budget < 0 || lds_budget > 160 * 1024) return fail(-2, "lds_budget must be 0 or at most 160 KiB");
    if (flags & ~(TUNE_NO_PAIRS | TUNE_INDEX_ORDER | TUNE_NO_XCD | TUNE_NO_PATCHES | ICGN3_REUSE_MOMENTS |
                  TUNE_ROW_MOMENTS))
        return fail(-2, "unknown tuning flags");
    Volumes V{f, gx, gy, gz, g, mask, (int)nz, (int)ny, (int)nx};
    const Subset s = make_subset(hx, hy, hz, stride);
    if (int e = check_common(V, s, mode)) return e;
    if (N < 0) return fail(-2, "negative node count");
    if (N == 0) return 0;
    if (!coords || !U_old || !U || !F_fixed || !vdual || !H_all || !bottomf || !valid || !n_iter || !status || !zncc)
        return fail(-2, "null pointer");
    if (N > INT32_MAX) return fail(-2, "too many nodes for one launch");
    if (split_keep && keep_bytes < (s.S + 7) / 8) return fail(-2, "keep rows shorter than ceil(S/8) bytes");
    if (sizeof(unsigned) * (size_t)((s.S + 31) / 32) > 48 * 1024)
        return fail(-2, "subset too large for the fused kernel's LDS voxel mask");
    const size_t need = aldvc_icgn3_workspace_bytes(N);
    if (!workspace || workspace_bytes < need) return fail(-3, "workspace too small (see aldvc_icgn3_workspace_bytes)");
    int64_t* count = (int64_t*)workspace;
    int64_t* list = (int64_t*)((char*)workspace + 256);
    if (gate) {
        hipLaunchKernelGGL(k_copy_gated, dim3((unsigned)std::min<int64_t>((3 * N + 255) / 256, 4096)), dim3(256), 0,
                           stream, U_old, U, 3 * N, gate);
    } else if (hipMemcpyAsync(U, U_old, sizeof(double) * 3 * (size_t)N, hipMemcpyDeviceToDevice, stream) != hipSuccess) {
        return fail(-4, "copy of U_old failed");
    }
    hipLaunchKernelGGL(k_init_outputs, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, stream, valid, N, n_iter,
                       status, zncc, gate);
    launch_compact(!(flags & TUNE_INDEX_ORDER), valid, coords, N, hx, hy, hz, V, list, count, stream);
    if (gate) hipLaunchKernelGGL(k_gate_count, dim3(1), dim3(64), 0, stream, count, gate);
    size_t list_bytes = (256 + sizeof(int64_t) * (size_t)N + 255) / 256 * 256;
    double* mom = (double*)((char*)workspace + list_bytes);
    IcgnArgs a{};
    a.V = V;
    a.sub = s;
    a.coords = coords;
    a.H_all = H_all;
    a.meanf = meanf;
    a.split_index = split_index;
    a.split_keep = split_keep;
    a.keep_bytes = keep_bytes;
    a.noise_gain = noise_gain > 0.0 ? noise_gain : 0.0;
    a.tol = tol;
    a.dp_tol = dp_tol;
    a.max_iter = max_iter;
    a.patience = patience;
    a.predictive = predictive;
    a.n_iter = n_iter;
    a.status = status;
    a.zncc = zncc;
    a.xcd = (flags & TUNE_NO_XCD) ? 0 : 1;
    a.F = F_fixed;
    a.U_old = U_old;
    a.vdual = vdual;
    a.bottomf = bottomf;
    a.U = U;
    a.mu = mu;
    RowMoments rm;
    const RowMoments* rmp = row_moments_at(rm, flags, variant, V, s, workspace, workspace_bytes, need);
    const bool stored = gx != nullptr;
#define ALDVC_DISPATCH3(MD)                                                                           \
    if (stored) launch_icgn3<MD, true>(variant, a, list, count, N, mom, lds_budget, flags, rmp, stream);    \
    else launch_icgn3<MD, false>(variant, a, list, count, N, mom, lds_budget, flags, rmp, stream);
    switch (mode) {
        case 0: ALDVC_DISPATCH3(0) break;
        case 1: ALDVC_DISPATCH3(1) break;
        default: ALDVC_DISPATCH3(2) break;
    }
#undef ALDVC_DISPATCH3
    return launch_status();
}

int aldvc_icgn3_gated(const int64_t* coords, int64_t N, const double* U_old, double* U, const double* F_fixed,
                      const double* vdual, int hx, int hy, int hz, int stride, const float* f, const float* gx,
                      const float* gy, const float* gz, const uint8_t* mask, const float* g, int64_t nz, int64_t ny,
                      int64_t nx, int mode, const double* H_all, const double* meanf, const double* bottomf,
                      const uint8_t* valid, double mu, double tol, double dp_tol, int max_iter, int patience,
                      double noise_gain, int predictive, const int64_t* split_index, const uint8_t* split_keep,
                      int64_t keep_bytes, int32_t* n_iter, int8_t* status, double* zncc, const int* gate,
                      void* workspace, size_t workspace_bytes, hipStream_t stream) {
    return aldvc_icgn3_tuned(coords, N, U_old, U, F_fixed, vdual, hx, hy, hz, stride, f, gx, gy, gz, mask, g, nz, ny,
                             nx, mode, H_all, meanf, bottomf, valid, mu, tol, dp_tol, max_iter, patience, noise_gain,
                             predictive, split_index, split_keep, keep_bytes, n_iter, status, zncc, gate, 1, 0, 0,
                             workspace, workspace_bytes, stream);
}

int aldvc_icgn3(const int64_t* coords, int64_t N, const double* U_old, double* U, const double* F_fixed,
                const double* vdual, int hx, int hy, int hz, int stride, const float* f, const float* gx,
                const float* gy, const float* gz, const uint8_t* mask, const float* g, int64_t nz, int64_t ny,
                int64_t nx, int mode, const double* H_all, const double* meanf, const double* bottomf,
                const uint8_t* valid, double mu, double tol, double dp_tol, int max_iter, int patience,
                double noise_gain, int predictive, const int64_t* split_index, const uint8_t* split_keep,
                int64_t keep_bytes, int32_t* n_iter, int8_t* status, double* zncc, void* workspace,
                size_t workspace_bytes, hipStream_t stream) {
    return aldvc_icgn3_gated(coords, N, U_old, U, F_fixed, vdual, hx, hy, hz, stride, f, gx, gy, gz, mask, g, nz, ny,
                             nx, mode, H_all, meanf, bottomf, valid, mu, tol, dp_tol, max_iter, patience, noise_gain,
                             predictive, split_index, split_keep, keep_bytes, n_iter, status, zncc, nullptr, workspace,
                             workspace_bytes, stream);
}

int aldvc_evaluate_zncc(const int64_t* coords, int64_t N, const double* P_all, int hx, int hy, int hz, int stride,
                        const float* f, const uint8_t* mask, const float* g, int64_t nz, int64_t ny, int64_t nx,
                        int mode, const uint8_t* valid, double* out, hipStream_t stream) {
    clear_error();
    Volumes V{f, nullptr, nullptr, nullptr, g, mask, (int)nz, (int)ny, (int)nx};
    const Subset s = make_subset(hx, hy, hz, stride);
    if (int e = check_common(V, s, mode)) return e;
    if (N < 0) return fail(-2, "negative node count");
    if (N == 0) return 0;
    if (!coords || !P_all || !valid || !out) return fail(-2, "null pointer");
    if (N > INT32_MAX) return fail(-2, "too many nodes for one launch");
    switch (mode) {
        case 0: hipLaunchKernelGGL(k_evaluate_zncc<0>, dim3((unsigned)N), dim3(BLOCK), 0, stream, V, s, coords, N, P_all, valid, out); break;
        case 1: hipLaunchKernelGGL(k_evaluate_zncc<1>, dim3((unsigned)N), dim3(BLOCK), 0, stream, V, s, coords, N, P_all, valid, out); break;
        default: hipLaunchKernelGGL(k_evaluate_zncc<2>, dim3((unsigned)N), dim3(BLOCK), 0, stream, V, s, coords, N, P_all, valid, out); break;
    }
    return launch_status();
}

}  // extern "C"
