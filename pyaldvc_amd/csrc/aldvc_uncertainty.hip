// Per-node displacement uncertainty from the IC-GN normal equations (solver/uncertainty.py:83-133).
//
// At a converged node with residual variance s^2 = (1 - ZNCC) bottomf^2 / n_valid the covariance of
// the 12 warp parameters is
//     Cov(P) = sigma_r^2 bottomf^2 H0^-1 + c s^4 H0^-1 (scale D) H0^-1,   H0 = H - c s^2 scale D lim,
// with sigma_r^2 = 2 (1 - ZNCC) / (n_valid - 12), c the 7-point stencil's noise gain, D the
// noise pattern of the Hessian (diagonal: parameter p pairs gradient component and subset
// coordinate, and the subset moment matrix is diagonal), scale = n_valid / D[9][9], and lim <= 1
// the factor that keeps at least 10 % of every translation diagonal entry. The standard deviation
// of u, v, w is the square root of the translation diagonal.
//
// Only the columns 9, 10, 11 of H0^-1 are needed: H0 x = e_j for the three translations. One node
// per 16-lane group (four per wave64): lane r < 12 holds row r of [H0 | e9 e10 e11] in registers
// and Gauss-Jordan elimination with partial pivoting runs column by column, the pivot chosen by a
// max-|a| reduction over the group and its row broadcast by lane shuffles. No rows move: the lane
// that pivots column k ends holding x_k. An exactly zero pivot is reported (the caller then uses
// the pseudo-inverse for the whole batch, as the reference does when its batched inverse fails).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "aldvc_abi_util.h"

using namespace aldvc;

namespace {

constexpr int GROUP = 16;  // lanes per node
constexpr int BLOCK = 256;
constexpr int NP = 12;      // warp parameters
constexpr int NC = NP + 3;  // augmented columns
constexpr int STATUS_CONVERGED = 0;
constexpr int MIN_EXTRA_DOF = 2;
constexpr double MIN_CORRECTED_FRACTION = 0.1;

struct UncArgs {
    const double* H;  // (n, 12, 12)
    const double* zncc;
    const int8_t* status;  // may be null: every node counts as converged
    const uint8_t* valid;
    const int64_t* n_valid;
    const double* bottomf;
    int64_t n;
    double mx, my, mz, mn;  // diagonal of the subset moment matrix (X^2, Y^2, Z^2 sums, voxel count)
    double gain;            // stencil noise gain
    int noise_correction;
    double* out;   // (n, 3)
    int* singular;  // set to 1 when an exactly zero pivot occurs at a node that gets a value
};

// Noise pattern entry of parameter p: P = [F row-major, u, v, w]; F_ab pairs gradient component a
// with coordinate b, the translations pair with the constant 1.
__device__ inline double pattern_diag(const UncArgs& a, int p) {
    const int coord = p < 9 ? p % 3 : 3;
    return coord == 0 ? a.mx : (coord == 1 ? a.my : (coord == 2 ? a.mz : a.mn));
}

__device__ inline double group_sum(double v) {
#pragma unroll
    for (int m = 1; m < GROUP; m <<= 1) v += __shfl_xor(v, m, GROUP);
    return v;
}

__global__ __launch_bounds__(BLOCK) void k_displacement_uncertainty(UncArgs a) {
    const int64_t node = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) / GROUP;
    const int r = threadIdx.x % GROUP;
    const bool in = node < a.n;
    const int64_t nd = in ? node : a.n - 1;  // groups past the end compute a copy and write nothing
    const double z = a.zncc[nd];
    const int64_t nv = a.n_valid[nd];
    bool ok = in && a.valid[nd] && std::isfinite(z) && nv > NP + MIN_EXTRA_DOF;
    if (a.status) ok = ok && a.status[nd] == STATUS_CONVERGED;
    const double one_minus = fmax(1.0 - z, 0.0);
    const double bf = a.bottomf[nd];
    const double factor = 2.0 * one_minus / (double)(nv - NP) * bf * bf;
    const double s2 = one_minus * bf * bf / (double)nv;
    const double scale = (double)nv / a.mn;
    const double* Hn = a.H + nd * NP * NP;
    const bool row = r < NP;
    double A[NC];
#pragma unroll
    for (int c = 0; c < NP; ++c) A[c] = row ? Hn[r * NP + c] : 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) A[NP + j] = (r == 9 + j) ? 1.0 : 0.0;
    if (a.noise_correction) {
        // the translation diagonal of the correction is c s2 scale D[9][9] = c s2 n_valid for all three
        const double cdiag = a.gain * s2 * scale * a.mn;
        const double t0 = __shfl(A[9], 9, GROUP), t1 = __shfl(A[10], 10, GROUP), t2 = __shfl(A[11], 11, GROUP);
        double lim = 1.0;
        if (cdiag > 0.0) {
            const double d = fmax(cdiag, 1e-300);
            lim = fmin(fmin((1.0 - MIN_CORRECTED_FRACTION) * t0 / d, (1.0 - MIN_CORRECTED_FRACTION) * t1 / d),
                       (1.0 - MIN_CORRECTED_FRACTION) * t2 / d);
        }
        lim = fmin(fmax(lim, 0.0), 1.0);
        const double corr = a.gain * s2 * scale * lim;
#pragma unroll
        for (int c = 0; c < NP; ++c)
            if (c == r) A[c] -= corr * pattern_diag(a, c);
    }
    // Gauss-Jordan with partial pivoting; kcol = the column this lane's row pivoted
    int kcol = row ? -1 : NP;
    bool zero_pivot = false;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        double best = kcol < 0 ? fabs(A[k]) : -1.0;
        if (best != best) best = INFINITY;  // a NaN entry wins the reduction alike in every lane
        int who = r;
#pragma unroll
        for (int m = 1; m < GROUP; m <<= 1) {
            const double ob = __shfl_xor(best, m, GROUP);
            const int ow = __shfl_xor(who, m, GROUP);
            if (ob > best || (ob == best && ow < who)) {
                best = ob;
                who = ow;
            }
        }
        if (best == 0.0) zero_pivot = true;  // exactly singular
        double prow[NC];
#pragma unroll
        for (int c = k; c < NC; ++c) prow[c] = __shfl(A[c], who, GROUP);
        const double inv_p = 1.0 / prow[k];
        if (r == who) {
            kcol = k;
#pragma unroll
            for (int c = k; c < NC; ++c) A[c] *= inv_p;
        } else {
            const double fct = A[k] * inv_p;
#pragma unroll
            for (int c = k; c < NC; ++c) A[c] -= fct * prow[c];
        }
    }
    // lane with kcol = p holds column p of rows 9..11 of H0^-1 (= rows of columns 9..11)
    double var[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double x = (row && kcol >= 0) ? A[NP + j] : 0.0;
        const double diag = group_sum(kcol == 9 + j ? x : 0.0);
        var[j] = factor * diag;
        if (a.noise_correction) {
            const double cross = group_sum(row ? x * x * pattern_diag(a, kcol) : 0.0) * scale;
            var[j] += a.gain * s2 * s2 * cross;
        }
    }
    if (!in) return;
    if (r < 3) {
        const double v = r == 0 ? var[0] : (r == 1 ? var[1] : var[2]);
        a.out[node * 3 + r] = ok ? (v < 0.0 ? 0.0 : sqrt(v)) : NAN;  // NaN stays NaN
    }
    if (r == 0 && ok && zero_pivot) atomicOr(a.singular, 1);
}

}  // namespace

extern "C" {

int aldvc_displacement_uncertainty(const double* H, const double* zncc, const int8_t* status, const uint8_t* valid,
                                   const int64_t* n_valid, const double* bottomf, int64_t n, const double* moment_diag,
                                   double gain, int noise_correction, double* out, int* singular, hipStream_t stream) {
    clear_error();
    if (n < 0) return fail(-2, "negative node count");
    if (n == 0) return 0;
    if (!H || !zncc || !valid || !n_valid || !bottomf || !moment_diag || !out || !singular) return fail(-2, "null pointer");
    if (!(moment_diag[3] > 0.0)) return fail(-2, "the subset moment matrix needs a positive voxel count");
    const int64_t blocks = (n * GROUP + BLOCK - 1) / BLOCK;
    if (blocks > INT32_MAX) return fail(-2, "too many nodes for one launch");
    UncArgs a{H, zncc, status, valid, n_valid, bottomf, n, moment_diag[0], moment_diag[1], moment_diag[2],
              moment_diag[3], gain, noise_correction, out, singular};
    hipLaunchKernelGGL(k_displacement_uncertainty, dim3((unsigned)blocks), dim3(BLOCK), 0, stream, a);
    return launch_status();
}

}  // extern "C"
