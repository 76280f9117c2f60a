// Jacobi-preconditioned conjugate gradient for three right-hand sides, controlled on the device.
//
// The iteration of solver/subpb2_solver.py:77-105 (pcg_multi) and of the spring-model fill
// (utils/inpaint.py:100-126): every scalar of the iteration (<r, z>, alpha, beta, |r|^2, |b|)
// lives in a control block in device memory, and the stopping rule is evaluated by the device, so a
// solve is a fixed sequence of kernel launches with no host read-back. A solve can therefore be
// recorded into a HIP graph once and replayed; launches after convergence return at once.
//
// One iteration is two kernels, and no kernel waits on another workgroup: each writes per-block
// partial sums, and every block of the NEXT kernel adds all of them up in block order (a few KB
// read from L2), so every block derives the same scalars with no fence, atomic or extra launch:
//   iter1(k): from iter2(k-1)'s partials: |R|^2 -> converged? (then stop), <R, Z> -> beta;
//             P = Z + beta P_old (formed on the fly for every stencil neighbour, written for the own
//             node into the other P buffer), AP = A P, partials of <P, AP>
//   iter2(k): alpha = <r, z> / <P, AP> from iter1(k)'s partials; X += alpha P, R -= alpha AP,
//             Z = D^-1 R, partials of <R, Z> and <R, R>
// Block 0 keeps the control block (<r, z> by iteration parity, |B|, the stop flag and the
// iteration count) for the later kernels and the host. The convergence test of an iteration is
// made by the following iter1 (or by the closing kernel after the last slot), so the iteration
// count and the iterate equal those of the textbook loop that tests after every update.
//
// An operator supplies, per node, its three diagonal inverses and the stencil product:
//   __device__ void inv_diag(int64_t a, double d[3]) const;
//   template <class X> __device__ void apply(int64_t a, const X& x, double y[3]) const;
// where x(b, c) is the value of column c at node b.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace aldvc {
namespace pcg {

constexpr int BLOCK = 256;
constexpr int MAX_BLOCKS = 512;  // partials a block of the next kernel adds up

struct Ctl {
    double rz[2][3];  // <r, z> of the last two iterations, by parity
    double bnorm[3];  // |B| per column (1 for a zero column)
    int done;         // 1: converged, max_iter reached, or gated off; later launches return at once
    int iterations;   // iterations executed (valid once done)
};

inline unsigned blocks_for(int64_t n) {
    const int64_t b = (n + BLOCK - 1) / BLOCK;
    return (unsigned)(b < MAX_BLOCKS ? (b > 0 ? b : 1) : MAX_BLOCKS);
}

// Sum over the 64 lanes of a wave by an xor butterfly: partners add the same two values, so every
// lane ends with the same bits.
__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum of K doubles per thread over the block in a fixed order (each wave by the butterfly, then the
// waves in order); the total is in red[k][0].
template <int K>
__device__ inline void block_sum(const double v[K], double (*red)[BLOCK]) {
    constexpr int NW = BLOCK / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double w[K];
    for (int k = 0; k < K; ++k) w[k] = wave_sum(v[k]);
    if (lane == 0)
        for (int k = 0; k < K; ++k) red[k][1 + wave] = w[k];
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double t = red[threadIdx.x][1];
        for (int i = 1; i < NW; ++i) t += red[threadIdx.x][1 + i];
        red[threadIdx.x][0] = t;
    }
    __syncthreads();
}

// Per-block partials of K sums: block `blk` writes part[K * blk + k].
template <int K>
__device__ inline void write_partials(const double v[K], double* part, double (*red)[BLOCK]) {
    block_sum<K>(v, red);
    if ((int)threadIdx.x < K) part[K * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Totals of nparts partials written by an earlier kernel, in block order (every caller gets the
// same bits).
template <int K>
__device__ inline void total_partials(const double* part, int nparts, double out[K], double (*red)[BLOCK]) {
    double s[K];
    for (int k = 0; k < K; ++k) s[k] = 0.0;
    for (int i = threadIdx.x; i < nparts; i += BLOCK)
        for (int k = 0; k < K; ++k) s[k] += part[K * i + k];
    block_sum<K>(s, red);
    for (int k = 0; k < K; ++k) out[k] = red[k][0];
    __syncthreads();  // red is reused by the caller
}

// This thread's share of nparts partials (K sums each) written by an earlier kernel, added in block
// order; the loads are issued at once, so a kernel can start them before anything it waits for.
template <int K>
__device__ inline void load_partials(const double* part, int nparts, double s[K]) {
    for (int k = 0; k < K; ++k) s[k] = 0.0;
    for (int i = threadIdx.x; i < nparts; i += BLOCK)
        for (int k = 0; k < K; ++k) s[k] += part[K * i + k];
}

// The totals of the shares: every block gets the same bits.
template <int K>
__device__ inline void sum_shares(const double s[K], double out[K], double (*red)[BLOCK]) {
    block_sum<K>(s, red);
    for (int k = 0; k < K; ++k) out[k] = red[k][0];
    __syncthreads();  // red is reused by the caller
}

// (kernels that are not templates have internal linkage: every source including this has its own)
static __global__ void k_reset(Ctl* ctl, const int* gate) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->iterations = 0;
        ctl->done = (gate && *gate) ? 1 : 0;
    }
}

// R = B - A X, Z = D^-1 R, P_old = 0; partials of <R, Z> and <B, B>
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_setup(Op op, int64_t N, const int* __restrict__ n_dev,
                                                 const double* __restrict__ B,
                                                 const double* __restrict__ X, double* __restrict__ R,
                                                 double* __restrict__ Z, double* __restrict__ Pold,
                                                 double* __restrict__ inv_diag, double* __restrict__ part2,
                                                 const Ctl* ctl) {
    if (ctl->done) return;
    if (n_dev) N = *n_dev;
    __shared__ double red[6][BLOCK];
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    auto xv = [X](int64_t b, int c) { return X[3 * b + c]; };
    for (int64_t a = (int64_t)blockIdx.x * BLOCK + threadIdx.x; a < N; a += (int64_t)gridDim.x * BLOCK) {
        double ax[3], d[3];
        op.apply(a, xv, ax);
        op.inv_diag(a, d);
        for (int c = 0; c < 3; ++c) {
            const double b = B[3 * a + c];
            const double r = b - ax[c];
            const double z = d[c] * r;
            R[3 * a + c] = r;
            Z[3 * a + c] = z;
            Pold[3 * a + c] = 0.0;
            inv_diag[3 * a + c] = d[c];
            s[c] += r * z;
            s[3 + c] += b * b;
        }
    }
    write_partials<6>(s, part2, red);
}

// iteration slot k: the stopping test of iteration k, then P = Z + beta P_old, AP = A P
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_iter1(Op op, int64_t N, const int* __restrict__ n_dev, int k,
                                                 const double* __restrict__ Z,
                                                 const double* __restrict__ Pold, double* __restrict__ Pnew,
                                                 double* __restrict__ AP, const double* __restrict__ part2,
                                                 double* __restrict__ part1, Ctl* ctl, double tol) {
    // the reads that depend on nothing of this kernel first, so that their latencies overlap
    const int done = ctl->done;
    if (n_dev) N = *n_dev;
    double sh[6];
    load_partials<6>(part2, gridDim.x, sh);
    double rz_prev[3], bn[3];
    for (int c = 0; c < 3; ++c) {
        rz_prev[c] = ctl->rz[(k - 1) & 1][c];  // the slot of iteration k - 1 (unused when k = 0)
        bn[c] = ctl->bnorm[c];
    }
    if (done) return;
    __shared__ double red[6][BLOCK];
    double t[6];
    sum_shares<6>(sh, t, red);
    double be[3] = {0.0, 0.0, 0.0};
    if (k == 0) {  // t = <R, Z>, <B, B> of the setup
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (int c = 0; c < 3; ++c) {
                ctl->rz[0][c] = t[c];
                ctl->bnorm[c] = t[3 + c] > 0.0 ? sqrt(t[3 + c]) : 1.0;  // a zero column counts as 1
            }
    } else {  // t = <R, Z>, <R, R> after iteration k
        bool conv = true;
        for (int c = 0; c < 3; ++c) conv = conv && (sqrt(t[3 + c]) / bn[c] < tol);
        if (conv) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                ctl->iterations = k;
                ctl->done = 1;
            }
            return;
        }
        for (int c = 0; c < 3; ++c) be[c] = rz_prev[c] == 0.0 ? 0.0 : t[c] / rz_prev[c];
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (int c = 0; c < 3; ++c) ctl->rz[k & 1][c] = t[c];
    }
    auto pv = [Z, Pold, &be](int64_t b, int c) { return Z[3 * b + c] + Pold[3 * b + c] * be[c]; };
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t a = (int64_t)blockIdx.x * BLOCK + threadIdx.x; a < N; a += (int64_t)gridDim.x * BLOCK) {
        double ap[3];
        op.apply(a, pv, ap);
        for (int c = 0; c < 3; ++c) {
            const double p = pv(a, c);
            Pnew[3 * a + c] = p;
            AP[3 * a + c] = ap[c];
            s[c] += p * ap[c];
        }
    }
    write_partials<3>(s, part1, red);
}

// iteration slot k: alpha = <r, z> / <P, AP>; X += alpha P, R -= alpha AP, Z = D^-1 R
static __global__ __launch_bounds__(BLOCK) void k_iter2(int64_t N, const int* __restrict__ n_dev, int k,
                                                        const double* __restrict__ P,
                                                        const double* __restrict__ AP,
                                                        const double* __restrict__ inv_diag, double* __restrict__ X,
                                                        double* __restrict__ R, double* __restrict__ Z,
                                                        const double* __restrict__ part1, double* __restrict__ part2,
                                                        const Ctl* ctl) {
    // the reads that depend on nothing of this kernel first (the partials, the control block and
    // this thread's first node, in memory for every node below the host's N), so that their
    // latencies overlap; each later node is read while the one before is updated
    const int done = ctl->done;
    const int64_t n = n_dev ? *n_dev : N;
    double sh[3];
    load_partials<3>(part1, gridDim.x, sh);
    double rz[3];
    for (int c = 0; c < 3; ++c) rz[c] = ctl->rz[k & 1][c];
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    int64_t a = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double vp[3], vap[3], vd[3], vx[3], vr[3];
    auto fetch = [&](int64_t b) {
        for (int c = 0; c < 3; ++c) {
            const int64_t i = 3 * b + c;
            vp[c] = P[i];
            vap[c] = AP[i];
            vd[c] = inv_diag[i];
            vx[c] = X[i];
            vr[c] = R[i];
        }
    };
    if (a < N) fetch(a);
    if (done) return;
    __shared__ double red[6][BLOCK];
    double t[3];
    sum_shares<3>(sh, t, red);
    double al[3];
    for (int c = 0; c < 3; ++c) al[c] = t[c] == 0.0 ? 0.0 : rz[c] / t[c];
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (; a < n; a += stride) {
        double xo[3], ro[3], zo[3];
        for (int c = 0; c < 3; ++c) {
            xo[c] = vx[c] + vp[c] * al[c];
            ro[c] = vr[c] - vap[c] * al[c];
            zo[c] = vd[c] * ro[c];
            s[c] += ro[c] * zo[c];
            s[3 + c] += ro[c] * ro[c];
        }
        if (a + stride < n) fetch(a + stride);
        for (int c = 0; c < 3; ++c) {
            const int64_t i = 3 * a + c;
            X[i] = xo[c];
            R[i] = ro[c];
            Z[i] = zo[c];
        }
    }
    write_partials<6>(s, part2, red);
}

// after the last slot: the iteration limit was reached (converged or not)
// After the last slot: a solve still running took all max_iter iterations; `capped` (nullable)
// records that it did (whether or not the last update met the tolerance).
static __global__ void k_close(int max_iter, Ctl* ctl, int* capped) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !ctl->done) {
        ctl->iterations = max_iter;
        ctl->done = 1;
        if (capped) *capped = 1;
    }
}

static __global__ void k_report(const Ctl* ctl, int* iterations) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *iterations = ctl->iterations;
}

// Workspace of a solve over N nodes: R, Z, two P buffers, AP and D^-1 (3N doubles each), the
// per-block partials and the control block.
inline size_t workspace_bytes(int64_t N) {
    return sizeof(double) * (6 * 3 * (size_t)N + 9 * (size_t)MAX_BLOCKS) + sizeof(Ctl) + 256;
}

struct Workspace {
    double *R, *Z, *P[2], *AP, *inv_diag, *part1, *part2;
    Ctl* ctl;
};

inline Workspace carve(void* ws, int64_t N) {
    Workspace w;
    double* p = static_cast<double*>(ws);
    const size_t n3 = 3 * (size_t)N;
    w.R = p;
    w.Z = p + n3;
    w.P[0] = p + 2 * n3;
    w.P[1] = p + 3 * n3;
    w.AP = p + 4 * n3;
    w.inv_diag = p + 5 * n3;
    w.part1 = p + 6 * n3;
    w.part2 = w.part1 + 3 * MAX_BLOCKS;
    w.ctl = reinterpret_cast<Ctl*>(w.part2 + 6 * MAX_BLOCKS);
    return w;
}

// Enqueue a whole solve of A X = B from the X given. With `sync_every` > 0 the host reads the
// control block after `sync_every` iterations, then at intervals that double up to 4 `sync_every`,
// and stops enqueueing once it is done (the
// iteration count is returned in *iterations_host); with 0 all `max_iter` iteration slots are
// enqueued (those after convergence return at once) and nothing is read back, so the sequence can
// be captured into a graph; *iterations_dev, when given, receives the count on the device.
// `gate` (nullable, device) switches the whole solve off when nonzero. `n_dev` (nullable, device)
// holds the number of nodes when only the device knows it (at most N, which sizes the workspace);
// `blocks` (0: as many as N calls for) sets the grid of such a solve.
template <class Op>
inline hipError_t solve(const Op& op, int64_t N, const double* B, double* X, double tol, int max_iter, void* ws,
                        const int* gate, int sync_every, int* iterations_host, int* iterations_dev,
                        hipStream_t st, int* capped = nullptr, const int* n_dev = nullptr, unsigned blocks = 0) {
    Workspace w = carve(ws, N);
    const unsigned nb = blocks ? (blocks < (unsigned)MAX_BLOCKS ? blocks : (unsigned)MAX_BLOCKS) : blocks_for(N);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, w.ctl, gate);
    hipLaunchKernelGGL(k_setup<Op>, dim3(nb), dim3(BLOCK), 0, st, op, N, n_dev, B, (const double*)X, w.R, w.Z, w.P[1],
                       w.inv_diag, w.part2, (const Ctl*)w.ctl);
    int next_poll = sync_every, interval = sync_every;
    for (int k = 0; k < max_iter; ++k) {
        double* Pold = w.P[(k + 1) & 1];
        double* Pnew = w.P[k & 1];
        hipLaunchKernelGGL(k_iter1<Op>, dim3(nb), dim3(BLOCK), 0, st, op, N, n_dev, k, (const double*)w.Z,
                           (const double*)Pold, Pnew, w.AP, (const double*)w.part2, w.part1, w.ctl, tol);
        hipLaunchKernelGGL(k_iter2, dim3(nb), dim3(BLOCK), 0, st, N, n_dev, k, (const double*)Pnew, (const double*)w.AP,
                           (const double*)w.inv_diag, X, w.R, w.Z, (const double*)w.part1, w.part2,
                           (const Ctl*)w.ctl);
        if (sync_every > 0 && k + 1 == next_poll && k + 1 < max_iter) {
            interval = interval < 4 * sync_every ? 2 * interval : interval;
            next_poll += interval;
            int done = 0;
            hipError_t e = hipMemcpyAsync(&done, &w.ctl->done, sizeof(int), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
            if (done) break;
        }
    }
    // not stopped by then: max_iter iterations (whether or not the last one converged)
    hipLaunchKernelGGL(k_close, dim3(1), dim3(64), 0, st, max_iter, w.ctl, capped);
    if (iterations_dev) hipLaunchKernelGGL(k_report, dim3(1), dim3(64), 0, st, (const Ctl*)w.ctl, iterations_dev);
    if (sync_every > 0 && iterations_host) {
        hipError_t e = hipMemcpyAsync(iterations_host, &w.ctl->iterations, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace pcg
}  // namespace aldvc
