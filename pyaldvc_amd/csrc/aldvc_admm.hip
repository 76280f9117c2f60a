// The rest of an ADMM iteration on the GPU: the local step's outlier cleanup and the dual update,
// so that a whole iteration (3-DOF IC-GN, cleanup, global step, dual update) runs without host
// round trips and can be recorded into a HIP graph.
//
// Reference:
//   solver/subpb1_solver.py:135-164   bad = status != CONVERGED, plus the universal median test on
//                                     the good nodes when more than 27 are good; bad nodes keep U_hat
//   utils/outlier_detection.py:98-160 normalized fluctuation |u - med| / (med |r| + eps) over the
//                                     26 neighbours (median_filter, mode "nearest"); unknown nodes
//                                     spring-filled first; cut meshes: reachable neighbours only,
//                                     nanmedian, fewer than 6 neighbours not judged (:24-90)
//   utils/inpaint.py:35-121           spring fill: the discrete Laplace equation on the unknown
//                                     nodes, known neighbours as data, 1e-10 Tikhonov term
//   core/pipeline.py:286-366          F2 = nodal gradient where finite else F1; W += F2 - F1,
//                                     v += U2 - U1 (or reset); RMS updates; stop when the RMS
//                                     update of U_global or U_local drops below admm_tol
//
// Every launch takes a device flag `stop` (nullable): when set, the launch does nothing, which is
// how the iterations after the stopping rule fired are skipped inside a replayed graph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "aldvc_abi_util.h"
#include "aldvc_pcg.h"

using namespace aldvc;

namespace {

constexpr int ABLOCK = 256;
constexpr double SPRING_TIKHONOV = 1e-10;  // inpaint.py:103-104
constexpr int MIN_NEIGHBOURS = 6;          // outlier_detection.py:21
constexpr int ST_CONVERGED = 0;

struct NodeGrid {
    int nx, ny, nz;
    __device__ int64_t n() const { return (int64_t)nx * ny * nz; }
    __device__ int dim(int ax) const { return ax == 0 ? nx : (ax == 1 ? ny : nz); }
    __device__ int64_t lin(const int p[3]) const { return ((int64_t)p[2] * ny + p[1]) * nx + p[0]; }
    __device__ void pos(int64_t a, int p[3]) const {
        p[0] = (int)(a % nx);
        p[1] = (int)((a / nx) % ny);
        p[2] = (int)(a / ((int64_t)nx * ny));
    }
};

// Control of one cleanup: whether the median test runs, and per column whether any node is known.
struct MedCtl {
    int skip;       // 1: no median test (stopped, threshold <= 0, or at most 27 good nodes)
    int col_on[3];  // column c has at least one known node
};

__device__ inline bool gated(const int* stop) { return stop && *stop; }

// ---------------------------------------------------------------------------- spring fill
// Unknown nodes of column c: bit c of unknown[a]. Springs join 6-neighbours inside the grid, except
// across cut edges (edge_ok (N, 3): the +x, +y, +z edge of every node; null for an uncut mesh).
//
// The springs of node a, as 6 bits in the order -x, +x, -y, +y, -z, +z. Worked out once per solve,
// so the operator's stencil is a byte and three strides (no index decode on the solve's path).
__device__ inline unsigned spring_code(const NodeGrid& g, const uint8_t* edge_ok, int64_t a) {
    int p[3];
    g.pos(a, p);
    const int64_t stride[3] = {1, (int64_t)g.nx, (int64_t)g.nx * g.ny};
    unsigned m = 0;
    for (int ax = 0; ax < 3; ++ax) {
        if (p[ax] > 0 && (!edge_ok || edge_ok[3 * (a - stride[ax]) + ax])) m |= 1u << (2 * ax);
        if (p[ax] < g.dim(ax) - 1 && (!edge_ok || edge_ok[3 * a + ax])) m |= 2u << (2 * ax);
    }
    return m;
}

struct SpringOp {
    NodeGrid g;
    const uint8_t* unknown;
    const uint8_t* code;  // spring_code of every node

    template <class F>
    __device__ void for_springs(int64_t a, F f) const {
        const unsigned m = code[a];
        const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
        if (m & 1u) f(a - 1);
        if (m & 2u) f(a + 1);
        if (m & 4u) f(a - sy);
        if (m & 8u) f(a + sy);
        if (m & 16u) f(a - sz);
        if (m & 32u) f(a + sz);
    }
    __device__ double diag(int64_t a) const { return (double)__popc(code[a]) + SPRING_TIKHONOV; }
    __device__ void inv_diag(int64_t a, double d[3]) const {
        const double dg = diag(a);
        for (int c = 0; c < 3; ++c) d[c] = (unknown[a] >> c) & 1 ? 1.0 / dg : 0.0;
    }
    // A node known in column c holds zero in that column of every vector the solve applies this to:
    // the right-hand side and the start are zero there, so the residual, the search direction and
    // the iterate stay zero. The coupling to a neighbour therefore needs no test of its own bits.
    template <class X>
    __device__ void apply(int64_t a, const X& x, double y[3]) const {
        const unsigned u = unknown[a];
        if (!u) {
            y[0] = y[1] = y[2] = 0.0;
            return;
        }
        const double dg = diag(a);
        for (int c = 0; c < 3; ++c) y[c] = (u >> c) & 1 ? dg * x(a, c) : 0.0;
        for_springs(a, [&](int64_t q) {
            for (int c = 0; c < 3; ++c)
                if ((u >> c) & 1) y[c] -= x(q, c);
        });
    }
};

__global__ void k_spring_codes(NodeGrid g, const uint8_t* __restrict__ edge_ok, uint8_t* __restrict__ code,
                               const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t N = g.n();
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x)
        code[a] = (uint8_t)spring_code(g, edge_ok, a);
}

// ---------------------------------------------------------------------------- compact fill
// The fill solved over its unknown nodes only: entry i of the compact vectors is node list[i] (the
// nodes with any unknown bit, ascending), nbr[6 i + k] the entry of its k-th spring neighbour in the
// order of spring_code (-1: no spring, or a neighbour known in every column) and meta[i] its
// unknown bits | spring count << 3. The iteration is the one over all nodes: a neighbour known in
// a column holds zero there, so leaving it out changes no sum.
struct CompactSpringOp {
    const uint8_t* meta;
    const int* nbr;

    __device__ void inv_diag(int64_t i, double d[3]) const {
        const unsigned m = meta[i];
        const double dg = (double)(m >> 3) + SPRING_TIKHONOV;
        for (int c = 0; c < 3; ++c) d[c] = (m >> c) & 1 ? 1.0 / dg : 0.0;
    }
    template <class X>
    __device__ void apply(int64_t i, const X& x, double y[3]) const {
        const unsigned m = meta[i];
        const double dg = (double)(m >> 3) + SPRING_TIKHONOV;
        for (int c = 0; c < 3; ++c) y[c] = (m >> c) & 1 ? dg * x(i, c) : 0.0;
        for (int k = 0; k < 6; ++k) {
            const int j = nbr[6 * i + k];
            if (j >= 0)
                for (int c = 0; c < 3; ++c)
                    if ((m >> c) & 1) y[c] -= x(j, c);
        }
    }
};

constexpr int LIST_THREADS = 1024;

// list, pos_of (entry of each listed node) and *count in one workgroup: each thread scans a
// contiguous run of nodes (16-byte loads of the unknown bytes: the run starts on a multiple of 16
// and the unknown array is padded to 256 bytes), the runs' counts are scanned in the block, then
// each thread writes its run's entries in order.
__global__ __launch_bounds__(LIST_THREADS) void k_fill_list(int64_t N, const uint8_t* __restrict__ unknown,
                                                            int* __restrict__ list, int* __restrict__ pos_of,
                                                            int* __restrict__ count, const int* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ int scan[LIST_THREADS];
    const int t = threadIdx.x;
    const int64_t run = ((N + LIST_THREADS - 1) / LIST_THREADS + 15) / 16 * 16;
    const int64_t lo = (int64_t)t * run, hi = lo + run < N ? lo + run : N;
    int n = 0;
    for (int64_t a = lo; a < hi; a += 16) {
        const uint4 w = *reinterpret_cast<const uint4*>(unknown + a);
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) n += (a + k < hi && ((ws[k >> 2] >> (8 * (k & 3))) & 0xffu)) ? 1 : 0;
    }
    scan[t] = n;
    __syncthreads();
    for (int off = 1; off < LIST_THREADS; off <<= 1) {  // inclusive scan (Hillis-Steele)
        const int v = t >= off ? scan[t - off] : 0;
        __syncthreads();
        scan[t] += v;
        __syncthreads();
    }
    int pos = scan[t] - n;
    for (int64_t a = lo; a < hi; a += 16) {
        const uint4 w = *reinterpret_cast<const uint4*>(unknown + a);
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        if (!(ws[0] | ws[1] | ws[2] | ws[3])) continue;
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (a + k < hi && ((ws[k >> 2] >> (8 * (k & 3))) & 0xffu)) {
                list[pos] = (int)(a + k);
                pos_of[a + k] = pos;
                ++pos;
            }
    }
    if (t == LIST_THREADS - 1) *count = scan[t];
}

// the compact operator's tables and right-hand side (b of the listed nodes)
__global__ void k_fill_tables(SpringOp op, const int* __restrict__ list, const int* __restrict__ pos_of,
                              const int* __restrict__ count, const double* __restrict__ b_full,
                              uint8_t* __restrict__ meta, int* __restrict__ nbr, double* __restrict__ bc,
                              double* __restrict__ Xc, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int M = *count;
    const int64_t sy = op.g.nx, sz = (int64_t)op.g.nx * op.g.ny;
    const int64_t step[6] = {-1, 1, -sy, sy, -sz, sz};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        const int64_t a = list[i];
        const unsigned code = op.code[a];
        meta[i] = (uint8_t)(op.unknown[a] | (__popc(code) << 3));
        for (int k = 0; k < 6; ++k) {
            int j = -1;
            if ((code >> k) & 1) {
                const int64_t q = a + step[k];
                if (op.unknown[q]) j = pos_of[q];
            }
            nbr[6 * i + k] = j;
        }
        for (int c = 0; c < 3; ++c) {
            bc[3 * i + c] = b_full[3 * a + c];
            Xc[3 * i + c] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------- one-workgroup fill
// A fill of at most SMALL_MAX unknown nodes is solved in one launch of three workgroups, one per
// column, each an independent Jacobi-preconditioned CG over the compact operator's column (the
// reference fills every column by a solve of its own: utils/inpaint.py:35-121). The column's
// search direction lives in LDS; each thread keeps the iterate, residual, A P and the tables of its
// SMALL_EPT entries (entry i = thread + SMALL_THREADS k) in registers. Scalars and the stopping
// rule (relative residual below tol, tested after every update but the last) are pcg::solve's,
// per column; the block sums run in a fixed order, so every thread derives the same scalars.
constexpr int SMALL_THREADS = 1024;
constexpr int SMALL_EPT = 7;
constexpr int SMALL_MAX = SMALL_THREADS * SMALL_EPT;  // 7168 entries: 56 KiB of LDS per vector
constexpr int SMALL_WAVES = SMALL_THREADS / 64;
constexpr unsigned SMALL_NONE = 0xffffu;  // no neighbour entry

struct SmallCtl {
    int gate;           // 1: the all-node solve is not needed (fill skipped, or solved here)
    int solved;         // 1: solved here
    int iterations[3];  // per column, of the solve made here
    int capped[3];      // per column: it took max_iter iterations
};

// totals of K sums over the workgroup: each wave by the butterfly, then the waves in order
template <int K>
__device__ inline void small_sum(const double v[K], double (*red)[SMALL_WAVES], double out[K]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double w[K];
    for (int k = 0; k < K; ++k) w[k] = pcg::wave_sum(v[k]);
    if (lane == 0)
        for (int k = 0; k < K; ++k) red[k][wave] = w[k];
    __syncthreads();
    for (int k = 0; k < K; ++k) {
        double t = red[k][0];
        for (int i = 1; i < SMALL_WAVES; ++i) t += red[k][i];
        out[k] = t;
    }
}

__global__ __launch_bounds__(SMALL_THREADS) void k_fill_small(const int* __restrict__ count,
                                                              const uint8_t* __restrict__ meta,
                                                              const int* __restrict__ nbr,
                                                              const double* __restrict__ bc, double* __restrict__ Xc,
                                                              double tol, int max_iter, const int* __restrict__ skip,
                                                              SmallCtl* sc) {
    const int t = threadIdx.x, col = blockIdx.x;
    const bool lead = t == 0 && col == 0;
    if (skip && *skip) {
        if (lead) {
            sc->gate = 1;
            sc->solved = 0;
        }
        return;
    }
    const int M = *count;
    if (M > SMALL_MAX) {
        if (lead) {
            sc->gate = 0;
            sc->solved = 0;
        }
        return;
    }
    __shared__ double Ps[SMALL_MAX], Xs[SMALL_MAX];  // search direction and iterate of the column
    __shared__ double redA[1][SMALL_WAVES], redB[2][SMALL_WAVES];
    __shared__ double dg_of[8], dinv_of[8];  // diagonal and its inverse by spring count
    unsigned nb[SMALL_EPT][3], mt[SMALL_EPT];  // neighbour entries (16 bits each), meta
    double R[SMALL_EPT], AP[SMALL_EPT];
    if (t < 8) {
        dg_of[t] = (double)t + SPRING_TIKHONOV;
        dinv_of[t] = 1.0 / dg_of[t];
    }
    __syncthreads();
    // the Jacobi inverse of entry e: zero where the column is known
    auto dinv = [&](int e) { return (mt[e] >> col) & 1 ? dinv_of[mt[e] >> 3] : 0.0; };
    double s2[2] = {0.0, 0.0};
#pragma unroll
    for (int e = 0; e < SMALL_EPT; ++e) {
        const int i = t + SMALL_THREADS * e;
        nb[e][0] = nb[e][1] = nb[e][2] = 0xffffffffu;
        mt[e] = 0;
        R[e] = AP[e] = 0.0;
        if (i < M) {
            mt[e] = meta[i];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int j = nbr[6 * i + k];
                const unsigned v = j >= 0 ? (unsigned)j : SMALL_NONE;
                nb[e][k >> 1] = (nb[e][k >> 1] & ~(0xffffu << (16 * (k & 1)))) | (v << (16 * (k & 1)));
            }
            const double b = bc[3 * i + col];
            const double z = dinv(e) * b;
            R[e] = b;
            Ps[i] = z;
            Xs[i] = 0.0;
            s2[0] += b * z;
            s2[1] += b * b;
        }
    }
    double tot[2];
    small_sum<2>(s2, redB, tot);  // its barrier also publishes P
    double rz = tot[0];
    const double bnorm = tot[1] > 0.0 ? sqrt(tot[1]) : 1.0;  // a zero column counts as 1
    int iterations = max_iter, capped = 1;
    for (int k = 0; k < max_iter; ++k) {
        // the tables count as rewritten here, so the compiler derives the LDS addresses inside the
        // loop instead of holding them in registers across it
#pragma unroll
        for (int e = 0; e < SMALL_EPT; ++e)
            asm volatile("" : "+v"(nb[e][0]), "+v"(nb[e][1]), "+v"(nb[e][2]), "+v"(mt[e]));
        // A P (the compact operator's terms in its order) and <P, A P>
        double s1[1] = {0.0};
#pragma unroll
        for (int e = 0; e < SMALL_EPT; ++e) {
            const int i = t + SMALL_THREADS * e;
            if (i < M) {
                const bool on = (mt[e] >> col) & 1;
                const double p = Ps[i];
                double y = on ? dg_of[mt[e] >> 3] * p : 0.0;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const unsigned j = (nb[e][q >> 1] >> (16 * (q & 1))) & 0xffffu;
                    if (j != SMALL_NONE && on) y -= Ps[j];
                }
                AP[e] = y;
                s1[0] += p * y;
            }
        }
        double pap[1];
        small_sum<1>(s1, redA, pap);
        const double al = pap[0] == 0.0 ? 0.0 : rz / pap[0];
        s2[0] = s2[1] = 0.0;
#pragma unroll
        for (int e = 0; e < SMALL_EPT; ++e) {
            const int i = t + SMALL_THREADS * e;
            if (i < M) {
                Xs[i] += Ps[i] * al;
                const double r = R[e] - AP[e] * al;
                R[e] = r;
                const double z = dinv(e) * r;
                s2[0] += r * z;
                s2[1] += r * r;
            }
        }
        small_sum<2>(s2, redB, tot);
        if (k + 1 == max_iter) break;  // the last slot: max_iter iterations, whatever the residual
        if (sqrt(tot[1]) / bnorm < tol) {
            iterations = k + 1;
            capped = 0;
            break;
        }
        const double be = rz == 0.0 ? 0.0 : tot[0] / rz;
        rz = tot[0];
#pragma unroll
        for (int e = 0; e < SMALL_EPT; ++e) {
            const int i = t + SMALL_THREADS * e;
            if (i < M) Ps[i] = dinv(e) * R[e] + Ps[i] * be;
        }
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < SMALL_EPT; ++e) {
        const int i = t + SMALL_THREADS * e;
        if (i < M) Xc[3 * i + col] = Xs[i];
    }
    if (t == 0) {
        sc->iterations[col] = iterations;
        sc->capped[col] = capped;
    }
    if (lead) {
        sc->gate = 1;
        sc->solved = 1;
    }
}

// the fill's iteration count (the largest over the columns) and capped flag from whichever solve ran
__global__ void k_fill_result(const SmallCtl* sc, const int* __restrict__ all_node_iterations, int* __restrict__ out,
                              int* __restrict__ capped) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (sc->solved) {
        int it = 0, cap = 0;
        for (int c = 0; c < 3; ++c) {
            it = sc->iterations[c] > it ? sc->iterations[c] : it;
            cap |= sc->capped[c];
        }
        *out = it;
        if (capped && cap) *capped = 1;
    } else {
        *out = *all_node_iterations;
    }
}

// the compact solution back into the all-node iterate (zero at the nodes not listed)
__global__ void k_fill_scatter(const int* __restrict__ list, const int* __restrict__ count,
                               const double* __restrict__ Xc, double* __restrict__ X, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int M = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) X[3 * (int64_t)list[i] + c] = Xc[3 * i + c];
}

// bad = status != CONVERGED; unknown bits (not good, or NaN); per-block counts of good nodes and
// of known nodes per column
__global__ __launch_bounds__(ABLOCK) void k_cleanup_setup(NodeGrid g, const int8_t* __restrict__ status,
                                                          const double* __restrict__ U, uint8_t* __restrict__ unknown,
                                                          double* __restrict__ part, const int* __restrict__ stop) {
    if (gated(stop)) return;
    __shared__ double red[4][pcg::BLOCK];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t N = g.n();
    for (int64_t a = (int64_t)blockIdx.x * ABLOCK + threadIdx.x; a < N; a += (int64_t)gridDim.x * ABLOCK) {
        const bool good = status[a] == ST_CONVERGED;
        unsigned u = 0;
        for (int c = 0; c < 3; ++c) {
            const bool unk = !good || isnan(U[3 * a + c]);
            u |= (unk ? 1u : 0u) << c;
            s[1 + c] += unk ? 0.0 : 1.0;
        }
        unknown[a] = (uint8_t)u;
        s[0] += good ? 1.0 : 0.0;
    }
    pcg::write_partials<4>(s, part, red);
}

// the median test's switches from the counts
__global__ __launch_bounds__(ABLOCK) void k_cleanup_switches(const double* __restrict__ part, int nparts,
                                                             double threshold, MedCtl* ctl,
                                                             const int* __restrict__ stop) {
    __shared__ double red[4][pcg::BLOCK];
    if (gated(stop)) {
        if (threadIdx.x == 0) ctl->skip = 1;
        return;
    }
    double t[4];
    pcg::total_partials<4>(part, nparts, t, red);
    if (threadIdx.x == 0) {
        ctl->skip = (threshold > 0.0 && t[0] > 27.0) ? 0 : 1;
        for (int c = 0; c < 3; ++c) ctl->col_on[c] = t[1 + c] > 0.0 ? 1 : 0;
    }
}

// right-hand side of the fill: the sum of the known spring neighbours; X = 0
__global__ void k_spring_rhs(SpringOp op, const double* __restrict__ vals, double* __restrict__ b,
                             double* __restrict__ X, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t N = op.g.n();
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        const unsigned u = op.unknown[a];
        double r[3] = {0.0, 0.0, 0.0};
        if (u) {
            op.for_springs(a, [&](int64_t q) {
                const unsigned kq = ~op.unknown[q] & u;  // known at q, unknown here
                for (int c = 0; c < 3; ++c)
                    if ((kq >> c) & 1) r[c] += vals[3 * q + c];
            });
        }
        for (int c = 0; c < 3; ++c) {
            b[3 * a + c] = r[c];
            X[3 * a + c] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------- median test
// The entries of v with ok set in ascending order, the others (+inf) after them: a bitonic network
// over 32 slots with every index fixed at compile time, so the values stay in registers.
template <int M>
__device__ inline void sort32(const double (&v)[M], const bool (&ok)[M], double (&s)[32]) {
    static_assert(M <= 32, "at most 32 values");
#pragma unroll
    for (int k = 0; k < 32; ++k) s[k] = (k < M && ok[k < M ? k : 0]) ? v[k < M ? k : 0] : __builtin_inf();
#pragma unroll
    for (int size = 2; size <= 32; size <<= 1)
#pragma unroll
        for (int stride = size / 2; stride > 0; stride >>= 1)
#pragma unroll
            for (int i = 0; i < 32; ++i) {
                const int j = i ^ stride;
                if (j > i) {
                    const double lo = fmin(s[i], s[j]), hi = fmax(s[i], s[j]);
                    const bool up = (i & size) == 0;
                    s[i] = up ? lo : hi;
                    s[j] = up ? hi : lo;
                }
            }
}

// s[r] (NaN when r is not below n)
__device__ inline double pick32(const double (&s)[32], int r, int n) {
    double o = __builtin_nan("");
#pragma unroll
    for (int k = 0; k < 32; ++k)
        if (k == r && r < n) o = s[k];
    return o;
}

// Move by d (+1 or -1) along axis ax from p through an ok edge, inside the grid.
__device__ inline bool step_ok(const NodeGrid& g, const uint8_t* edge_ok, const int p[3], int ax, int d) {
    if (d > 0) return p[ax] < g.dim(ax) - 1 && edge_ok[3 * g.lin(p) + ax];
    if (p[ax] <= 0) return false;
    int q[3] = {p[0], p[1], p[2]};
    q[ax] -= 1;
    return edge_ok[3 * g.lin(q) + ax] != 0;
}

// Neighbour at offset (dx, dy, dz) reached through ok edges, the axis steps taken in some order
// (outlier_detection.py:41-70).
__device__ inline bool reachable(const NodeGrid& g, const uint8_t* edge_ok, const int p[3], const int off[3]) {
    constexpr int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    bool any = false;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int q[3] = {p[0], p[1], p[2]};
        bool ok = true;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int ax = perms[k][t];
            if (off[ax] != 0) {
                ok = ok && step_ok(g, edge_ok, q, ax, off[ax]);
                q[ax] += off[ax];
            }
        }
        any = any || ok;
    }
    return any;
}

// The 26 neighbours of node p: uncut, indices clamped to the grid (mode "nearest"); cut, the
// reachable ones (ok false for the others).
template <bool CUT, class V>
__device__ inline void gather26(const NodeGrid& g, const uint8_t* edge_ok, const int p[3], int c, const V& val,
                                double (&v)[26], bool (&ok)[26]) {
    int k = 0;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                const int off[3] = {dx, dy, dz};
                int q[3];
                bool in = true;
                for (int ax = 0; ax < 3; ++ax) {
                    q[ax] = p[ax] + off[ax];
                    if (!CUT) q[ax] = q[ax] < 0 ? 0 : (q[ax] >= g.dim(ax) ? g.dim(ax) - 1 : q[ax]);
                }
                if (CUT) in = reachable(g, edge_ok, p, off);
                double x = in ? val(g.lin(q), c) : __builtin_nan("");
                v[k] = x;
                ok[k] = in && !isnan(x);
                ++k;
            }
}

// median of the ok entries: uncut, rank 13 of all 26 (median_filter's rank for 26 values); cut,
// numpy's nanmedian (mean of the two middle values for an even count; NaN when there is none)
template <bool CUT>
__device__ inline double median26(const double (&v)[26], const bool (&ok)[26], int& n) {
    n = 0;
#pragma unroll
    for (int k = 0; k < 26; ++k) n += ok[k] ? 1 : 0;
    double s[32];
    sort32<26>(v, ok, s);
    if (!CUT) return pick32(s, 13, n);
    if (n == 0) return __builtin_nan("");
    const double lo = pick32(s, (n - 1) / 2, n), hi = pick32(s, n / 2, n);
    return (n & 1) ? lo : (lo + hi) / 2.0;
}

// pass 1: absf = |x - med(neighbours)| per column (cut: also the neighbour count)
template <bool CUT>
__global__ __launch_bounds__(ABLOCK) void k_median_pass1(NodeGrid g, const uint8_t* __restrict__ edge_ok,
                                                         const uint8_t* __restrict__ unknown,
                                                         const double* __restrict__ U, const double* __restrict__ X,
                                                         double* __restrict__ absf, uint8_t* __restrict__ cnt,
                                                         const MedCtl* ctl) {
    if (ctl->skip) return;
    // uncut: unknown nodes carry their spring-filled value X; cut: they are NaN
    auto val = [&](int64_t n, int c) {
        const bool unk = (unknown[n] >> c) & 1;
        return unk ? (CUT ? __builtin_nan("") : X[3 * n + c]) : U[3 * n + c];
    };
    const int64_t N = g.n();
    for (int64_t a = (int64_t)blockIdx.x * ABLOCK + threadIdx.x; a < N; a += (int64_t)gridDim.x * ABLOCK) {
        int p[3];
        g.pos(a, p);
        for (int c = 0; c < 3; ++c) {
            double v[26];
            bool ok[26];
            gather26<CUT>(g, edge_ok, p, c, val, v, ok);
            int n;
            const double med = median26<CUT>(v, ok, n);
            absf[3 * a + c] = fabs(val(a, c) - med);
            cnt[3 * a + c] = (uint8_t)n;
        }
    }
}

// pass 2: nf = absf / (med(neighbours' absf) + eps), the flag, and the cleaned field:
// bad = !good | flag; U_out = bad ? U_hat : U
template <bool CUT>
__global__ __launch_bounds__(ABLOCK) void k_median_pass2(NodeGrid g, const uint8_t* __restrict__ edge_ok,
                                                         const int8_t* __restrict__ status,
                                                         const uint8_t* __restrict__ unknown,
                                                         const double* __restrict__ U,
                                                         const double* __restrict__ U_hat,
                                                         const double* __restrict__ absf,
                                                         const uint8_t* __restrict__ cnt, double threshold, double eps,
                                                         double* __restrict__ U_out, uint8_t* __restrict__ bad_out,
                                                         const MedCtl* ctl, const int* __restrict__ stop) {
    if (gated(stop)) return;
    const bool test = ctl->skip == 0;
    auto af = [absf](int64_t n, int c) { return absf[3 * n + c]; };
    const int64_t N = g.n();
    for (int64_t a = (int64_t)blockIdx.x * ABLOCK + threadIdx.x; a < N; a += (int64_t)gridDim.x * ABLOCK) {
        const bool good = status[a] == ST_CONVERGED;
        bool flag = false;
        if (test && good) {
            int p[3];
            g.pos(a, p);
            double mag2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                double nf = 0.0;
                if (ctl->col_on[c] && !((unknown[a] >> c) & 1)) {
                    double v[26];
                    bool ok[26];
                    gather26<CUT>(g, edge_ok, p, c, af, v, ok);
                    int n;
                    const double med_res = median26<CUT>(v, ok, n);
                    nf = absf[3 * a + c] / (med_res + eps);
                    if (CUT && (!isfinite(nf) || cnt[3 * a + c] < MIN_NEIGHBOURS)) nf = 0.0;
                }
                mag2 += nf * nf;
            }
            flag = sqrt(mag2) > threshold;
        }
        const bool bad = !good || flag;
        for (int c = 0; c < 3; ++c) U_out[3 * a + c] = bad ? U_hat[3 * a + c] : U[3 * a + c];
        bad_out[a] = bad ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------- dual update
// hist[4 * step + k]: RMS of U2n - U2 (global update), U1n - U1 (local update), U2n - U1n and
// F2n - F2 (primal residuals), then the state moves on: F2 = finite(F2n) ? F2n : F2, W, v, U2 = U2n,
// U1 = U1n. With check_stop the stopping rule sets *stop and records the step in *steps.
__global__ __launch_bounds__(ABLOCK) void k_dual_update(int64_t N, const double* U1n,
                                                        const double* __restrict__ U2n,
                                                        const double* __restrict__ F2n, double* U1,
                                                        double* __restrict__ U2, double* __restrict__ F2,
                                                        double* __restrict__ W, double* __restrict__ v,
                                                        int accumulate, double* __restrict__ part,
                                                        const int* __restrict__ stop) {
    if (gated(stop)) return;
    __shared__ double red[4][pcg::BLOCK];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t a = (int64_t)blockIdx.x * ABLOCK + threadIdx.x; a < N; a += (int64_t)gridDim.x * ABLOCK) {
        for (int c = 0; c < 3; ++c) {
            const int64_t i = 3 * a + c;
            const double u1n = U1n[i], u2n = U2n[i];
            const double dg = u2n - U2[i], dl = u1n - U1[i], ru = u2n - u1n;
            s[0] += dg * dg;
            s[1] += dl * dl;
            s[2] += ru * ru;
            v[i] = accumulate ? v[i] + ru : ru;
            U1[i] = u1n;
            U2[i] = u2n;
        }
        for (int k = 0; k < 9; ++k) {
            const int64_t i = 9 * a + k;
            const double f1 = F2[i], fr = F2n[i];
            const double f2 = isfinite(fr) ? fr : f1;
            const double df = f2 - f1;
            s[3] += df * df;
            W[i] = accumulate ? W[i] + df : df;
            F2[i] = f2;
        }
    }
    pcg::write_partials<4>(s, part, red);
}

// the RMS values of the step from the partials, and the stopping rule
__global__ __launch_bounds__(ABLOCK) void k_dual_final(const double* __restrict__ part, int nparts, int64_t N,
                                                       int check_stop, double admm_tol, int step,
                                                       double* __restrict__ hist, int* __restrict__ stop,
                                                       int* __restrict__ steps) {
    if (gated(stop)) return;
    __shared__ double red[4][pcg::BLOCK];
    double t[4];
    pcg::total_partials<4>(part, nparts, t, red);
    if (threadIdx.x == 0) {
        const double n3 = 3.0 * (double)N, n9 = 9.0 * (double)N;
        const double r[4] = {sqrt(t[0] / n3), sqrt(t[1] / n3), sqrt(t[2] / n3), sqrt(t[3] / n9)};
        for (int k = 0; k < 4; ++k) hist[4 * step + k] = r[k];
        *steps = step;
        if (check_stop && (r[0] < admm_tol || r[1] < admm_tol)) *stop = 1;
    }
}

__global__ void k_unknown_from_nan(int64_t N, const double* __restrict__ vals, uint8_t* __restrict__ unknown) {
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        unsigned u = 0;
        for (int c = 0; c < 3; ++c) u |= (isnan(vals[3 * a + c]) ? 1u : 0u) << c;
        unknown[a] = (uint8_t)u;
    }
}

__global__ void k_fill_merge(int64_t N, const uint8_t* __restrict__ unknown, const double* __restrict__ X,
                             double* __restrict__ vals) {
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c)
            if ((unknown[a] >> c) & 1) vals[3 * a + c] = X[3 * a + c];
}

// Carves consecutive 256-byte aligned regions out of a workspace; with a null base it only
// measures, so the size query and the carving follow the same sequence.
//
// The first region starts at the first 256-byte boundary of the workspace (the measured size
// counts the 255 bytes that may take), so every region is aligned for 16-byte loads.
struct Carver {
    char* base;
    size_t used;
    explicit Carver(char* b)
        : base(b ? reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(b) + 255) & ~uintptr_t(255)) : nullptr),
          used(b ? 0 : 255) {}
    template <class T>
    T* take(size_t count) {
        T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
        used += (sizeof(T) * count + 255) / 256 * 256;
        return p;
    }
};

// The compact fill's tables and vectors, sized for every node unknown.
struct CompactFill {
    int *list, *pos_of, *nbr, *count, *iterations;
    uint8_t* meta;
    double *bc, *Xc;
    SmallCtl* small;
    void take(Carver& c, int64_t N) {
        list = c.take<int>(N);
        pos_of = c.take<int>(N);
        nbr = c.take<int>(6 * N);
        meta = c.take<uint8_t>(N);
        bc = c.take<double>(3 * N);
        Xc = c.take<double>(3 * N);
        count = c.take<int>(1);
        iterations = c.take<int>(2);
        small = c.take<SmallCtl>(1);
    }
};

struct CleanupWs {
    uint8_t *unknown, *code, *cnt;
    double *absf, *b, *X, *part;
    MedCtl* ctl;
    CompactFill cf;
    void* pcg_ws;
    size_t bytes;
    CleanupWs(void* ws, int64_t N) {
        Carver c{static_cast<char*>(ws)};
        unknown = c.take<uint8_t>(N);
        code = c.take<uint8_t>(N);
        cnt = c.take<uint8_t>(3 * N);
        absf = c.take<double>(3 * N);
        b = c.take<double>(3 * N);
        X = c.take<double>(3 * N);
        part = c.take<double>(4 * pcg::MAX_BLOCKS);
        ctl = c.take<MedCtl>(1);
        cf.take(c, N);
        pcg_ws = c.take<char>(pcg::workspace_bytes(N));
        bytes = c.used;
    }
};

struct FillWs {
    uint8_t *unknown, *code;
    double *b, *X;
    CompactFill cf;
    void* pcg_ws;
    size_t bytes;
    FillWs(void* ws, int64_t N) {
        Carver c{static_cast<char*>(ws)};
        unknown = c.take<uint8_t>(N);
        code = c.take<uint8_t>(N);
        b = c.take<double>(3 * N);
        X = c.take<double>(3 * N);
        cf.take(c, N);
        pcg_ws = c.take<char>(pcg::workspace_bytes(N));
        bytes = c.used;
    }
};

int check_node_grid(int nx, int ny, int nz) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(-2, "empty node grid");
    if ((int64_t)nx * ny * nz > INT32_MAX) return fail(-2, "more than 2^31 - 1 nodes");
    return 0;
}

// Spring fill of b (all nodes) into X (all nodes, zero on entry) over the unknown nodes only:
// list them, build the compact tables, solve (in one workgroup when they are few, else by
// pcg::solve, which the one-workgroup kernel gates off when it has solved), scatter back. `skip`
// (nullable, device) switches the whole fill off; the other arguments are those of pcg::solve.
hipError_t fill_solve(const SpringOp& op, int64_t N, const double* b, double* X, const CompactFill& cf, double tol,
                      int max_iter, void* pcg_ws, const int* skip, int sync_every, int* iterations_host,
                      int* iterations_dev, int* capped, hipStream_t st) {
    constexpr unsigned COMPACT_BLOCKS = 128;
    const unsigned nbc = pcg::blocks_for(N) < COMPACT_BLOCKS ? pcg::blocks_for(N) : COMPACT_BLOCKS;
    hipLaunchKernelGGL(k_fill_list, dim3(1), dim3(LIST_THREADS), 0, st, N, op.unknown, cf.list, cf.pos_of, cf.count,
                       skip);
    hipLaunchKernelGGL(k_fill_tables, dim3(nbc), dim3(ABLOCK), 0, st, op, (const int*)cf.list,
                       (const int*)cf.pos_of, (const int*)cf.count, b, cf.meta, cf.nbr, cf.bc, cf.Xc, skip);
    hipLaunchKernelGGL(k_fill_small, dim3(3), dim3(SMALL_THREADS), 0, st, (const int*)cf.count,
                       (const uint8_t*)cf.meta, (const int*)cf.nbr, (const double*)cf.bc, cf.Xc, tol, max_iter, skip,
                       cf.small);
    const CompactSpringOp cop{cf.meta, cf.nbr};
    hipError_t e = pcg::solve(cop, N, (const double*)cf.bc, cf.Xc, tol, max_iter, pcg_ws, &cf.small->gate,
                              sync_every, nullptr, cf.iterations, st, capped, (const int*)cf.count, nbc);
    if (e != hipSuccess) return e;
    int* out = iterations_dev ? iterations_dev : cf.iterations + 1;
    hipLaunchKernelGGL(k_fill_result, dim3(1), dim3(64), 0, st, (const SmallCtl*)cf.small, (const int*)cf.iterations,
                       out, capped);
    hipLaunchKernelGGL(k_fill_scatter, dim3(nbc), dim3(ABLOCK), 0, st, (const int*)cf.list, (const int*)cf.count,
                       (const double*)cf.Xc, X, skip);
    if (iterations_host) {
        e = hipMemcpyAsync(iterations_host, out, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace

extern "C" {

size_t aldvc_subpb1_cleanup_workspace_bytes(int nx, int ny, int nz) {
    return CleanupWs(nullptr, (int64_t)nx * ny * nz).bytes;
}

int aldvc_subpb1_cleanup(int nx, int ny, int nz, const uint8_t* edge_ok, const int8_t* status, const double* U,
                         const double* U_hat, double threshold, double eps, double fill_tol, int fill_max_iter,
                         double* U_out, uint8_t* bad, const int* stop, int sync_every, int* fill_iterations,
                         int* capped, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    if (int rc = check_node_grid(nx, ny, nz)) return rc;
    if (!status || !U || !U_hat || !U_out || !bad || !workspace) return fail(-2, "null pointer");
    if (fill_max_iter < 1) return fail(-2, "fill_max_iter must be at least 1");
    if (sync_every < 0) return fail(-2, "sync_every must be 0 (all iterations enqueued) or positive");
    if (workspace_bytes < aldvc_subpb1_cleanup_workspace_bytes(nx, ny, nz)) return fail(-3, "workspace too small");
    const NodeGrid g{nx, ny, nz};
    const int64_t N = (int64_t)nx * ny * nz;
    const CleanupWs w(workspace, N);
    uint8_t* unknown = w.unknown;
    uint8_t* cnt = w.cnt;
    double *absf = w.absf, *b = w.b, *X = w.X, *part = w.part;
    MedCtl* ctl = w.ctl;
    void* pcg_ws = w.pcg_ws;
    const unsigned nb = pcg::blocks_for(N);
    const SpringOp op{g, unknown, w.code};  // the median test fills without cut edges (outlier_detection.py:127)

    hipLaunchKernelGGL(k_cleanup_setup, dim3(nb), dim3(ABLOCK), 0, stream, g, status, U, unknown, part, stop);
    hipLaunchKernelGGL(k_cleanup_switches, dim3(1), dim3(ABLOCK), 0, stream, (const double*)part, (int)nb, threshold,
                       ctl, stop);
    if (!edge_ok) {
        hipLaunchKernelGGL(k_spring_codes, dim3(nb), dim3(ABLOCK), 0, stream, g, (const uint8_t*)nullptr, w.code,
                           (const int*)&ctl->skip);
        hipLaunchKernelGGL(k_spring_rhs, dim3(nb), dim3(ABLOCK), 0, stream, op, U, b, X, (const int*)&ctl->skip);
        if (fill_solve(op, N, b, X, w.cf, fill_tol, fill_max_iter, pcg_ws, &ctl->skip, sync_every, nullptr,
                       fill_iterations, capped, stream) != hipSuccess)
            return fail(-5, "the cleanup's spring-fill solve failed");
        hipLaunchKernelGGL(k_median_pass1<false>, dim3(nb), dim3(ABLOCK), 0, stream, g, edge_ok,
                           (const uint8_t*)unknown, U, (const double*)X, absf, cnt, (const MedCtl*)ctl);
        hipLaunchKernelGGL(k_median_pass2<false>, dim3(nb), dim3(ABLOCK), 0, stream, g, edge_ok, status,
                           (const uint8_t*)unknown, U, U_hat, (const double*)absf, (const uint8_t*)cnt, threshold,
                           eps, U_out, bad, (const MedCtl*)ctl, stop);
    } else {
        hipLaunchKernelGGL(k_median_pass1<true>, dim3(nb), dim3(ABLOCK), 0, stream, g, edge_ok,
                           (const uint8_t*)unknown, U, (const double*)X, absf, cnt, (const MedCtl*)ctl);
        hipLaunchKernelGGL(k_median_pass2<true>, dim3(nb), dim3(ABLOCK), 0, stream, g, edge_ok, status,
                           (const uint8_t*)unknown, U, U_hat, (const double*)absf, (const uint8_t*)cnt, threshold,
                           eps, U_out, bad, (const MedCtl*)ctl, stop);
    }
    return launch_status();
}

size_t aldvc_spring_fill_workspace_bytes(int nx, int ny, int nz) {
    return FillWs(nullptr, (int64_t)nx * ny * nz).bytes;
}

int aldvc_spring_fill(int nx, int ny, int nz, const uint8_t* edge_ok, double* vals, double tol, int max_iter,
                      int* iterations, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    if (int rc = check_node_grid(nx, ny, nz)) return rc;
    if (!vals || !iterations || !workspace) return fail(-2, "null pointer");
    if (max_iter < 1) return fail(-2, "max_iter must be at least 1");
    if (workspace_bytes < aldvc_spring_fill_workspace_bytes(nx, ny, nz)) return fail(-3, "workspace too small");
    const NodeGrid g{nx, ny, nz};
    const int64_t N = (int64_t)nx * ny * nz;
    const FillWs w(workspace, N);
    uint8_t* unknown = w.unknown;
    double *b = w.b, *X = w.X;
    void* pcg_ws = w.pcg_ws;
    const unsigned nb = pcg::blocks_for(N);
    const SpringOp op{g, unknown, w.code};
    hipLaunchKernelGGL(k_spring_codes, dim3(nb), dim3(ABLOCK), 0, stream, g, edge_ok, w.code, (const int*)nullptr);
    hipLaunchKernelGGL(k_unknown_from_nan, dim3(nb), dim3(ABLOCK), 0, stream, N, (const double*)vals, unknown);
    hipLaunchKernelGGL(k_spring_rhs, dim3(nb), dim3(ABLOCK), 0, stream, op, (const double*)vals, b, X,
                       (const int*)nullptr);
    *iterations = 0;
    if (fill_solve(op, N, b, X, w.cf, tol, max_iter, pcg_ws, nullptr, 8, iterations, nullptr, nullptr, stream) !=
        hipSuccess)
        return fail(-5, "the spring-fill solve failed");
    hipLaunchKernelGGL(k_fill_merge, dim3(nb), dim3(ABLOCK), 0, stream, N, (const uint8_t*)unknown, (const double*)X,
                       vals);
    return launch_status();
}

size_t aldvc_dual_update_workspace_bytes(int64_t N) {
    (void)N;
    return sizeof(double) * 4 * pcg::MAX_BLOCKS;
}

int aldvc_dual_update(int64_t N, const double* U1n, const double* U2n, const double* F2n, double* U1, double* U2,
                      double* F2, double* W, double* v, int accumulate, int check_stop, double admm_tol, int step,
                      double* history, int* stop, int* steps, void* workspace, size_t workspace_bytes,
                      hipStream_t stream) {
    clear_error();
    if (N < 1) return fail(-2, "no nodes");
    if (!U1n || !U2n || !F2n || !U1 || !U2 || !F2 || !W || !v || !history || !stop || !steps || !workspace)
        return fail(-2, "null pointer");
    if (step < 0) return fail(-2, "negative step");
    if (workspace_bytes < aldvc_dual_update_workspace_bytes(N)) return fail(-3, "workspace too small");
    double* part = static_cast<double*>(workspace);
    const unsigned nb = pcg::blocks_for(N);
    hipLaunchKernelGGL(k_dual_update, dim3(nb), dim3(ABLOCK), 0, stream, N, U1n, U2n, F2n, U1, U2, F2, W, v, accumulate,
                       part, (const int*)stop);
    hipLaunchKernelGGL(k_dual_final, dim3(1), dim3(ABLOCK), 0, stream, (const double*)part, (int)nb, N, check_stop,
                       admm_tol, step, history, stop, steps);
    return launch_status();
}

}  // extern "C"
