// ADMM subproblem 2 on the GPU: the hex8 FEM operators of the uniform node grid applied
// matrix-free in float64, the right-hand side, the nodal gradient, and the Jacobi-preconditioned
// conjugate gradient for the three displacement components at once.
//
// Reference: solver/global_operators.py:58-70 (build_fem_operators), :147-170 (nodal_gradient),
// solver/subpb2_solver.py:45-155 (build_global_system, pcg_multi, global_rhs, solve_subpb2),
// mesh/hex8.py:74-105 (hex8_box_matrices).
//
// Every element of the grid is the same box, so an operator is one 8x8 element matrix and the
// per-element activity (elements dropped by the mask carry no equations). A node's row is the sum,
// over its up to 8 active elements, of the element matrix row of the node's corner: 64
// multiply-adds per component, no assembled sparse matrix (the reference's CSR at 161k nodes is
// 27 nonzeros per row, 35 MB per operator). Vectors are (N, 3) float64 over the whole grid, zero
// at inactive nodes, which is the algebra of the reference's active-node restriction. Dot products
// are reduced in a fixed order (per-block partials, then the block finishing last), so the
// iteration is deterministic, and the stopping rule is evaluated on the device (aldvc_pcg.h): the
// host either checks the control block every few iterations or, for graph capture, enqueues every
// iteration slot up front.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "aldvc_abi_util.h"
#include "aldvc_pcg.h"

using namespace aldvc;

namespace {

constexpr int GBLOCK = 256;
constexpr int PCG_SYNC_EVERY = 8;  // host-controlled solves read the control block this often

struct Grid {
    int nx, ny, nz;  // nodes per axis
    const uint8_t* elem_active;  // (nz-1, ny-1, nx-1)
    const uint8_t* node_active;  // (nz, ny, nx)
};

// corner c of an element at offset (dx, dy, dz): the hex8 order of mesh_setup (grid_mesh.py:96-107)
__device__ __constant__ int CORNER_DX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ __constant__ int CORNER_DY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ __constant__ int CORNER_DZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

__device__ inline int corner_of(int dx, int dy, int dz) {
    // inverse of the table above
    const int lo = dz * 4;
    if (dy == 0) return lo + dx;
    return lo + (dx ? 2 : 3);
}

// y[a] = sum over active elements e of a, corners b: Ae[la][lb] x[b]   (3 components)
// With TRANSPOSE the element matrix is read as Ae[lb][la].
template <bool TRANSPOSE, class X>
__device__ inline void element_row_of(const Grid& g, const double* __restrict__ Ae, const X& x, int ix, int iy, int iz,
                                      double acc[3]) {
    for (int e = 0; e < 8; ++e) {
        const int dx = e & 1, dy = (e >> 1) & 1, dz = e >> 2;  // this node's offset inside the element
        const int ex = ix - dx, ey = iy - dy, ez = iz - dz;
        if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx - 1 || ey >= g.ny - 1 || ez >= g.nz - 1) continue;
        if (!g.elem_active[((int64_t)ez * (g.ny - 1) + ey) * (g.nx - 1) + ex]) continue;
        const int la = corner_of(dx, dy, dz);
        for (int lb = 0; lb < 8; ++lb) {
            const int64_t b = ((int64_t)(ez + CORNER_DZ[lb]) * g.ny + (ey + CORNER_DY[lb])) * g.nx + (ex + CORNER_DX[lb]);
            const double w = TRANSPOSE ? Ae[lb * 8 + la] : Ae[la * 8 + lb];
            acc[0] += w * x(b, 0);
            acc[1] += w * x(b, 1);
            acc[2] += w * x(b, 2);
        }
    }
}

template <bool TRANSPOSE>
__device__ inline void element_row(const Grid& g, const double* __restrict__ Ae, const double* __restrict__ x, int ix,
                                   int iy, int iz, double acc[3]) {
    element_row_of<TRANSPOSE>(g, Ae, [x](int64_t b, int c) { return x[3 * b + c]; }, ix, iy, iz, acc);
}

__device__ inline void node_xyz(const Grid& g, int64_t a, int& ix, int& iy, int& iz) {
    ix = (int)(a % g.nx);
    iy = (int)((a / g.nx) % g.ny);
    iz = (int)(a / ((int64_t)g.nx * g.ny));
}

// y = A x on active nodes (0 elsewhere), A = element matrix Ae
__global__ void k_apply(Grid g, const double* __restrict__ Ae, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t N = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        double acc[3] = {0.0, 0.0, 0.0};
        if (g.node_active[a]) {
            int ix, iy, iz;
            node_xyz(g, a, ix, iy, iz);
            element_row<false>(g, Ae, x, ix, iy, iz, acc);
        }
        y[3 * a] = acc[0];
        y[3 * a + 1] = acc[1];
        y[3 * a + 2] = acc[2];
    }
}

// 1 / diagonal of A at a node (1 where the diagonal vanishes, subpb2_solver.py:52-53)
__device__ inline double inv_diag_at(const Grid& g, const double* __restrict__ Ae, int64_t a) {
    int ix, iy, iz;
    node_xyz(g, a, ix, iy, iz);
    double d = 0.0;
    for (int e = 0; e < 8; ++e) {
        const int dx = e & 1, dy = (e >> 1) & 1, dz = e >> 2;
        const int ex = ix - dx, ey = iy - dy, ez = iz - dz;
        if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx - 1 || ey >= g.ny - 1 || ez >= g.nz - 1) continue;
        if (!g.elem_active[((int64_t)ez * (g.ny - 1) + ey) * (g.nx - 1) + ex]) continue;
        d += Ae[corner_of(dx, dy, dz) * 9];
    }
    return fabs(d) > 1e-300 ? 1.0 / d : 1.0;
}

// The global system A = (beta + alpha) K + mu M over the active elements, for the PCG core.
struct FemOp {
    Grid g;
    const double* Ae;
    __device__ void inv_diag(int64_t a, double d[3]) const {
        const double v = inv_diag_at(g, Ae, a);
        d[0] = d[1] = d[2] = v;
    }
    template <class X>
    __device__ void apply(int64_t a, const X& x, double y[3]) const {
        y[0] = y[1] = y[2] = 0.0;
        if (!g.node_active[a]) return;
        int ix, iy, iz;
        node_xyz(g, a, ix, iy, iz);
        element_row_of<false>(g, Ae, x, ix, iy, iz, y);
    }
};

// b = beta sum_j G_j (F - W)[:, :, j] + mu M (U - v) + alpha K U   (subpb2_solver.py:108-127);
// mats = [K, M, G_x, G_y, G_z] (5 x 64)
__global__ void k_rhs(Grid g, const double* __restrict__ mats, double beta, double mu, double alpha,
                      const double* __restrict__ FmW_j /* (3, N, 3): (F - W)[:, i, j] per j */,
                      const double* __restrict__ Umv, const double* __restrict__ U, double* __restrict__ b) {
    const int64_t N = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        node_xyz(g, a, ix, iy, iz);
        double out[3] = {0.0, 0.0, 0.0};
        double t[3] = {0.0, 0.0, 0.0};
        element_row<false>(g, mats + 64, Umv, ix, iy, iz, t);
        for (int c = 0; c < 3; ++c) out[c] = mu * t[c];
        for (int j = 0; j < 3; ++j) {
            double s[3] = {0.0, 0.0, 0.0};
            element_row<false>(g, mats + 64 * (2 + j), FmW_j + 3 * N * j, ix, iy, iz, s);
            for (int c = 0; c < 3; ++c) out[c] += beta * s[c];
        }
        if (alpha != 0.0) {
            double k[3] = {0.0, 0.0, 0.0};
            element_row<false>(g, mats, U, ix, iy, iz, k);
            for (int c = 0; c < 3; ++c) out[c] += alpha * k[c];
        }
        b[3 * a] = out[0];
        b[3 * a + 1] = out[1];
        b[3 * a + 2] = out[2];
    }
}

// F[a, i, j] = (G_j^T U[:, i])[a] / mL[a], NaN at inactive nodes (global_operators.py:147-170)
__global__ void k_nodal_gradient(Grid g, const double* __restrict__ mats, const double* __restrict__ U,
                                 double* __restrict__ F, const int* __restrict__ gate) {
    if (gate && *gate) return;
    const int64_t N = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        if (!g.node_active[a]) {
            for (int k = 0; k < 9; ++k) F[9 * a + k] = __builtin_nan("");
            continue;
        }
        int ix, iy, iz;
        node_xyz(g, a, ix, iy, iz);
        // lumped mass: row sum of M
        double mL = 0.0;
        for (int e = 0; e < 8; ++e) {
            const int dx = e & 1, dy = (e >> 1) & 1, dz = e >> 2;
            const int ex = ix - dx, ey = iy - dy, ez = iz - dz;
            if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx - 1 || ey >= g.ny - 1 || ez >= g.nz - 1) continue;
            if (!g.elem_active[((int64_t)ez * (g.ny - 1) + ey) * (g.nx - 1) + ex]) continue;
            const int la = corner_of(dx, dy, dz);
            for (int lb = 0; lb < 8; ++lb) mL += mats[64 + la * 8 + lb];
        }
        const double inv = mL > 1e-15 ? 1.0 / mL : 1.0;
        for (int j = 0; j < 3; ++j) {
            double s[3] = {0.0, 0.0, 0.0};
            element_row<true>(g, mats + 64 * (2 + j), U, ix, iy, iz, s);
            for (int i = 0; i < 3; ++i) F[9 * a + 3 * i + j] = s[i] * inv;
        }
    }
}

// ---------------------------------------------------------------------------- ADMM global step
// b = beta sum_j G_j (F - W)[:, :, j] + mu M (U - v) + alpha K U from the per-node arrays
// F, W (N, 3, 3) and U, v (N, 3) directly; rows of inactive nodes are zero (subpb2_solver.py:
// 108-127, :143-146). X = U: the solve starts from U and inactive nodes keep it.
__global__ void k_admm_rhs(Grid g, const double* __restrict__ mats, double beta, double mu, double alpha,
                           const double* __restrict__ F, const double* __restrict__ W, const double* __restrict__ U,
                           const double* __restrict__ v, double* __restrict__ b, double* __restrict__ X,
                           const int* __restrict__ gate) {
    if (gate && *gate) return;
    const int64_t N = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < 3; ++c) X[3 * a + c] = U[3 * a + c];
        double out[3] = {0.0, 0.0, 0.0};
        if (g.node_active[a]) {
            int ix, iy, iz;
            node_xyz(g, a, ix, iy, iz);
            double t[3] = {0.0, 0.0, 0.0};
            element_row_of<false>(g, mats + 64, [U, v](int64_t n, int c) { return U[3 * n + c] - v[3 * n + c]; }, ix,
                                  iy, iz, t);
            for (int c = 0; c < 3; ++c) out[c] = mu * t[c];
            for (int j = 0; j < 3; ++j) {
                double s[3] = {0.0, 0.0, 0.0};
                element_row_of<false>(
                    g, mats + 64 * (2 + j), [F, W, j](int64_t n, int i) { return F[9 * n + 3 * i + j] - W[9 * n + 3 * i + j]; },
                    ix, iy, iz, s);
                for (int c = 0; c < 3; ++c) out[c] += beta * s[c];
            }
            if (alpha != 0.0) {
                double k[3] = {0.0, 0.0, 0.0};
                element_row<false>(g, mats, U, ix, iy, iz, k);
                for (int c = 0; c < 3; ++c) out[c] += alpha * k[c];
            }
        }
        b[3 * a] = out[0];
        b[3 * a + 1] = out[1];
        b[3 * a + 2] = out[2];
    }
}

// ---------------------------------------------------------------------------- finite differences
// The FD discretisation (global_operators.py:74-137): per axis j a difference operator D_j whose
// row a is (x[a+s] - x[a-s]) / 2h with both neighbours usable, the one-sided difference with one,
// and zero with none or at an invalid node; Kg = sum_j D_j^T D_j, M = diag(valid), G_j = D_j^T and
// the active nodes are the valid ones. Everything a row needs is one byte per node, its code:
// bit 0 the node is valid; for axis j (0 = x, 1 = y, 2 = z) bit 1 + 2j a usable neighbour one step
// back, bit 2 + 2j one step forward (on the grid, valid, and not across a cut edge). Invalid nodes
// have code 0, so a set neighbour bit implies both ends of the edge are valid. Kg is applied as the
// 13-point stencil it is: the column of node b in D_j has entries from rows b - s, b and b + s,
// each row value a difference of two entries of x.
struct FdGrid {
    int nx, ny, nz;
    const uint8_t* code;  // (nz, ny, nx)
    double ih[3];         // 1 / spacing along x, y, z
};

__device__ inline bool fd_back(uint8_t c, int j) { return (c >> (1 + 2 * j)) & 1u; }
__device__ inline bool fd_fwd(uint8_t c, int j) { return (c >> (2 + 2 * j)) & 1u; }

__device__ inline void fd_pos(const FdGrid& g, int64_t a, int j, int& p, int& size, int64_t& s) {
    if (j == 0) {
        p = (int)(a % g.nx), size = g.nx, s = 1;
    } else if (j == 1) {
        p = (int)((a / g.nx) % g.ny), size = g.ny, s = g.nx;
    } else {
        p = (int)(a / ((int64_t)g.nx * g.ny)), size = g.nz, s = (int64_t)g.nx * g.ny;
    }
}

// (D_j x)(r) for column c, from the row's code (0 for a zero row)
template <class X>
__device__ inline double fd_row(uint8_t cr, int j, double ih, int64_t r, int64_t s, const X& x, int c) {
    const bool b = fd_back(cr, j), f = fd_fwd(cr, j);
    if (b && f) return (-0.5 * ih) * x(r - s, c) + (0.5 * ih) * x(r + s, c);
    if (f) return (-ih) * x(r, c) + ih * x(r + s, c);
    if (b) return (-ih) * x(r - s, c) + ih * x(r, c);
    return 0.0;
}

// The entries of column a of D_j: w[0] from row a - s, w[1] from row a, w[2] from row a + s (0
// where the row does not reach a).
__device__ inline void fd_column(const FdGrid& g, int64_t a, int j, double w[3], int64_t& s) {
    int p, size;
    fd_pos(g, a, j, p, size, s);
    const double ih = g.ih[j];
    const uint8_t ca = g.code[a];
    w[0] = w[1] = w[2] = 0.0;
    if (p > 0) {  // row a - s reaches a when it has its forward neighbour
        const uint8_t cp = g.code[a - s];
        if (fd_fwd(cp, j)) w[0] = fd_back(cp, j) ? 0.5 * ih : ih;
    }
    if (fd_back(ca, j) != fd_fwd(ca, j)) w[1] = fd_fwd(ca, j) ? -ih : ih;  // one-sided rows use the node itself
    if (p < size - 1) {  // row a + s reaches a when it has its backward neighbour
        const uint8_t cn = g.code[a + s];
        if (fd_back(cn, j)) w[2] = fd_fwd(cn, j) ? -0.5 * ih : -ih;
    }
}

// (D_j^T y)(a) for column c of y, y read only at rows that reach a (so only at valid nodes)
template <class Y>
__device__ inline double fd_transpose(const FdGrid& g, int64_t a, int j, const Y& y, int c) {
    double w[3];
    int64_t s;
    fd_column(g, a, j, w, s);
    double acc = 0.0;
    if (w[0] != 0.0) acc += w[0] * y(a - s, c);
    if (w[1] != 0.0) acc += w[1] * y(a, c);
    if (w[2] != 0.0) acc += w[2] * y(a + s, c);
    return acc;
}

// (Kg x)(a) = sum_j (D_j^T (D_j x))(a), three columns
template <class X>
__device__ inline void fd_stiffness(const FdGrid& g, int64_t a, const X& x, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    for (int j = 0; j < 3; ++j) {
        double w[3];
        int64_t s;
        fd_column(g, a, j, w, s);
        const double ih = g.ih[j];
        for (int k = 0; k < 3; ++k) {
            if (w[k] == 0.0) continue;
            const int64_t r = a + (k - 1) * s;
            const uint8_t cr = g.code[r];
            for (int c = 0; c < 3; ++c) out[c] += w[k] * fd_row(cr, j, ih, r, s, x, c);
        }
    }
}

// A = cK Kg + cM M on the valid nodes, for the PCG core (subpb2_solver.py:45-74)
struct FdOp {
    FdGrid g;
    double cK, cM;
    __device__ void inv_diag(int64_t a, double d[3]) const {
        double v = 0.0;
        if (g.code[a] & 1u) {
            for (int j = 0; j < 3; ++j) {
                double w[3];
                int64_t s;
                fd_column(g, a, j, w, s);
                v += w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
            }
            v = cK * v + cM;
        }
        const double inv = fabs(v) > 1e-300 ? 1.0 / v : 1.0;
        d[0] = d[1] = d[2] = inv;
    }
    template <class X>
    __device__ void apply(int64_t a, const X& x, double y[3]) const {
        y[0] = y[1] = y[2] = 0.0;
        if (!(g.code[a] & 1u)) return;
        double k[3];
        fd_stiffness(g, a, x, k);
        for (int c = 0; c < 3; ++c) y[c] = cK * k[c] + cM * x(a, c);
    }
};

__global__ void k_fd_codes(int nx, int ny, int nz, const uint8_t* __restrict__ valid,
                           const uint8_t* __restrict__ edge_ok, uint8_t* __restrict__ code) {
    const FdGrid g{nx, ny, nz, nullptr, {0.0, 0.0, 0.0}};
    const int64_t N = (int64_t)nx * ny * nz;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        uint8_t c = 0;
        if (valid[a]) {
            c = 1;
            for (int j = 0; j < 3; ++j) {
                int p, size;
                int64_t s;
                fd_pos(g, a, j, p, size, s);
                if (p > 0 && valid[a - s] && (!edge_ok || edge_ok[3 * (a - s) + j])) c |= (uint8_t)(1u << (1 + 2 * j));
                if (p < size - 1 && valid[a + s] && (!edge_ok || edge_ok[3 * a + j])) c |= (uint8_t)(1u << (2 + 2 * j));
            }
        }
        code[a] = c;
    }
}

__global__ void k_fd_apply(FdOp op, int64_t N, const double* __restrict__ x, double* __restrict__ y) {
    auto xv = [x](int64_t b, int c) { return x[3 * b + c]; };
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        double out[3];
        op.apply(a, xv, out);
        for (int c = 0; c < 3; ++c) y[3 * a + c] = out[c];
    }
}

// b = mu M (U - v) + beta sum_j G_j (F - W)[:, :, j] + alpha Kg U in the reference's order of
// terms (subpb2_solver.py:108-127); rows of invalid nodes are zero. With X given, X = U (the
// solve starts from U and inactive nodes keep it).
__global__ void k_fd_rhs(FdGrid g, double beta, double mu, double alpha, const double* __restrict__ F,
                         const double* __restrict__ W, const double* __restrict__ U, const double* __restrict__ v,
                         double* __restrict__ b, double* __restrict__ X, const int* __restrict__ gate) {
    if (gate && *gate) return;
    const int64_t N = (int64_t)g.nx * g.ny * g.nz;
    auto uv = [U](int64_t n, int c) { return U[3 * n + c]; };
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        if (X)
            for (int c = 0; c < 3; ++c) X[3 * a + c] = U[3 * a + c];
        double out[3] = {0.0, 0.0, 0.0};
        if (g.code[a] & 1u) {
            for (int c = 0; c < 3; ++c) out[c] = mu * (U[3 * a + c] - v[3 * a + c]);
            for (int j = 0; j < 3; ++j) {
                auto fw = [F, W, j](int64_t n, int i) { return F[9 * n + 3 * i + j] - W[9 * n + 3 * i + j]; };
                for (int c = 0; c < 3; ++c) out[c] += beta * fd_transpose(g, a, j, fw, c);
            }
            if (alpha != 0.0) {
                double k[3];
                fd_stiffness(g, a, uv, k);
                for (int c = 0; c < 3; ++c) out[c] += alpha * k[c];
            }
        }
        b[3 * a] = out[0];
        b[3 * a + 1] = out[1];
        b[3 * a + 2] = out[2];
    }
}

// F[a, i, j] = (D_j U[:, i])(a), NaN at invalid nodes (global_operators.py:151-170)
__global__ void k_fd_gradient(FdGrid g, const double* __restrict__ U, double* __restrict__ F,
                              const int* __restrict__ gate) {
    if (gate && *gate) return;
    const int64_t N = (int64_t)g.nx * g.ny * g.nz;
    auto uv = [U](int64_t n, int c) { return U[3 * n + c]; };
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t ca = g.code[a];
        if (!(ca & 1u)) {
            for (int k = 0; k < 9; ++k) F[9 * a + k] = __builtin_nan("");
            continue;
        }
        for (int j = 0; j < 3; ++j) {
            int p, size;
            int64_t s;
            fd_pos(g, a, j, p, size, s);
            for (int i = 0; i < 3; ++i) F[9 * a + 3 * i + j] = fd_row(ca, j, g.ih[j], a, s, uv, i);
        }
    }
}

unsigned blocks_for(int64_t n) {
    const int64_t b = (n + GBLOCK - 1) / GBLOCK;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

int check_grid(int nx, int ny, int nz, const uint8_t* ea, const uint8_t* na) {
    if (nx < 2 || ny < 2 || nz < 2) return fail(-2, "the node grid needs at least 2 nodes per axis");
    if (!ea || !na) return fail(-2, "null element or node activity");
    return 0;
}

// the FD operators of a grid from its codes and spacing; a non-zero return is the error code
int fd_grid(const uint8_t* codes, int nx, int ny, int nz, double hx, double hy, double hz, FdGrid& g) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(-2, "the node grid needs at least 1 node per axis");
    if (!codes) return fail(-2, "null node codes");
    if (!(hx > 0.0 && hy > 0.0 && hz > 0.0)) return fail(-2, "the grid spacing must be positive");
    g = FdGrid{nx, ny, nz, codes, {1.0 / hx, 1.0 / hy, 1.0 / hz}};
    return 0;
}

}  // namespace

extern "C" {

int aldvc_fem_apply(const double* Ae, int nx, int ny, int nz, const uint8_t* elem_active, const uint8_t* node_active,
                    const double* x, double* y, hipStream_t stream) {
    clear_error();
    if (int rc = check_grid(nx, ny, nz, elem_active, node_active)) return rc;
    if (!Ae || !x || !y) return fail(-2, "null pointer");
    const Grid g{nx, ny, nz, elem_active, node_active};
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_apply, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, Ae, x, y);
    return launch_status();
}

int aldvc_fem_rhs(const double* mats, int nx, int ny, int nz, const uint8_t* elem_active, const uint8_t* node_active,
                  double beta, double mu, double alpha, const double* FmW_j, const double* Umv, const double* U,
                  double* b, hipStream_t stream) {
    clear_error();
    if (int rc = check_grid(nx, ny, nz, elem_active, node_active)) return rc;
    if (!mats || !FmW_j || !Umv || !U || !b) return fail(-2, "null pointer");
    const Grid g{nx, ny, nz, elem_active, node_active};
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_rhs, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, mats, beta, mu, alpha, FmW_j, Umv, U, b);
    return launch_status();
}

int aldvc_fem_nodal_gradient(const double* mats, int nx, int ny, int nz, const uint8_t* elem_active,
                             const uint8_t* node_active, const double* U, double* F, hipStream_t stream) {
    clear_error();
    if (int rc = check_grid(nx, ny, nz, elem_active, node_active)) return rc;
    if (!mats || !U || !F) return fail(-2, "null pointer");
    const Grid g{nx, ny, nz, elem_active, node_active};
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_nodal_gradient, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, mats, U, F, (const int*)nullptr);
    return launch_status();
}

size_t aldvc_pcg_workspace_bytes(int nx, int ny, int nz) {
    return pcg::workspace_bytes((int64_t)nx * ny * nz);
}

int aldvc_pcg_fem(const double* Ae, int nx, int ny, int nz, const uint8_t* elem_active, const uint8_t* node_active,
                  const double* B, double* X, double tol, int max_iter, int* iterations, void* workspace,
                  size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    if (int rc = check_grid(nx, ny, nz, elem_active, node_active)) return rc;
    if (!Ae || !B || !X || !iterations || !workspace) return fail(-2, "null pointer");
    if (max_iter < 1) return fail(-2, "max_iter must be at least 1");
    if (workspace_bytes < aldvc_pcg_workspace_bytes(nx, ny, nz)) return fail(-3, "workspace too small");
    const FemOp op{Grid{nx, ny, nz, elem_active, node_active}, Ae};
    *iterations = 0;
    if (pcg::solve(op, (int64_t)nx * ny * nz, B, X, tol, max_iter, workspace, nullptr, PCG_SYNC_EVERY, iterations,
                   nullptr, stream) != hipSuccess)
        return fail(-5, "the conjugate-gradient solve failed");
    return launch_status();
}

int aldvc_pcg_fem_async(const double* Ae, int nx, int ny, int nz, const uint8_t* elem_active,
                        const uint8_t* node_active, const double* B, double* X, double tol, int max_iter,
                        int* iterations_dev, const int* gate, void* workspace, size_t workspace_bytes,
                        hipStream_t stream) {
    clear_error();
    if (int rc = check_grid(nx, ny, nz, elem_active, node_active)) return rc;
    if (!Ae || !B || !X || !workspace) return fail(-2, "null pointer");
    if (max_iter < 1) return fail(-2, "max_iter must be at least 1");
    if (workspace_bytes < aldvc_pcg_workspace_bytes(nx, ny, nz)) return fail(-3, "workspace too small");
    const FemOp op{Grid{nx, ny, nz, elem_active, node_active}, Ae};
    (void)pcg::solve(op, (int64_t)nx * ny * nz, B, X, tol, max_iter, workspace, gate, 0, nullptr, iterations_dev, stream);
    return launch_status();
}

size_t aldvc_admm_global_workspace_bytes(int nx, int ny, int nz) {
    const size_t N = (size_t)nx * ny * nz;
    return 256 + sizeof(double) * 3 * N + aldvc_pcg_workspace_bytes(nx, ny, nz);
}

int aldvc_admm_global_step(const double* mats, const double* Ae, int nx, int ny, int nz, const uint8_t* elem_active,
                           const uint8_t* node_active, double beta, double mu, double alpha, const double* U,
                           const double* F, const double* W, const double* v, double tol, int max_iter, double* U2,
                           double* F2, int* iterations_dev, const int* gate, int sync_every, int* capped,
                           void* workspace, size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    if (int rc = check_grid(nx, ny, nz, elem_active, node_active)) return rc;
    if (!mats || !Ae || !U || !F || !W || !v || !U2 || !F2 || !workspace) return fail(-2, "null pointer");
    if (max_iter < 1) return fail(-2, "max_iter must be at least 1");
    if (sync_every < 0) return fail(-2, "sync_every must be 0 (all iterations enqueued) or positive");
    if (workspace_bytes < aldvc_admm_global_workspace_bytes(nx, ny, nz)) return fail(-3, "workspace too small");
    const Grid g{nx, ny, nz, elem_active, node_active};
    const int64_t N = (int64_t)nx * ny * nz;
    double* b = static_cast<double*>(workspace);
    void* pcg_ws = reinterpret_cast<char*>(workspace) + 256 + sizeof(double) * 3 * (size_t)N;
    hipLaunchKernelGGL(k_admm_rhs, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, mats, beta, mu, alpha, F, W, U, v,
                       b, U2, gate);
    if (pcg::solve(FemOp{g, Ae}, N, (const double*)b, U2, tol, max_iter, pcg_ws, gate, sync_every, nullptr,
                   iterations_dev, stream, capped) != hipSuccess)
        return fail(-5, "the global-step solve failed");
    hipLaunchKernelGGL(k_nodal_gradient, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, mats, (const double*)U2, F2,
                       gate);
    return launch_status();
}

// ---------------------------------------------------------------------------- finite differences
int aldvc_fd_codes(int nx, int ny, int nz, const uint8_t* node_valid, const uint8_t* edge_ok, uint8_t* codes,
                   hipStream_t stream) {
    clear_error();
    if (nx < 1 || ny < 1 || nz < 1) return fail(-2, "the node grid needs at least 1 node per axis");
    if (!node_valid || !codes) return fail(-2, "null pointer");
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_fd_codes, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, nx, ny, nz, node_valid, edge_ok, codes);
    return launch_status();
}

int aldvc_fd_apply(const uint8_t* codes, int nx, int ny, int nz, double hx, double hy, double hz, double cK, double cM,
                   const double* x, double* y, hipStream_t stream) {
    clear_error();
    FdGrid g;
    if (int rc = fd_grid(codes, nx, ny, nz, hx, hy, hz, g)) return rc;
    if (!x || !y) return fail(-2, "null pointer");
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_fd_apply, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, FdOp{g, cK, cM}, N, x, y);
    return launch_status();
}

int aldvc_fd_rhs(const uint8_t* codes, int nx, int ny, int nz, double hx, double hy, double hz, double beta, double mu,
                 double alpha, const double* U, const double* F, const double* W, const double* v, double* b,
                 hipStream_t stream) {
    clear_error();
    FdGrid g;
    if (int rc = fd_grid(codes, nx, ny, nz, hx, hy, hz, g)) return rc;
    if (!U || !F || !W || !v || !b) return fail(-2, "null pointer");
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_fd_rhs, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, beta, mu, alpha, F, W, U, v, b,
                       (double*)nullptr, (const int*)nullptr);
    return launch_status();
}

int aldvc_fd_nodal_gradient(const uint8_t* codes, int nx, int ny, int nz, double hx, double hy, double hz,
                            const double* U, double* F, hipStream_t stream) {
    clear_error();
    FdGrid g;
    if (int rc = fd_grid(codes, nx, ny, nz, hx, hy, hz, g)) return rc;
    if (!U || !F) return fail(-2, "null pointer");
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_fd_gradient, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, U, F, (const int*)nullptr);
    return launch_status();
}

int aldvc_pcg_fd(const uint8_t* codes, int nx, int ny, int nz, double hx, double hy, double hz, double cK, double cM,
                 const double* B, double* X, double tol, int max_iter, int* iterations, void* workspace,
                 size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    FdGrid g;
    if (int rc = fd_grid(codes, nx, ny, nz, hx, hy, hz, g)) return rc;
    if (!B || !X || !iterations || !workspace) return fail(-2, "null pointer");
    if (max_iter < 1) return fail(-2, "max_iter must be at least 1");
    if (workspace_bytes < aldvc_pcg_workspace_bytes(nx, ny, nz)) return fail(-3, "workspace too small");
    *iterations = 0;
    if (pcg::solve(FdOp{g, cK, cM}, (int64_t)nx * ny * nz, B, X, tol, max_iter, workspace, nullptr, PCG_SYNC_EVERY,
                   iterations, nullptr, stream) != hipSuccess)
        return fail(-5, "the conjugate-gradient solve failed");
    return launch_status();
}

int aldvc_admm_global_step_fd(const uint8_t* codes, int nx, int ny, int nz, double hx, double hy, double hz,
                              double beta, double mu, double alpha, const double* U, const double* F, const double* W,
                              const double* v, double tol, int max_iter, double* U2, double* F2, int* iterations_dev,
                              const int* gate, int sync_every, int* capped, void* workspace, size_t workspace_bytes,
                              hipStream_t stream) {
    clear_error();
    FdGrid g;
    if (int rc = fd_grid(codes, nx, ny, nz, hx, hy, hz, g)) return rc;
    if (!U || !F || !W || !v || !U2 || !F2 || !workspace) return fail(-2, "null pointer");
    if (max_iter < 1) return fail(-2, "max_iter must be at least 1");
    if (sync_every < 0) return fail(-2, "sync_every must be 0 (all iterations enqueued) or positive");
    if (workspace_bytes < aldvc_admm_global_workspace_bytes(nx, ny, nz)) return fail(-3, "workspace too small");
    const int64_t N = (int64_t)nx * ny * nz;
    double* b = static_cast<double*>(workspace);
    void* pcg_ws = reinterpret_cast<char*>(workspace) + 256 + sizeof(double) * 3 * (size_t)N;
    hipLaunchKernelGGL(k_fd_rhs, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, beta, mu, alpha, F, W, U, v, b, U2,
                       gate);
    if (pcg::solve(FdOp{g, beta + alpha, mu}, N, (const double*)b, U2, tol, max_iter, pcg_ws, gate, sync_every, nullptr,
                   iterations_dev, stream, capped) != hipSuccess)
        return fail(-5, "the global-step solve failed");
    hipLaunchKernelGGL(k_fd_gradient, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, (const double*)U2, F2, gate);
    return launch_status();
}

}  // extern "C"
