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
// are reduced in a fixed order (per-block partials, then one block), so the iteration is
// deterministic; the host reads the three residual norms once per iteration for the stopping rule.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "aldvc_abi_util.h"

using namespace aldvc;

namespace {

constexpr int GBLOCK = 256;
constexpr int DOT_BLOCKS = 512;  // partial sums per dot product

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
template <bool TRANSPOSE>
__device__ inline void element_row(const Grid& g, const double* __restrict__ Ae, const double* __restrict__ x, int ix,
                                   int iy, int iz, double acc[3]) {
    for (int e = 0; e < 8; ++e) {
        const int dx = e & 1, dy = (e >> 1) & 1, dz = e >> 2;  // this node's offset inside the element
        const int ex = ix - dx, ey = iy - dy, ez = iz - dz;
        if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx - 1 || ey >= g.ny - 1 || ez >= g.nz - 1) continue;
        if (!g.elem_active[((int64_t)ez * (g.ny - 1) + ey) * (g.nx - 1) + ex]) continue;
        const int la = corner_of(dx, dy, dz);
        for (int lb = 0; lb < 8; ++lb) {
            const int64_t b = ((int64_t)(ez + CORNER_DZ[lb]) * g.ny + (ey + CORNER_DY[lb])) * g.nx + (ex + CORNER_DX[lb]);
            const double w = TRANSPOSE ? Ae[lb * 8 + la] : Ae[la * 8 + lb];
            acc[0] += w * x[3 * b];
            acc[1] += w * x[3 * b + 1];
            acc[2] += w * x[3 * b + 2];
        }
    }
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

// diagonal of A on active nodes and 1 / diag (1 where it vanishes, subpb2_solver.py:52-53)
__global__ void k_inv_diag(Grid g, const double* __restrict__ Ae, double* __restrict__ inv_diag) {
    const int64_t N = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N; a += (int64_t)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        node_xyz(g, a, ix, iy, iz);
        double d = 0.0;
        for (int e = 0; e < 8; ++e) {
            const int dx = e & 1, dy = (e >> 1) & 1, dz = e >> 2;
            const int ex = ix - dx, ey = iy - dy, ez = iz - dz;
            if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx - 1 || ey >= g.ny - 1 || ez >= g.nz - 1) continue;
            if (!g.elem_active[((int64_t)ez * (g.ny - 1) + ey) * (g.nx - 1) + ex]) continue;
            const int la = corner_of(dx, dy, dz);
            d += Ae[la * 9];
        }
        inv_diag[a] = fabs(d) > 1e-300 ? 1.0 / d : 1.0;
    }
}

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
                                 double* __restrict__ F) {
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

// ---------------------------------------------------------------------------- reductions
// part[blk][c] = sum over this block's nodes of u[a][c] * v[a][c]
__global__ __launch_bounds__(GBLOCK) void k_dot3_part(const double* __restrict__ u, const double* __restrict__ v,
                                                      int64_t N, double* __restrict__ part) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t a = (int64_t)blockIdx.x * GBLOCK + threadIdx.x; a < N; a += (int64_t)gridDim.x * GBLOCK)
        for (int c = 0; c < 3; ++c) s[c] += u[3 * a + c] * v[3 * a + c];
    __shared__ double red[3][GBLOCK];
    for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = s[c];
    __syncthreads();
    for (int off = GBLOCK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__device__ inline void sum_parts(const double* __restrict__ part, int nparts, double out[3]) {
    __shared__ double red[3][GBLOCK];
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nparts; i += GBLOCK)
        for (int c = 0; c < 3; ++c) s[c] += part[3 * i + c];
    for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = s[c];
    __syncthreads();
    for (int off = GBLOCK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + off];
        __syncthreads();
    }
    for (int c = 0; c < 3; ++c) out[c] = red[c][0];
}

// Scalars of the iteration, in device memory: [0:3] rz, [3:6] alpha, [6:9] beta, [9:12] |R|^2
enum { S_RZ = 0, S_ALPHA = 3, S_BETA = 6, S_RR = 9, S_COUNT = 12 };

// which = 0: rz = <R, Z>; 1: alpha = rz / <P, AP> (0 when the denominator vanishes);
// 2: |R|^2; 3: beta = <R, Z> / rz (0 when rz vanishes), then rz = <R, Z>
__global__ __launch_bounds__(GBLOCK) void k_dot3_final(const double* __restrict__ part, int nparts, int which,
                                                       double* __restrict__ sc) {
    double d[3];
    sum_parts(part, nparts, d);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 3; ++c) {
            if (which == 0) {
                sc[S_RZ + c] = d[c];
            } else if (which == 1) {
                sc[S_ALPHA + c] = d[c] == 0.0 ? 0.0 : sc[S_RZ + c] / d[c];
            } else if (which == 2) {
                sc[S_RR + c] = d[c];
            } else {
                const double rz = sc[S_RZ + c];
                sc[S_BETA + c] = rz == 0.0 ? 0.0 : d[c] / rz;
                sc[S_RZ + c] = d[c];
            }
        }
    }
}

// R = B - A X (A X already in R), Z = inv_diag R, P = Z
__global__ void k_pcg_init(const double* __restrict__ B, const double* __restrict__ inv_diag, int64_t N,
                           double* __restrict__ R, double* __restrict__ Z, double* __restrict__ P) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * N; i += (int64_t)gridDim.x * blockDim.x) {
        const double r = B[i] - R[i];
        R[i] = r;
        Z[i] = inv_diag[i / 3] * r;
        P[i] = Z[i];
    }
}

__global__ void k_pcg_xr(const double* __restrict__ sc, const double* __restrict__ P, const double* __restrict__ AP,
                         int64_t N, double* __restrict__ X, double* __restrict__ R) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * N; i += (int64_t)gridDim.x * blockDim.x) {
        const double al = sc[S_ALPHA + i % 3];
        X[i] += P[i] * al;
        R[i] -= AP[i] * al;
    }
}

__global__ void k_pcg_z(const double* __restrict__ inv_diag, const double* __restrict__ R, int64_t N,
                        double* __restrict__ Z) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * N; i += (int64_t)gridDim.x * blockDim.x)
        Z[i] = inv_diag[i / 3] * R[i];
}

__global__ void k_pcg_p(const double* __restrict__ sc, const double* __restrict__ Z, int64_t N, double* __restrict__ P) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * N; i += (int64_t)gridDim.x * blockDim.x)
        P[i] = Z[i] + P[i] * sc[S_BETA + i % 3];
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
    hipLaunchKernelGGL(k_nodal_gradient, dim3(blocks_for(N)), dim3(GBLOCK), 0, stream, g, mats, U, F);
    return launch_status();
}

size_t aldvc_pcg_workspace_bytes(int nx, int ny, int nz) {
    const size_t N = (size_t)nx * ny * nz;
    return sizeof(double) * (4 * 3 * N + N + 3 * DOT_BLOCKS + S_COUNT + 16);
}

int aldvc_pcg_fem(const double* Ae, int nx, int ny, int nz, const uint8_t* elem_active, const uint8_t* node_active,
                  const double* B, double* X, double tol, int max_iter, int* iterations, void* workspace,
                  size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    if (int rc = check_grid(nx, ny, nz, elem_active, node_active)) return rc;
    if (!Ae || !B || !X || !iterations || !workspace) return fail(-2, "null pointer");
    if (max_iter < 1) return fail(-2, "max_iter must be at least 1");
    if (workspace_bytes < aldvc_pcg_workspace_bytes(nx, ny, nz)) return fail(-3, "workspace too small");
    const Grid g{nx, ny, nz, elem_active, node_active};
    const int64_t N = (int64_t)nx * ny * nz;
    double* R = static_cast<double*>(workspace);
    double* Z = R + 3 * N;
    double* P = Z + 3 * N;
    double* AP = P + 3 * N;
    double* inv_diag = AP + 3 * N;
    double* part = inv_diag + N;
    double* sc = part + 3 * DOT_BLOCKS;
    const unsigned nb = blocks_for(N), nv = blocks_for(3 * N);
    auto dot = [&](const double* u, const double* v, int which) {
        hipLaunchKernelGGL(k_dot3_part, dim3(DOT_BLOCKS), dim3(GBLOCK), 0, stream, u, v, N, part);
        hipLaunchKernelGGL(k_dot3_final, dim3(1), dim3(GBLOCK), 0, stream, (const double*)part, DOT_BLOCKS, which, sc);
    };
    // |B| per column, on the host: zero norms count as 1 (subpb2_solver.py:92-93)
    double bn2[S_COUNT];
    dot(B, B, 2);
    if (hipMemcpyAsync(bn2, sc, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fail(-5, "reading the right-hand side norm failed");
    double bnorm[3];
    for (int c = 0; c < 3; ++c) bnorm[c] = bn2[S_RR + c] > 0.0 ? sqrt(bn2[S_RR + c]) : 1.0;

    hipLaunchKernelGGL(k_inv_diag, dim3(nb), dim3(GBLOCK), 0, stream, g, Ae, inv_diag);
    hipLaunchKernelGGL(k_apply, dim3(nb), dim3(GBLOCK), 0, stream, g, Ae, (const double*)X, R);
    hipLaunchKernelGGL(k_pcg_init, dim3(nv), dim3(GBLOCK), 0, stream, B, (const double*)inv_diag, N, R, Z, P);
    dot(R, Z, 0);
    int it = 0;
    for (it = 1; it <= max_iter; ++it) {
        hipLaunchKernelGGL(k_apply, dim3(nb), dim3(GBLOCK), 0, stream, g, Ae, (const double*)P, AP);
        dot(P, AP, 1);
        hipLaunchKernelGGL(k_pcg_xr, dim3(nv), dim3(GBLOCK), 0, stream, (const double*)sc, (const double*)P,
                           (const double*)AP, N, X, R);
        dot(R, R, 2);
        double s[S_COUNT];
        if (hipMemcpyAsync(s, sc, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)
            return fail(-5, "reading the residual norms failed");
        bool done = true;
        for (int c = 0; c < 3; ++c) done = done && (sqrt(s[S_RR + c]) / bnorm[c] < tol);
        if (done) break;
        hipLaunchKernelGGL(k_pcg_z, dim3(nv), dim3(GBLOCK), 0, stream, (const double*)inv_diag, (const double*)R, N, Z);
        dot(R, Z, 3);
        hipLaunchKernelGGL(k_pcg_p, dim3(nv), dim3(GBLOCK), 0, stream, (const double*)sc, (const double*)Z, N, P);
    }
    *iterations = it > max_iter ? max_iter : it;
    return launch_status();
}

}  // extern "C"
