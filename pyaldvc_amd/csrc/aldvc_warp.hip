// Differentiable warp-residual: the ZNSSD of every subset under a displacement field given on a
// coarse hex8 node grid (trilinear shape functions), and its backward pass with respect to the
// node displacements and the deformed volume.
//
// Reference: the forward is the reference's ZNCC of a warp (solver/numba_kernels.py:601-616,
// :1354-1387, evaluate_zncc_parallel) with the per-voxel affine warp replaced by the FE
// interpolation of the node field (mesh/hex8.py:58-61, hex8_shape); the reference has no backward.
//
// Forward, one 256-lane workgroup per subset: voxel p of the subset moves to p + u(p), with
// u(p) = sum_a N_a(p) u_a over the 8 nodes of the grid cell holding p; g is sampled there
// (linear / Keys cubic / cubic B-spline); float64 sums give the centred statistics and
// C = sum (f~ - g~)^2 = 2 (1 - ZNCC), f~ = (f - m_f) / b_f, g~ = (g - m_g) / b_g.
// Backward: dC/dg_v = -2 (f~_v - ZNCC g~_v) / b_g per sampled voxel. The interpolation weights and
// their derivatives are recomputed (nothing per voxel is stored between the passes):
// dL/dG[tap] += dL/dg_v w_tap (float32 atomics into the volume gradient) and
// dL/du_a += N_a(p) dL/dg_v grad g(p + u(p)), accumulated per workgroup in LDS over the nodes the
// subset touches, then added to the node gradient with one float64 atomic per node and component.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "aldvc_abi_util.h"
#include "aldvc_device.h"

using namespace aldvc;

namespace {

constexpr int WBLOCK = 256;
constexpr int WWAVES = WBLOCK / 64;
constexpr int LDS_NODES = 1024;  // node gradients accumulated in LDS per workgroup
constexpr int NSTAT = 8;         // per-subset statistics: n, m_f, b_f, m_g, b_g, ZNCC, valid, C

struct NodeGrid {
    const float* u;  // (gnz, gny, gnx, 3)
    int gnx, gny, gnz;
    float ox, oy, oz;   // position of node (0, 0, 0)
    float isx, isy, isz;  // 1 / spacing
};

struct Cell {
    int cx, cy, cz;      // cell index (node of the low corner)
    float tx, ty, tz;    // local coordinates in [0, 1] inside the grid
};

// cell of absolute voxel position (x, y, z); positions beyond the grid extrapolate the edge cell
__device__ inline Cell cell_of(const NodeGrid& G, float x, float y, float z) {
    Cell c;
    const float qx = (x - G.ox) * G.isx, qy = (y - G.oy) * G.isy, qz = (z - G.oz) * G.isz;
    c.cx = min(max((int)floorf(qx), 0), G.gnx - 2);
    c.cy = min(max((int)floorf(qy), 0), G.gny - 2);
    c.cz = min(max((int)floorf(qz), 0), G.gnz - 2);
    c.tx = qx - (float)c.cx;
    c.ty = qy - (float)c.cy;
    c.tz = qz - (float)c.cz;
    return c;
}

// trilinear shape function of corner (i, j, k) of the cell (hex8_shape in grid-aligned form)
__device__ inline float shape(const Cell& c, int i, int j, int k) {
    return (i ? c.tx : 1.0f - c.tx) * (j ? c.ty : 1.0f - c.ty) * (k ? c.tz : 1.0f - c.tz);
}

__device__ inline void fe_displacement(const NodeGrid& G, const Cell& c, float u[3]) {
    u[0] = u[1] = u[2] = 0.0f;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const float N = shape(c, i, j, k);
                const float* un = G.u + 3 * (((int64_t)(c.cz + k) * G.gny + (c.cy + j)) * G.gnx + (c.cx + i));
                u[0] += N * un[0];
                u[1] += N * un[1];
                u[2] += N * un[2];
            }
}

// 1-D interpolation weights and their derivatives (keys4 / bspline4 / linear of the sampler)
template <int MODE>
__device__ inline void weights_d(float t, float w[4], float d[4]) {
    if constexpr (MODE == INTERP_CUBIC) {
        keys4(t, w[0], w[1], w[2], w[3]);
        d[0] = (-1.5f * t + 2.0f) * t - 0.5f;
        d[1] = (4.5f * t - 5.0f) * t;
        d[2] = (-4.5f * t + 4.0f) * t + 0.5f;
        d[3] = (1.5f * t - 1.0f) * t;
    } else if constexpr (MODE == INTERP_BSPLINE) {
        bspline4(t, w[0], w[1], w[2], w[3]);
        const float s = 1.0f - t;
        d[0] = -0.5f * s * s;
        d[1] = (1.5f * t - 2.0f) * t;
        d[2] = (-1.5f * t + 1.0f) * t + 0.5f;
        d[3] = 0.5f * t * t;
    } else {
        w[0] = 1.0f - t;
        w[1] = t;
        w[2] = w[3] = 0.0f;
        d[0] = -1.0f;
        d[1] = 1.0f;
        d[2] = d[3] = 0.0f;
    }
}

// Sample g at the anchor-relative position (rx, ry, rz) and, with GRAD, its spatial gradient;
// with SCATTER, add `scale * weight` of every tap into gg (the volume gradient) instead.
// The caller has checked the sampling domain; the stencil base is clamped to n - 3 as the sampler's.
template <int MODE, bool GRAD, bool SCATTER>
__device__ inline float sample_grad(const float* __restrict__ base, int rowp, int planep, float rx, float ry, float rz,
                                    int kx, int ky, int kz, float grad[3], float* gg, float scale) {
    constexpr int T = MODE == INTERP_LINEAR ? 2 : 4, LO = MODE == INTERP_LINEAR ? 0 : 1;
    int ix = (int)floorf(rx), iy = (int)floorf(ry), iz = (int)floorf(rz);
    ix = ix > kx ? kx : ix;
    iy = iy > ky ? ky : iy;
    iz = iz > kz ? kz : iz;
    float wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
    weights_d<MODE>(rx - (float)ix, wx, dx);
    weights_d<MODE>(ry - (float)iy, wy, dy);
    weights_d<MODE>(rz - (float)iz, wz, dz);
    const int64_t o = (int64_t)(iz - LO) * planep + (int64_t)(iy - LO) * rowp + (ix - LO);
    float val = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
    for (int k = 0; k < T; ++k) {
        float pv = 0.0f, py = 0.0f, px = 0.0f;
#pragma unroll
        for (int j = 0; j < T; ++j) {
            const int64_t r = o + (int64_t)k * planep + (int64_t)j * rowp;
            float rv = 0.0f, rd = 0.0f;
#pragma unroll
            for (int i = 0; i < T; ++i) {
                if constexpr (SCATTER) {
                    unsafeAtomicAdd(gg + r + i, scale * wx[i] * wy[j] * wz[k]);
                } else {
                    const float s = base[r + i];
                    rv += wx[i] * s;
                    if constexpr (GRAD) rd += dx[i] * s;
                }
            }
            pv += wy[j] * rv;
            if constexpr (GRAD) {
                py += dy[j] * rv;
                px += wy[j] * rd;
            }
        }
        val += wz[k] * pv;
        if constexpr (GRAD) {
            gz += dz[k] * pv;
            gy += wz[k] * py;
            gx += wz[k] * px;
        }
    }
    if constexpr (GRAD) {
        grad[0] = gx;
        grad[1] = gy;
        grad[2] = gz;
    }
    return val;
}

struct WarpArgs {
    Volumes V;
    Subset s;
    NodeGrid G;
    const int64_t* centers;  // (S, 3) [x, y, z]
    int64_t S;
    double* C;      // (S)
    double* stats;  // (S, NSTAT)
};

template <int MODE>
__global__ __launch_bounds__(WBLOCK) void k_warp_znssd(WarpArgs a) {
    const int64_t n = blockIdx.x;
    if (n >= a.S) return;
    const Volumes& V = a.V;
    const Subset& s = a.s;
    const int tid = threadIdx.x;
    const int64_t x0 = a.centers[3 * n], y0 = a.centers[3 * n + 1], z0 = a.centers[3 * n + 2];
    double* st = a.stats + NSTAT * n;
    const bool inside_vol = x0 - s.hx >= 0 && x0 + s.hx < V.nx && y0 - s.hy >= 0 && y0 + s.hy < V.ny &&
                            z0 - s.hz >= 0 && z0 + s.hz < V.nz;
    if (!inside_vol) {
        if (tid == 0) {
            a.C[n] = NAN;
            for (int q = 0; q < NSTAT; ++q) st[q] = 0.0;
        }
        return;
    }
    constexpr int NQ = 7;  // n, oob, f', f'^2, g', g'^2, f'g'
    __shared__ double red[WWAVES * NQ];
    __shared__ double res[NQ];
    const Domain dom = make_domain(V, (int)x0, (int)y0, (int)z0);
    const int64_t obase = lin(V, z0, y0, x0);
    const float* gbase = V.g + obase;
    const int rowp = V.nx, planep = V.nx * V.ny;
    const float pf = V.f[obase];
    float m[NQ] = {0, 0, 0, 0, 0, 0, 0};
    for (int v = tid; v < s.S; v += WBLOCK) {
        const int iz = v / (s.sx * s.sy), rem = v - iz * s.sx * s.sy, iy = rem / s.sx, ix = rem - iy * s.sx;
        const int dx = -s.hx + s.stride * ix, dy = -s.hy + s.stride * iy, dz = -s.hz + s.stride * iz;
        const float fv = V.f[obase + (int64_t)dz * planep + (int64_t)dy * rowp + dx];
        const Cell c = cell_of(a.G, (float)(x0 + dx), (float)(y0 + dy), (float)(z0 + dz));
        float u[3];
        fe_displacement(a.G, c, u);
        const float rx = (float)dx + u[0], ry = (float)dy + u[1], rz = (float)dz + u[2];
        if (!inside(dom, rx, ry, rz)) {
            m[1] = 1.0f;
            continue;
        }
        const float gv = sample_grad<MODE, false, false>(gbase, rowp, planep, rx, ry, rz, dom.kx, dom.ky, dom.kz,
                                                          nullptr, nullptr, 0.0f);
        if (gv != gv) continue;
        const float fp = fv - pf, gp = gv - pf;
        m[0] += 1.0f;
        m[2] += fp;
        m[3] += fp * fp;
        m[4] += gp;
        m[5] += gp * gp;
        m[6] += fp * gp;
    }
    block_sum<NQ, WWAVES>(m, res, red);
    if (tid == 0) {
        const double nv = res[0];
        bool ok = !(res[1] > 0.0) && nv >= MIN_SUBSET_VOXELS && nv >= MIN_VALID_FRACTION * (double)s.S;
        double mf = 0, mg = 0, bf = 0, bg = 0, z = NAN;
        if (ok) {
            mf = res[2] / nv;
            mg = res[4] / nv;
            double ssf = res[3] - nv * mf * mf, ssg = res[5] - nv * mg * mg;
            ssf = ssf < 0.0 ? 0.0 : ssf;
            ssg = ssg < 0.0 ? 0.0 : ssg;
            bf = sqrt(ssf > 1e-30 ? ssf : 1e-30);
            bg = sqrt(ssg > 1e-30 ? ssg : 1e-30);
            z = (res[6] - nv * mf * mg) / (bf * bg);
            mf += pf;
            mg += pf;
        }
        st[0] = nv;
        st[1] = mf;
        st[2] = bf;
        st[3] = mg;
        st[4] = bg;
        st[5] = z;
        st[6] = ok ? 1.0 : 0.0;
        st[7] = ok ? 2.0 * (1.0 - z) : NAN;
        a.C[n] = st[7];
    }
}

struct WarpBackArgs {
    WarpArgs f;
    const double* grad_C;  // (S)
    double* grad_u;        // (gnz, gny, gnx, 3), accumulated
    float* grad_g;         // (nz, ny, nx) or null, accumulated
};

template <int MODE>
__global__ __launch_bounds__(WBLOCK) void k_warp_znssd_backward(WarpBackArgs b) {
    const WarpArgs& a = b.f;
    const int64_t n = blockIdx.x;
    if (n >= a.S) return;
    const double* st = a.stats + NSTAT * n;
    const double gC = b.grad_C[n];
    if (st[6] == 0.0 || gC == 0.0 || !isfinite(gC)) return;  // invalid subset or no upstream gradient
    const Volumes& V = a.V;
    const Subset& s = a.s;
    const NodeGrid& G = a.G;
    const int tid = threadIdx.x;
    const int64_t x0 = a.centers[3 * n], y0 = a.centers[3 * n + 1], z0 = a.centers[3 * n + 2];
    // the nodes this subset's voxels can touch: the cells of its corners, plus one
    const Cell lo = cell_of(G, (float)(x0 - s.hx), (float)(y0 - s.hy), (float)(z0 - s.hz));
    const Cell hi = cell_of(G, (float)(x0 + s.hx), (float)(y0 + s.hy), (float)(z0 + s.hz));
    const int bx = hi.cx - lo.cx + 2, by = hi.cy - lo.cy + 2, bz = hi.cz - lo.cz + 2;
    const bool in_lds = bx * by * bz <= LDS_NODES;
    __shared__ double acc[3 * LDS_NODES];
    if (in_lds)
        for (int i = tid; i < 3 * bx * by * bz; i += WBLOCK) acc[i] = 0.0;
    __syncthreads();
    const double mf = st[1], bf = st[2], mg = st[3], bg = st[4], z = st[5];
    const float cf = (float)(-2.0 * gC / bg);  // dL/dg_v = cf (f~ - ZNCC g~)
    const float imf = (float)mf, ibf = (float)(1.0 / bf), img = (float)mg, ibg = (float)(1.0 / bg), zf = (float)z;
    const Domain dom = make_domain(V, (int)x0, (int)y0, (int)z0);
    const int64_t obase = lin(V, z0, y0, x0);
    const float* gbase = V.g + obase;
    float* ggbase = b.grad_g ? b.grad_g + obase : nullptr;
    const int rowp = V.nx, planep = V.nx * V.ny;
    for (int v = tid; v < s.S; v += WBLOCK) {
        const int iz = v / (s.sx * s.sy), rem = v - iz * s.sx * s.sy, iy = rem / s.sx, ix = rem - iy * s.sx;
        const int dx = -s.hx + s.stride * ix, dy = -s.hy + s.stride * iy, dz = -s.hz + s.stride * iz;
        const float fv = V.f[obase + (int64_t)dz * planep + (int64_t)dy * rowp + dx];
        const Cell c = cell_of(G, (float)(x0 + dx), (float)(y0 + dy), (float)(z0 + dz));
        float u[3];
        fe_displacement(G, c, u);
        const float rx = (float)dx + u[0], ry = (float)dy + u[1], rz = (float)dz + u[2];
        float gr[3];
        const float gv = sample_grad<MODE, true, false>(gbase, rowp, planep, rx, ry, rz, dom.kx, dom.ky, dom.kz, gr,
                                                         nullptr, 0.0f);
        if (gv != gv) continue;  // excluded from the statistics in the forward pass
        const float dLdg = cf * ((fv - imf) * ibf - zf * (gv - img) * ibg);
        if (ggbase)
            sample_grad<MODE, false, true>(nullptr, rowp, planep, rx, ry, rz, dom.kx, dom.ky, dom.kz, nullptr, ggbase,
                                           dLdg);
        const float d[3] = {dLdg * gr[0], dLdg * gr[1], dLdg * gr[2]};
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const float N = shape(c, i, j, k);
                    if (in_lds) {
                        const int l = 3 * (((c.cz + k - lo.cz) * by + (c.cy + j - lo.cy)) * bx + (c.cx + i - lo.cx));
                        atomicAdd(&acc[l], (double)(N * d[0]));
                        atomicAdd(&acc[l + 1], (double)(N * d[1]));
                        atomicAdd(&acc[l + 2], (double)(N * d[2]));
                    } else {
                        double* gu = b.grad_u + 3 * (((int64_t)(c.cz + k) * G.gny + (c.cy + j)) * G.gnx + (c.cx + i));
                        unsafeAtomicAdd(gu, (double)(N * d[0]));
                        unsafeAtomicAdd(gu + 1, (double)(N * d[1]));
                        unsafeAtomicAdd(gu + 2, (double)(N * d[2]));
                    }
                }
    }
    __syncthreads();
    if (in_lds) {
        for (int i = tid; i < bx * by * bz; i += WBLOCK) {
            const int kx = i % bx, ky = (i / bx) % by, kz = i / (bx * by);
            const int nx_ = lo.cx + kx, ny_ = lo.cy + ky, nz_ = lo.cz + kz;
            if (nx_ >= G.gnx || ny_ >= G.gny || nz_ >= G.gnz) continue;
            double* gu = b.grad_u + 3 * (((int64_t)nz_ * G.gny + ny_) * G.gnx + nx_);
            for (int c3 = 0; c3 < 3; ++c3)
                if (acc[3 * i + c3] != 0.0) unsafeAtomicAdd(gu + c3, acc[3 * i + c3]);
        }
    }
}

int check_warp(const float* f, const float* g, int64_t nz, int64_t ny, int64_t nx, int mode, const int64_t* centers,
               int hx, int hy, int hz, const float* node_u, int gnx, int gny, int gnz, double sx, double sy, double sz) {
    if (!f || !g || !centers || !node_u) return fail(-2, "null pointer");
    if (nz < 4 || ny < 4 || nx < 4 || nz * ny * nx > INT32_MAX) return fail(-2, "bad volume shape");
    if (mode < 0 || mode > 2) return fail(-2, "interp mode must be 0 (linear), 1 (cubic) or 2 (bspline)");
    if (hx < 0 || hy < 0 || hz < 0) return fail(-2, "bad subset half-widths");
    if (gnx < 2 || gny < 2 || gnz < 2) return fail(-2, "the node grid needs at least 2 nodes per axis");
    if (!(sx > 0) || !(sy > 0) || !(sz > 0)) return fail(-2, "node spacing must be positive");
    return 0;
}

WarpArgs make_args(const float* f, const float* g, int64_t nz, int64_t ny, int64_t nx, const int64_t* centers, int64_t S,
                   int hx, int hy, int hz, const float* node_u, int gnx, int gny, int gnz, double ox, double oy,
                   double oz, double sx, double sy, double sz, double* C, double* stats) {
    WarpArgs a;
    a.V = Volumes{f, nullptr, nullptr, nullptr, g, nullptr, (int)nz, (int)ny, (int)nx};
    a.s = make_subset(hx, hy, hz, 1);
    a.G = NodeGrid{node_u, gnx, gny, gnz, (float)ox, (float)oy, (float)oz, (float)(1.0 / sx), (float)(1.0 / sy),
                   (float)(1.0 / sz)};
    a.centers = centers;
    a.S = S;
    a.C = C;
    a.stats = stats;
    return a;
}

}  // namespace

extern "C" {

int aldvc_warp_znssd(const float* f, const float* g, int64_t nz, int64_t ny, int64_t nx, int mode,
                     const int64_t* centers, int64_t S, int hx, int hy, int hz, const float* node_u, int gnx, int gny,
                     int gnz, double ox, double oy, double oz, double sx, double sy, double sz, double* C,
                     double* stats, hipStream_t stream) {
    clear_error();
    if (int rc = check_warp(f, g, nz, ny, nx, mode, centers, hx, hy, hz, node_u, gnx, gny, gnz, sx, sy, sz)) return rc;
    if (!C || !stats) return fail(-2, "null pointer");
    if (S <= 0) return 0;
    if (S > INT32_MAX) return fail(-2, "too many subsets for one launch");
    const WarpArgs a = make_args(f, g, nz, ny, nx, centers, S, hx, hy, hz, node_u, gnx, gny, gnz, ox, oy, oz, sx, sy, sz,
                                 C, stats);
    switch (mode) {
        case 0: hipLaunchKernelGGL(k_warp_znssd<0>, dim3((unsigned)S), dim3(WBLOCK), 0, stream, a); break;
        case 1: hipLaunchKernelGGL(k_warp_znssd<1>, dim3((unsigned)S), dim3(WBLOCK), 0, stream, a); break;
        default: hipLaunchKernelGGL(k_warp_znssd<2>, dim3((unsigned)S), dim3(WBLOCK), 0, stream, a); break;
    }
    return launch_status();
}

int aldvc_warp_znssd_backward(const float* f, const float* g, int64_t nz, int64_t ny, int64_t nx, int mode,
                              const int64_t* centers, int64_t S, int hx, int hy, int hz, const float* node_u, int gnx,
                              int gny, int gnz, double ox, double oy, double oz, double sx, double sy, double sz,
                              const double* stats, const double* grad_C, double* grad_u, float* grad_g,
                              hipStream_t stream) {
    clear_error();
    if (int rc = check_warp(f, g, nz, ny, nx, mode, centers, hx, hy, hz, node_u, gnx, gny, gnz, sx, sy, sz)) return rc;
    if (!stats || !grad_C || !grad_u) return fail(-2, "null pointer");
    if (S <= 0) return 0;
    if (S > INT32_MAX) return fail(-2, "too many subsets for one launch");
    WarpBackArgs b;
    b.f = make_args(f, g, nz, ny, nx, centers, S, hx, hy, hz, node_u, gnx, gny, gnz, ox, oy, oz, sx, sy, sz, nullptr,
                    const_cast<double*>(stats));
    b.grad_C = grad_C;
    b.grad_u = grad_u;
    b.grad_g = grad_g;
    switch (mode) {
        case 0: hipLaunchKernelGGL(k_warp_znssd_backward<0>, dim3((unsigned)S), dim3(WBLOCK), 0, stream, b); break;
        case 1: hipLaunchKernelGGL(k_warp_znssd_backward<1>, dim3((unsigned)S), dim3(WBLOCK), 0, stream, b); break;
        default: hipLaunchKernelGGL(k_warp_znssd_backward<2>, dim3((unsigned)S), dim3(WBLOCK), 0, stream, b); break;
    }
    return launch_status();
}

}  // extern "C"
