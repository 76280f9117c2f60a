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
// Backward (the reference has none): dC_n/dg_v = -2 dL/dC_n (f~_v - ZNCC g~_v) / b_g per sampled voxel.
// All subsets sample g through the same warp phi(v) = v + u(v), so the per-subset terms fold into one
// field over the reference voxels, D(v); dL/du and dL/dg follow from D by fixed-order sums (no
// float atomics: two runs give bit-identical gradients). Interpolation weights and their derivatives
// are recomputed from the node field; see the comment above k_warp_bins.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>

#include "aldvc_abi_util.h"
#include "aldvc_device.h"

using namespace aldvc;

namespace {

constexpr int WBLOCK = 256;
constexpr int WWAVES = WBLOCK / 64;
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

// Range of u over the voxels of the box [vlo, vhi] (absolute, per axis) that cell (cx, cy, cz) owns
// (cell_of: edge cells also own the voxels beyond the grid), merged into lo/hi. u is multilinear in
// the cell's local coordinates, so its extremes over a box of them lie at the box's corners.
__device__ inline void cell_u_bounds(const NodeGrid& G, int cx, int cy, int cz, const float vlo[3], const float vhi[3],
                                     float lo[3], float hi[3]) {
    const int c[3] = {cx, cy, cz}, last[3] = {G.gnx - 2, G.gny - 2, G.gnz - 2};
    const float org[3] = {G.ox, G.oy, G.oz}, isp[3] = {G.isx, G.isy, G.isz};
    float t0[3], t1[3];
    for (int k = 0; k < 3; ++k) {
        float a = (vlo[k] - org[k]) * isp[k] - (float)c[k], b = (vhi[k] - org[k]) * isp[k] - (float)c[k];
        if (c[k] > 0) a = fmaxf(a, 0.0f);
        if (c[k] < last[k]) b = fminf(b, 1.0f);
        if (a > b) return;
        t0[k] = a;
        t1[k] = b;
    }
    for (int corner = 0; corner < 8; ++corner) {
        Cell cc;
        cc.cx = cx;
        cc.cy = cy;
        cc.cz = cz;
        cc.tx = (corner & 1) ? t1[0] : t0[0];
        cc.ty = (corner & 2) ? t1[1] : t0[1];
        cc.tz = (corner & 4) ? t1[2] : t0[2];
        float u[3];
        fe_displacement(G, cc, u);
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], u[k]);
            hi[k] = fmaxf(hi[k], u[k]);
        }
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

// Sample g at the anchor-relative position (rx, ry, rz) and, with GRAD, its spatial gradient.
// The caller has checked the sampling domain; the stencil base is clamped to n - 3 as the sampler's.
template <int MODE, bool GRAD>
__device__ inline float sample_grad(const float* __restrict__ base, int rowp, int planep, float rx, float ry, float rz,
                                    int kx, int ky, int kz, float grad[3]) {
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
                const float s = base[r + i];
                rv += wx[i] * s;
                if constexpr (GRAD) rd += dx[i] * s;
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
        const float gv = sample_grad<MODE, false>(gbase, rowp, planep, rx, ry, rz, dom.kx, dom.ky, dom.kz, nullptr);
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

// ---------------------------------------------------------------------------- backward
// Every subset samples g through the same warp phi(v) = v + u(v), so the per-subset gradients fold
// into one field over the reference voxels: D(v) = sum over the valid subsets n holding v of
// dC_n/dg at v = alpha_n f_v - beta_n g(phi(v)) + gamma_n, with alpha_n = c_n / b_f, beta_n = c_n z_n / b_g,
// gamma_n = c_n (z_n m_g / b_g - m_f / b_f) and c_n = -2 dL/dC_n / b_g. Then
//   dL/du_a = sum_v N_a(v) D(v) grad g(phi(v))                  (per-cell sums, then per node)
//   dL/dg[q] = sum_v D(v) w(q - phi(v))                           (a gather per output voxel)
// Every sum runs in a fixed order: no float atomics, bit-identical results run to run.

constexpr int BIN_BLOCK = 1024;
constexpr int FIELD_BLOCK = 256;
constexpr int CELL_BLOCK = 256;
constexpr int CELL_SUMS = 24;    // 8 corners x 3 components
constexpr int SPLAT_TILE = 8;    // output tile edge of the gather
constexpr int SPLAT_BLOCK = 256;
constexpr int OFF_BITS = 10;     // packed per-axis stencil offsets floor(u): signed 10-bit fields

struct WarpBackArgs {
    WarpArgs f;
    const double* grad_C;  // (S)
    double* grad_u;        // (gnz, gny, gnx, 3), accumulated
    float* grad_g;         // (nz, ny, nx) or null, accumulated
    // workspace
    int* bin_start;    // (nbins + 1)
    int* bin_ids;      // (S) subset ids grouped by bin, ascending inside a bin
    int* bin_fill;     // (nbins) fill cursors
    double* coef;      // (S, 3) alpha, beta, gamma of the contributing subsets
    float4* field;     // (voxels) frac of phi (x, y, z), D
    int* offs;         // (voxels) packed floor(u) per axis, or INT_MIN where D is not defined
    float4* dgrad;     // (voxels) D grad g(phi), 0
    double* cell_part; // (cells, 24)
    float* ubound;     // (6): min (x, y, z), max (x, y, z) of u over the volume
    int nbx, nby, nbz;  // bins of (2h + 1) voxels per axis over the volume
};

__device__ inline int pack_offs(int ox, int oy, int oz) {
    const int m = (1 << OFF_BITS) - 1;
    return (ox & m) | ((oy & m) << OFF_BITS) | ((oz & m) << (2 * OFF_BITS));
}

__device__ inline int unpack_off(int p, int k) {
    return (p << (32 - OFF_BITS * (k + 1))) >> (32 - OFF_BITS);  // sign-extended field k
}

// One workgroup: the per-subset coefficients and the bin index of the subsets that contribute
// (valid, finite nonzero upstream gradient). Counts are integer atomics; ids are sorted in each
// bin afterwards, so the index does not depend on the order of the atomics.
__global__ __launch_bounds__(BIN_BLOCK) void k_warp_bins(WarpBackArgs b) {
    const WarpArgs& a = b.f;
    const int nbins = b.nbx * b.nby * b.nbz;
    const int tid = threadIdx.x;
    for (int i = tid; i < nbins; i += BIN_BLOCK) b.bin_fill[i] = 0;
    __syncthreads();
    auto bin_of = [&](int64_t n) {
        const int64_t* c = a.centers + 3 * n;
        const int bx = (int)min(max(c[0] / (2 * a.s.hx + 1), (int64_t)0), (int64_t)b.nbx - 1);
        const int by = (int)min(max(c[1] / (2 * a.s.hy + 1), (int64_t)0), (int64_t)b.nby - 1);
        const int bz = (int)min(max(c[2] / (2 * a.s.hz + 1), (int64_t)0), (int64_t)b.nbz - 1);
        return (bz * b.nby + by) * b.nbx + bx;
    };
    for (int64_t n = tid; n < a.S; n += BIN_BLOCK) {
        const double* st = a.stats + NSTAT * n;
        const double gC = b.grad_C[n];
        const bool use = st[6] != 0.0 && gC != 0.0 && isfinite(gC);
        double* cf = b.coef + 3 * n;
        if (use) {
            const double c = -2.0 * gC / st[4];
            cf[0] = c / st[2];
            cf[1] = c * st[5] / st[4];
            cf[2] = c * (st[5] * st[3] / st[4] - st[1] / st[2]);
            atomicAdd(&b.bin_fill[bin_of(n)], 1);
        } else {
            cf[0] = cf[1] = cf[2] = 0.0;
        }
    }
    __syncthreads();
    if (tid == 0) {  // exclusive scan (bins are few: one per (2h+1)^3 voxels)
        int run = 0;
        for (int i = 0; i < nbins; ++i) {
            b.bin_start[i] = run;
            run += b.bin_fill[i];
            b.bin_fill[i] = b.bin_start[i];
        }
        b.bin_start[nbins] = run;
    }
    __syncthreads();
    for (int64_t n = tid; n < a.S; n += BIN_BLOCK) {
        const double* cf = b.coef + 3 * n;
        if (cf[0] != 0.0 || cf[1] != 0.0 || cf[2] != 0.0) b.bin_ids[atomicAdd(&b.bin_fill[bin_of(n)], 1)] = (int)n;
    }
    __syncthreads();
    for (int i = tid; i < nbins; i += BIN_BLOCK) {  // insertion sort of each bin's ids
        int* ids = b.bin_ids + b.bin_start[i];
        const int m = b.bin_start[i + 1] - b.bin_start[i];
        for (int p = 1; p < m; ++p) {
            const int key = ids[p];
            int q = p - 1;
            while (q >= 0 && ids[q] > key) {
                ids[q + 1] = ids[q];
                --q;
            }
            ids[q + 1] = key;
        }
    }
    // range of u over the whole volume (min / max reductions: the order does not matter)
    const NodeGrid& G = a.G;
    const float vlo[3] = {0.0f, 0.0f, 0.0f};
    const float vhi[3] = {(float)(a.V.nx - 1), (float)(a.V.ny - 1), (float)(a.V.nz - 1)};
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const int64_t ncx = G.gnx - 1, ncy = G.gny - 1, cells = ncx * ncy * (G.gnz - 1);
    for (int64_t c = tid; c < cells; c += BIN_BLOCK)
        cell_u_bounds(G, (int)(c % ncx), (int)((c / ncx) % ncy), (int)(c / (ncx * ncy)), vlo, vhi, lo, hi);
    __shared__ float red[6][BIN_BLOCK / 64];
    const int lane = tid & 63, wave = tid >> 6;
    for (int k = 0; k < 3; ++k) {
        float l = lo[k], h = hi[k];
        for (int off = 32; off > 0; off >>= 1) {
            l = fminf(l, __shfl_xor(l, off, 64));
            h = fmaxf(h, __shfl_xor(h, off, 64));
        }
        if (lane == 0) {
            red[k][wave] = l;
            red[3 + k][wave] = h;
        }
    }
    __syncthreads();
    if (tid < 6) {
        float v = red[tid][0];
        for (int w = 1; w < BIN_BLOCK / 64; ++w) v = tid < 3 ? fminf(v, red[tid][w]) : fmaxf(v, red[tid][w]);
        b.ubound[tid] = v;
    }
}

// Per reference voxel: D(v) from the subsets holding it, the warped position's stencil offsets and
// fractions, g and grad g there. One workgroup per row segment (fixed y, z; FIELD_BLOCK voxels along
// x): the subsets whose y and z ranges hold the row are gathered once into LDS, in bin order and
// ascending id inside a bin (an order-keeping block compaction), and every voxel then sums those
// whose x range holds it, in that order.
constexpr int ROW_LIST = 1024;  // subsets of one row held in LDS

template <int MODE>
__global__ __launch_bounds__(FIELD_BLOCK) void k_warp_field(WarpBackArgs b) {
    const WarpArgs& a = b.f;
    const Volumes& V = a.V;
    const Subset& s = a.s;
    const int segs = (V.nx + FIELD_BLOCK - 1) / FIELD_BLOCK;
    const int seg = blockIdx.x % segs;
    const int y = (blockIdx.x / segs) % V.ny, z = blockIdx.x / (segs * V.ny);
    const int xs0 = seg * FIELD_BLOCK, xs1 = min(xs0 + FIELD_BLOCK, V.nx) - 1;
    const int Bx = 2 * s.hx + 1, By = 2 * s.hy + 1, Bz = 2 * s.hz + 1;
    const int bx0 = max((xs0 - s.hx) / Bx, 0), bx1 = min((xs1 + s.hx) / Bx, b.nbx - 1);
    const int by0 = max((y - s.hy) / By, 0), by1 = min((y + s.hy) / By, b.nby - 1);
    const int bz0 = max((z - s.hz) / Bz, 0), bz1 = min((z + s.hz) / Bz, b.nbz - 1);
    __shared__ int lcx[ROW_LIST];
    __shared__ double lcoef[3][ROW_LIST];
    __shared__ int wcount[FIELD_BLOCK / 64];
    __shared__ int total;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x = xs0 + tid;
    double al = 0.0, be = 0.0, ga = 0.0;
    bool any = false;
    auto consume = [&]() {  // the listed subsets whose x range holds this voxel, in list order
        for (int k = 0; k < total; ++k) {
            if (x < lcx[k] - s.hx || x > lcx[k] + s.hx) continue;
            al += lcoef[0][k];
            be += lcoef[1][k];
            ga += lcoef[2][k];
            any = true;
        }
    };
    if (tid == 0) total = 0;
    __syncthreads();
    // candidates: the bins of (bz, by, bx) in that order, each bin's ids ascending; the list is
    // consumed whenever the next round of candidates might not fit it
    for (int bz = bz0; bz <= bz1; ++bz)
        for (int by = by0; by <= by1; ++by) {
            const int first = b.bin_start[(bz * b.nby + by) * b.nbx + bx0];
            const int last = b.bin_start[(bz * b.nby + by) * b.nbx + bx1 + 1];
            for (int k0 = first; k0 < last; k0 += FIELD_BLOCK) {
                if (total + FIELD_BLOCK > ROW_LIST) {
                    consume();
                    __syncthreads();
                    if (tid == 0) total = 0;
                    __syncthreads();
                }
                const int k = k0 + tid;
                int n = -1;
                if (k < last) {
                    n = b.bin_ids[k];
                    const int64_t* c = a.centers + 3 * n;
                    if (y < c[1] - s.hy || y > c[1] + s.hy || z < c[2] - s.hz || z > c[2] + s.hz || c[0] + s.hx < xs0 ||
                        c[0] - s.hx > xs1)
                        n = -1;
                }
                const unsigned long long keep = __ballot(n >= 0);
                const int before = __popcll(keep & ((1ull << lane) - 1ull));
                if (lane == 0) wcount[wave] = __popcll(keep);
                __syncthreads();
                int base = total;
                for (int w = 0; w < wave; ++w) base += wcount[w];
                if (n >= 0) {
                    lcx[base + before] = (int)a.centers[3 * n];
                    for (int q = 0; q < 3; ++q) lcoef[q][base + before] = b.coef[3 * n + q];
                }
                __syncthreads();
                if (tid == 0)
                    for (int w = 0; w < FIELD_BLOCK / 64; ++w) total += wcount[w];
                __syncthreads();
            }
        }
    consume();
    if (x > xs1) return;
    const int64_t i = ((int64_t)z * V.ny + y) * V.nx + x;
    b.offs[i] = INT_MIN;
    b.field[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    b.dgrad[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!any) return;
    const Cell c = cell_of(a.G, (float)x, (float)y, (float)z);
    float u[3];
    fe_displacement(a.G, c, u);
    const Domain dom = make_domain(V, x, y, z);  // positions relative to the voxel itself
    if (!inside(dom, u[0], u[1], u[2])) return;
    float gr[3];
    const float gv = sample_grad<MODE, true>(V.g + i, V.nx, V.nx * V.ny, u[0], u[1], u[2], dom.kx, dom.ky, dom.kz, gr);
    if (gv != gv) return;  // excluded from the statistics in the forward pass
    const float D = (float)(al * (double)V.f[i] - be * (double)gv + ga);
    int ix = (int)floorf(u[0]), iy = (int)floorf(u[1]), iz = (int)floorf(u[2]);
    ix = ix > dom.kx ? dom.kx : ix;
    iy = iy > dom.ky ? dom.ky : iy;
    iz = iz > dom.kz ? dom.kz : iz;
    b.offs[i] = pack_offs(ix, iy, iz);
    b.field[i] = make_float4(u[0] - (float)ix, u[1] - (float)iy, u[2] - (float)iz, D);
    b.dgrad[i] = make_float4(D * gr[0], D * gr[1], D * gr[2], 0.f);
}

// One workgroup per node-grid cell: the 24 sums N_a(v) D(v) grad g over the voxels the cell owns
// (cell_of, which also gives the grid's edge cells the voxels beyond the grid).
__global__ __launch_bounds__(CELL_BLOCK) void k_warp_cells(WarpBackArgs b) {
    const WarpArgs& a = b.f;
    const Volumes& V = a.V;
    const NodeGrid& G = a.G;
    const int ncx = G.gnx - 1, ncy = G.gny - 1;
    const int cell = blockIdx.x;
    const int cx = cell % ncx, cy = (cell / ncx) % ncy, cz = cell / (ncx * ncy);
    auto range = [](int c, int ncell, float o, float sp, int n, int& lo, int& hi) {
        lo = c == 0 ? 0 : max((int)floorf(o + sp * (float)c) - 1, 0);
        hi = c == ncell - 1 ? n - 1 : min((int)ceilf(o + sp * (float)(c + 1)) + 1, n - 1);
    };
    int x0, x1, y0, y1, z0, z1;
    range(cx, ncx, G.ox, 1.0f / G.isx, V.nx, x0, x1);
    range(cy, ncy, G.oy, 1.0f / G.isy, V.ny, y0, y1);
    range(cz, G.gnz - 1, G.oz, 1.0f / G.isz, V.nz, z0, z1);
    const int sx = x1 - x0 + 1, sy = y1 - y0 + 1;
    const int64_t count = (x1 >= x0 && y1 >= y0 && z1 >= z0) ? (int64_t)sx * sy * (z1 - z0 + 1) : 0;
    double acc[CELL_SUMS];
#pragma unroll
    for (int q = 0; q < CELL_SUMS; ++q) acc[q] = 0.0;
    for (int64_t k = threadIdx.x; k < count; k += CELL_BLOCK) {
        const int x = x0 + (int)(k % sx), y = y0 + (int)((k / sx) % sy), z = z0 + (int)(k / ((int64_t)sx * sy));
        const Cell c = cell_of(G, (float)x, (float)y, (float)z);
        if (c.cx != cx || c.cy != cy || c.cz != cz) continue;
        const float4 d = b.dgrad[((int64_t)z * V.ny + y) * V.nx + x];
        if (d.x == 0.f && d.y == 0.f && d.z == 0.f) continue;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int ii = 0; ii < 2; ++ii) {
                    const double N = (double)shape(c, ii, j, kk);
                    const int q = 3 * ((kk * 2 + j) * 2 + ii);
                    acc[q] += N * (double)d.x;
                    acc[q + 1] += N * (double)d.y;
                    acc[q + 2] += N * (double)d.z;
                }
    }
    __shared__ double red[CELL_BLOCK / 64][CELL_SUMS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < CELL_SUMS; ++q) {
        const double t = wave_sum(acc[q]);
        if (lane == 0) red[wave][q] = t;
    }
    __syncthreads();
    if (threadIdx.x < CELL_SUMS) {
        double t = 0.0;
        for (int w = 0; w < CELL_BLOCK / 64; ++w) t += red[w][threadIdx.x];
        b.cell_part[(int64_t)cell * CELL_SUMS + threadIdx.x] = t;
    }
}

// Per node: its corner sums from the (up to 8) cells around it, in (z, y, x) cell order.
__global__ void k_warp_nodes(WarpBackArgs b) {
    const NodeGrid& G = b.f.G;
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nn = (int64_t)G.gnx * G.gny * G.gnz;
    if (a >= nn) return;
    const int ax = (int)(a % G.gnx), ay = (int)((a / G.gnx) % G.gny), az = (int)(a / ((int64_t)G.gnx * G.gny));
    const int ncx = G.gnx - 1, ncy = G.gny - 1, ncz = G.gnz - 1;
    double t[3] = {0.0, 0.0, 0.0};
    for (int kk = 1; kk >= 0; --kk)
        for (int j = 1; j >= 0; --j)
            for (int ii = 1; ii >= 0; --ii) {  // the cell whose corner (ii, j, kk) is this node
                const int cx = ax - ii, cy = ay - j, cz = az - kk;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= ncx || cy >= ncy || cz >= ncz) continue;
                const double* p = b.cell_part + (((int64_t)cz * ncy + cy) * ncx + cx) * CELL_SUMS + 3 * ((kk * 2 + j) * 2 + ii);
                t[0] += p[0];
                t[1] += p[1];
                t[2] += p[2];
            }
    b.grad_u[3 * a] += t[0];
    b.grad_u[3 * a + 1] += t[1];
    b.grad_u[3 * a + 2] += t[2];
}

// The interpolation weight of tap k (0..T-1) at fraction t.
template <int MODE>
__device__ inline float tap_weight(float t, int k) {
    float w[4], d[4];
    weights_d<MODE>(t, w, d);
    return k == 0 ? w[0] : k == 1 ? w[1] : k == 2 ? w[2] : w[3];
}

// dL/dg as a gather: one thread per output voxel q of an 8^3 tile; the reference voxels whose
// stencil can reach q lie in q - u + [-(T - LO), LO + 1] per axis, with u bounded by the node
// values of the cells around the tile (trilinear values stay within their cell's corners).
template <int MODE>
__global__ __launch_bounds__(SPLAT_BLOCK) void k_warp_splat(WarpBackArgs b) {
    constexpr int T = MODE == INTERP_LINEAR ? 2 : 4, LO = MODE == INTERP_LINEAR ? 0 : 1;
    const WarpArgs& a = b.f;
    const Volumes& V = a.V;
    const NodeGrid& G = a.G;
    const int tnx = (V.nx + SPLAT_TILE - 1) / SPLAT_TILE, tny = (V.ny + SPLAT_TILE - 1) / SPLAT_TILE;
    const int tile = blockIdx.x;
    const int qx0 = (tile % tnx) * SPLAT_TILE, qy0 = ((tile / tnx) % tny) * SPLAT_TILE,
              qz0 = (tile / (tnx * tny)) * SPLAT_TILE;
    // v feeds q along an axis when q - v - floor(u(v)) + LO lies in [0, T): v in
    // [q + LO - T + 1 - ceil(umax), q + LO - floor(umin)]. The volume-wide range of u gives the
    // voxels that can feed the tile; the range over the cells owning those voxels narrows it.
    __shared__ int win[6];
    if (threadIdx.x == 0) {
        const float eps = 1e-3f;  // rounding between these bounds and the u of the field kernel
        const float q0[3] = {(float)qx0, (float)qy0, (float)qz0};
        const int n[3] = {V.nx, V.ny, V.nz};
        float vlo[3], vhi[3];
        for (int k = 0; k < 3; ++k) {
            vlo[k] = fmaxf(q0[k] + (float)(LO - T + 1) - ceilf(b.ubound[3 + k] + eps), 0.0f);
            vhi[k] = fminf(q0[k] + (float)(SPLAT_TILE - 1 + LO) - floorf(b.ubound[k] - eps), (float)(n[k] - 1));
        }
        const Cell c0 = cell_of(G, vlo[0], vlo[1], vlo[2]), c1 = cell_of(G, vhi[0], vhi[1], vhi[2]);
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int cz = c0.cz; cz <= c1.cz; ++cz)
            for (int cy = c0.cy; cy <= c1.cy; ++cy)
                for (int cx = c0.cx; cx <= c1.cx; ++cx) cell_u_bounds(G, cx, cy, cz, vlo, vhi, lo, hi);
        for (int k = 0; k < 3; ++k) {
            if (!(lo[k] <= hi[k])) {  // no voxel can feed the tile
                lo[k] = 0.0f;
                hi[k] = -(float)(n[k] + 8);
            }
            win[k] = LO - T + 1 - (int)ceilf(hi[k] + eps);
            win[3 + k] = LO - (int)floorf(lo[k] - eps);
        }
    }
    __syncthreads();
    const int wx0 = win[0], wy0 = win[1], wz0 = win[2], wx1 = win[3], wy1 = win[4], wz1 = win[5];
    for (int k = threadIdx.x; k < SPLAT_TILE * SPLAT_TILE * SPLAT_TILE; k += SPLAT_BLOCK) {
        const int qx = qx0 + (k % SPLAT_TILE), qy = qy0 + ((k / SPLAT_TILE) % SPLAT_TILE),
                  qz = qz0 + k / (SPLAT_TILE * SPLAT_TILE);
        if (qx >= V.nx || qy >= V.ny || qz >= V.nz) continue;
        float sum = 0.0f;
        for (int vz = max(qz + wz0, 0); vz <= min(qz + wz1, V.nz - 1); ++vz)
            for (int vy = max(qy + wy0, 0); vy <= min(qy + wy1, V.ny - 1); ++vy)
                for (int vx = max(qx + wx0, 0); vx <= min(qx + wx1, V.nx - 1); ++vx) {
                    const int64_t vi = ((int64_t)vz * V.ny + vy) * V.nx + vx;
                    const int p = b.offs[vi];
                    if (p == INT_MIN) continue;
                    const int tx = qx - (vx + unpack_off(p, 0) - LO);
                    const int ty = qy - (vy + unpack_off(p, 1) - LO);
                    const int tz = qz - (vz + unpack_off(p, 2) - LO);
                    if ((unsigned)tx >= (unsigned)T || (unsigned)ty >= (unsigned)T || (unsigned)tz >= (unsigned)T) continue;
                    const float4 fd = b.field[vi];
                    sum += fd.w * tap_weight<MODE>(fd.x, tx) * tap_weight<MODE>(fd.y, ty) * tap_weight<MODE>(fd.z, tz);
                }
        if (sum != 0.0f) b.grad_g[((int64_t)qz * V.ny + qy) * V.nx + qx] += sum;
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

size_t aldvc_warp_backward_workspace_bytes(int64_t nz, int64_t ny, int64_t nx, int64_t S, int hx, int hy, int hz,
                                           int gnx, int gny, int gnz) {
    const int64_t nbins = ((nx + 2 * hx) / (2 * hx + 1)) * ((ny + 2 * hy) / (2 * hy + 1)) * ((nz + 2 * hz) / (2 * hz + 1));
    const int64_t nvox = nx * ny * nz, cells = (int64_t)(gnx - 1) * (gny - 1) * (gnz - 1);
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    return up(sizeof(int) * (size_t)(nbins + 1)) + up(sizeof(int) * (size_t)(S > 0 ? S : 1)) +
           up(sizeof(int) * (size_t)nbins) + up(sizeof(double) * 3 * (size_t)(S > 0 ? S : 1)) +
           up(sizeof(float4) * (size_t)nvox) + up(sizeof(int) * (size_t)nvox) + up(sizeof(float4) * (size_t)nvox) +
           up(sizeof(double) * CELL_SUMS * (size_t)(cells > 0 ? cells : 1)) + up(sizeof(float) * 8);
}

int aldvc_warp_znssd_backward(const float* f, const float* g, int64_t nz, int64_t ny, int64_t nx, int mode,
                              const int64_t* centers, int64_t S, int hx, int hy, int hz, const float* node_u, int gnx,
                              int gny, int gnz, double ox, double oy, double oz, double sx, double sy, double sz,
                              const double* stats, const double* grad_C, double* grad_u, float* grad_g,
                              void* workspace, size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    if (int rc = check_warp(f, g, nz, ny, nx, mode, centers, hx, hy, hz, node_u, gnx, gny, gnz, sx, sy, sz)) return rc;
    if (!stats || !grad_C || !grad_u) return fail(-2, "null pointer");
    if (S <= 0) return 0;
    if (S > INT32_MAX) return fail(-2, "too many subsets for one launch");
    if (!workspace || workspace_bytes < aldvc_warp_backward_workspace_bytes(nz, ny, nx, S, hx, hy, hz, gnx, gny, gnz))
        return fail(-3, "workspace too small (see aldvc_warp_backward_workspace_bytes)");
    WarpBackArgs b;
    b.f = make_args(f, g, nz, ny, nx, centers, S, hx, hy, hz, node_u, gnx, gny, gnz, ox, oy, oz, sx, sy, sz, nullptr,
                    const_cast<double*>(stats));
    b.grad_C = grad_C;
    b.grad_u = grad_u;
    b.grad_g = grad_g;
    b.nbx = (int)((nx + 2 * hx) / (2 * hx + 1));
    b.nby = (int)((ny + 2 * hy) / (2 * hy + 1));
    b.nbz = (int)((nz + 2 * hz) / (2 * hz + 1));
    const int64_t nbins = (int64_t)b.nbx * b.nby * b.nbz, nvox = nx * ny * nz;
    const int64_t cells = (int64_t)(gnx - 1) * (gny - 1) * (gnz - 1);
    char* w = static_cast<char*>(workspace);
    auto take = [&](size_t bytes) {
        char* p = w;
        w += (bytes + 255) / 256 * 256;
        return p;
    };
    b.bin_start = (int*)take(sizeof(int) * (size_t)(nbins + 1));
    b.bin_ids = (int*)take(sizeof(int) * (size_t)S);
    b.bin_fill = (int*)take(sizeof(int) * (size_t)nbins);
    b.coef = (double*)take(sizeof(double) * 3 * (size_t)S);
    b.field = (float4*)take(sizeof(float4) * (size_t)nvox);
    b.offs = (int*)take(sizeof(int) * (size_t)nvox);
    b.dgrad = (float4*)take(sizeof(float4) * (size_t)nvox);
    b.cell_part = (double*)take(sizeof(double) * CELL_SUMS * (size_t)cells);
    b.ubound = (float*)take(sizeof(float) * 8);
    if (2 * hx + 1 > 511 || 2 * hy + 1 > 511 || 2 * hz + 1 > 511) return fail(-2, "subsets beyond 511 voxels per axis");
    const unsigned vblocks = (unsigned)(((nx + FIELD_BLOCK - 1) / FIELD_BLOCK) * ny * nz);
    const unsigned tiles = (unsigned)(((nx + SPLAT_TILE - 1) / SPLAT_TILE) * ((ny + SPLAT_TILE - 1) / SPLAT_TILE) *
                                      ((nz + SPLAT_TILE - 1) / SPLAT_TILE));
    hipLaunchKernelGGL(k_warp_bins, dim3(1), dim3(BIN_BLOCK), 0, stream, b);
#define ALDVC_WARP_BACK(MD)                                                                                       \
    hipLaunchKernelGGL(k_warp_field<MD>, dim3(vblocks), dim3(FIELD_BLOCK), 0, stream, b);                         \
    hipLaunchKernelGGL(k_warp_cells, dim3((unsigned)cells), dim3(CELL_BLOCK), 0, stream, b);                      \
    hipLaunchKernelGGL(k_warp_nodes, dim3((unsigned)((gnx * gny * gnz + 255) / 256)), dim3(256), 0, stream, b);   \
    if (grad_g) hipLaunchKernelGGL(k_warp_splat<MD>, dim3(tiles), dim3(SPLAT_BLOCK), 0, stream, b);
    switch (mode) {
        case 0: ALDVC_WARP_BACK(0) break;
        case 1: ALDVC_WARP_BACK(1) break;
        default: ALDVC_WARP_BACK(2) break;
    }
#undef ALDVC_WARP_BACK
    return launch_status();
}

}  // extern "C"
