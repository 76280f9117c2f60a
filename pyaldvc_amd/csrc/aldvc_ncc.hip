// Integer NCC search on gfx950: per-node ZNCC maps over the search window and the fused peak
// epilogue (first-index argmax, PCE, 27-point quadratic sub-voxel fit).
//
// Map semantics (solver/integer_search.py:44-151, cuda_kernels.py:1296-1389): the template is
// the (2h+1)^3 reference box around the node, centred by its mean; for every offset o of the
// (2r+1)^3 search grid, ZNCC(o) = sum T~ g / sqrt(var_w * sum T~^2) with
// var_w = sum g^2 - (sum g)^2 / n over the window at o; var_w < 1e-20 gives 0, values are
// clipped to [-1, 1], and a flat template (sum T~^2 < 1e-20) or t_ok == 0 gives a map of -2.
#include "aldvc_device.h"

#include <algorithm>

#include "aldvc_abi_util.h"

using namespace aldvc;

namespace {

constexpr int BLOCK = 256;
constexpr int NWAVES = BLOCK / 64;
constexpr int RX = 8;  // consecutive x offsets per lane (register blocking of the window row)
constexpr int LDS_FLOATS_MAX = 160 * 1024 / 4 - 1024;
constexpr size_t SMALL_LDS_MAX = 80 * 1024;  // two small-window workgroups per CU at the most

struct NccArgs {
    const float* f;
    const float* g;
    int nz, ny, nx;
    const int64_t* coords;
    int64_t n;
    int hx, hy, hz, rx, ry, rz;
    const int64_t* wx0;
    const int64_t* wy0;
    const int64_t* wz0;
    const uint8_t* t_ok;
    float* out;
};

template <int NT = BLOCK>
__device__ inline void fill_map(float* o, int64_t n_off, float v) {
    for (int64_t i = threadIdx.x; i < n_off; i += NT) o[i] = v;
}

// floor(n / d) for 0 <= n < 2^22 from inv = 1.0f / d: (n + 0.5) / d lies at least 0.5 / d from an
// integer, and the float rounding of the product (under (n + 0.5) 2^-23 / d) stays inside that
// margin, so the truncation is exact.
__device__ inline int qdiv(int n, float inv) { return (int)(((float)n + 0.5f) * inv); }

// The node's template into LDS, centred by its mean: float64 mean and energy (sum T~^2). Returns
// false for a flat template (energy below 1e-20). Block-uniform; every thread must call it.
template <int NT = BLOCK>
__device__ inline bool centred_template(const NccArgs& a, float* tpl, int64_t x0, int64_t y0, int64_t z0, int sx,
                                        int sy, int n_tpl, float& mf_out, double& ssf_out) {
    __shared__ double red[NT / 64];
    __shared__ double bc[2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double s = 0.0;
    const float inv_plane = 1.0f / (float)(sy * sx), inv_sx = 1.0f / (float)sx;
    for (int v = tid; v < n_tpl; v += NT) {
        const int k = qdiv(v, inv_plane), r = v - k * sy * sx, j = qdiv(r, inv_sx), i = r - j * sx;
        const float val = a.f[((z0 - a.hz + k) * a.ny + (y0 - a.hy + j)) * (int64_t)a.nx + (x0 - a.hx + i)];
        tpl[v] = val;
        s += val;
    }
    s = wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < NT / 64; ++w) t += red[w];
        bc[0] = t / n_tpl;
    }
    __syncthreads();
    const float mf = (float)bc[0];
    double ss = 0.0;
    for (int v = tid; v < n_tpl; v += NT) {
        const float val = tpl[v] - mf;
        tpl[v] = val;
        ss += (double)val * val;
    }
    ss = wave_sum(ss);
    __syncthreads();
    if (lane == 0) red[wave] = ss;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < NT / 64; ++w) t += red[w];
        bc[1] = t;
    }
    __syncthreads();
    mf_out = mf;
    ssf_out = bc[1];
    return !(ssf_out < 1e-20);
}

// Direct float32 correlation: one workgroup per node, the centred template in LDS, every lane
// owns RX consecutive x offsets of one (dz, dy) output row and slides along the window rows.
__global__ __launch_bounds__(BLOCK) void k_zncc_map_f32(NccArgs a) {
    extern __shared__ float tpl[];
    const int64_t node = blockIdx.x;
    const int tid = threadIdx.x;
    const int sx = 2 * a.hx + 1, sy = 2 * a.hy + 1, sz = 2 * a.hz + 1;
    const int vx = 2 * a.rx + 1, vy = 2 * a.ry + 1, vz = 2 * a.rz + 1;
    const int wx = sx + vx - 1;
    const int n_tpl = sx * sy * sz;
    const int64_t n_off = (int64_t)vx * vy * vz;
    float* o = a.out + node * n_off;
    const int64_t x0 = a.coords[3 * node], y0 = a.coords[3 * node + 1], z0 = a.coords[3 * node + 2];
    const int64_t xb = a.wx0[node], yb = a.wy0[node], zb = a.wz0[node];
    // a template or a search window that leaves the volume is treated like t_ok == 0
    const bool tpl_in = x0 - a.hx >= 0 && x0 + a.hx < a.nx && y0 - a.hy >= 0 && y0 + a.hy < a.ny && z0 - a.hz >= 0 &&
                        z0 + a.hz < a.nz;
    const bool win_in = xb >= 0 && xb + wx <= a.nx && yb >= 0 && yb + sy + vy - 1 <= a.ny && zb >= 0 &&
                        zb + sz + vz - 1 <= a.nz;
    if (!a.t_ok[node] || !tpl_in || !win_in) {
        fill_map(o, n_off, -2.0f);
        return;
    }
    double ssf;
    float mf;
    if (!centred_template(a, tpl, x0, y0, z0, sx, sy, n_tpl, mf, ssf)) {
        fill_map(o, n_off, -2.0f);
        return;
    }
    const float inv_n = 1.0f / (float)n_tpl;
    const float sqrt_ssf = (float)sqrt(ssf);
    const int nxb = (vx + RX - 1) / RX;
    const int64_t items = (int64_t)vz * vy * nxb;
    const float c = mf;  // window pilot: sum T~ (g - c) = sum T~ g because sum T~ = 0
    for (int64_t it = tid; it < items; it += BLOCK) {
        const int oz = (int)(it / ((int64_t)vy * nxb));
        const int rem = (int)(it - (int64_t)oz * vy * nxb);
        const int oy = rem / nxb;
        const int ox0 = (rem - oy * nxb) * RX;
        float acc[RX], s1[RX], s2[RX];
#pragma unroll
        for (int r = 0; r < RX; ++r) acc[r] = s1[r] = s2[r] = 0.0f;
        for (int tz = 0; tz < sz; ++tz) {
            for (int ty = 0; ty < sy; ++ty) {
                const float* grow = a.g + ((zb + oz + tz) * a.ny + (yb + oy + ty)) * (int64_t)a.nx + xb;
                const float* trow = tpl + (tz * sy + ty) * sx;
                float win[RX];
#pragma unroll
                for (int r = 0; r < RX - 1; ++r) win[r] = grow[min(ox0 + r, wx - 1)] - c;
                for (int tx = 0; tx < sx; ++tx) {
                    win[RX - 1] = grow[min(ox0 + tx + RX - 1, wx - 1)] - c;
                    const float t = trow[tx];
#pragma unroll
                    for (int r = 0; r < RX; ++r) {
                        acc[r] += t * win[r];
                        s1[r] += win[r];
                        s2[r] += win[r] * win[r];
                    }
#pragma unroll
                    for (int r = 0; r < RX - 1; ++r) win[r] = win[r + 1];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RX; ++r) {
            if (ox0 + r >= vx) break;
            const float var = s2[r] - s1[r] * s1[r] * inv_n;
            float val;
            if (var < 1e-20f) {
                val = 0.0f;
            } else {
                val = acc[r] / (sqrtf(var) * sqrt_ssf);
                val = val > 1.0f ? 1.0f : (val < -1.0f ? -1.0f : val);
            }
            o[((int64_t)oz * vy + oy) * vx + ox0 + r] = val;
        }
    }
}

// ---------------------------------------------------------------------------- small search windows
// Search radii up to 4 along x (the pyramid's refinement and coarse levels, 5^3 .. 9^3 offsets):
// k_zncc_map_f32 would leave most of a workgroup idle (one lane per output row). Here the whole
// centred window and template sit in LDS and the work is split over template planes as well:
// item (tz, oz, oy) correlates template plane tz with the window rows of output row (oz, oy),
// RX = 2 rx + 1 outputs per lane with a sliding register window, and writes a partial numerator;
// the partials are then summed over tz in a fixed order (deterministic). The window variance comes
// from separable float32 box sums (x, then y, then z) of the centred window instead of being
// re-accumulated per output. LDS (floats): template sx sy sz + window wx wy wz + max(x-pass sums
// 2 wz wy vx, partials sz vz vy vx) + y-pass sums 2 wz vy vx.
__host__ __device__ inline int64_t small_lds_floats(int sx, int sy, int sz, int vx, int vy, int vz) {
    const int64_t wx = sx + vx - 1, wy = sy + vy - 1, wz = sz + vz - 1;
    const int64_t xs = 2 * wz * wy * vx, part = (int64_t)sz * vz * vy * vx;
    return (int64_t)sx * sy * sz + wx * wy * wz + (xs > part ? xs : part) + 2 * wz * vy * vx;
}

// SX > 0: the template width known at compile time (it must equal 2 hx + 1), so the sliding
// register window of a template row unrolls into plain reuse of one row of loads, summed in the
// same order as the general loop.
template <int RX, int NT, int SX = 0>
__global__ __launch_bounds__(NT) void k_zncc_map_small(NccArgs a) {
    extern __shared__ float sm[];
    const int64_t node = blockIdx.x;
    const int tid = threadIdx.x;
    const int sx = 2 * a.hx + 1, sy = 2 * a.hy + 1, sz = 2 * a.hz + 1;
    const int vy = 2 * a.ry + 1, vz = 2 * a.rz + 1;  // vx == RX
    const int wx = sx + RX - 1, wy = sy + vy - 1, wz = sz + vz - 1;
    const int n_tpl = sx * sy * sz;
    const int n_off = RX * vy * vz;
    float* o = a.out + node * (int64_t)n_off;
    const int64_t x0 = a.coords[3 * node], y0 = a.coords[3 * node + 1], z0 = a.coords[3 * node + 2];
    const int64_t xb = a.wx0[node], yb = a.wy0[node], zb = a.wz0[node];
    const bool tpl_in = x0 - a.hx >= 0 && x0 + a.hx < a.nx && y0 - a.hy >= 0 && y0 + a.hy < a.ny && z0 - a.hz >= 0 &&
                        z0 + a.hz < a.nz;
    const bool win_in = xb >= 0 && xb + wx <= a.nx && yb >= 0 && yb + wy <= a.ny && zb >= 0 && zb + wz <= a.nz;
    if (!a.t_ok[node] || !tpl_in || !win_in) {
        fill_map<NT>(o, n_off, -2.0f);
        return;
    }
    float* tpl = sm;
    float* win = tpl + n_tpl;                  // wz x wy x wx, centred by the template mean
    float* xs1 = win + wz * wy * wx;           // wz x wy x RX row sums (then the partials)
    float* xs2 = xs1 + wz * wy * RX;
    float* part = xs1;                         // sz x n_off partial numerators, after the y pass
    const int n_x = wz * wy * RX, n_part = sz * n_off;
    float* ys1 = xs1 + (2 * n_x > n_part ? 2 * n_x : n_part);  // wz x vy x RX box sums over (x, y)
    float* ys2 = ys1 + wz * vy * RX;
    // the window's first WPF values per lane are read before the template's reductions, so their
    // latency overlaps them; they are centred and stored once the template mean is known
    constexpr int WPF = 20;
    const int n_win = wz * wy * wx;
    const float inv_wplane = 1.0f / (float)(wy * wx), inv_wx = 1.0f / (float)wx;
    auto gwin = [&](int v) {
        const int k = qdiv(v, inv_wplane), r = v - k * wy * wx, j = qdiv(r, inv_wx), i = r - j * wx;
        return a.g[((zb + k) * a.ny + (yb + j)) * (int64_t)a.nx + (xb + i)];
    };
    float pre[WPF];
#pragma unroll
    for (int q = 0; q < WPF; ++q) pre[q] = tid + q * NT < n_win ? gwin(tid + q * NT) : 0.0f;
    double ssf;
    float mf;
    if (!centred_template<NT>(a, tpl, x0, y0, z0, sx, sy, n_tpl, mf, ssf)) {
        fill_map<NT>(o, n_off, -2.0f);
        return;
    }
    const float c = mf;  // window pilot: sum T~ (g - c) = sum T~ g because sum T~ = 0
#pragma unroll
    for (int q = 0; q < WPF; ++q)
        if (tid + q * NT < n_win) win[tid + q * NT] = pre[q] - c;
    for (int v = tid + WPF * NT; v < n_win; v += NT) win[v] = gwin(v) - c;
    __syncthreads();
    // x pass: sums over sx consecutive columns of every window row
    for (int v = tid; v < n_x; v += NT) {
        const int row = v / RX, ox = v - row * RX;
        const float* w = win + row * wx + ox;
        float p1 = 0.0f, p2 = 0.0f;
        for (int tx = 0; tx < sx; ++tx) {
            p1 += w[tx];
            p2 += w[tx] * w[tx];
        }
        xs1[v] = p1;
        xs2[v] = p2;
    }
    __syncthreads();
    // y pass: sums over sy consecutive rows
    const int n_y = wz * vy * RX;
    const float inv_yrow = 1.0f / (float)(vy * RX);
    for (int v = tid; v < n_y; v += NT) {
        const int z = qdiv(v, inv_yrow), r = v - z * vy * RX, oy = r / RX, ox = r - oy * RX;
        const int b = (z * wy + oy) * RX + ox;
        float p1 = 0.0f, p2 = 0.0f;
        for (int ty = 0; ty < sy; ++ty) {
            p1 += xs1[b + ty * RX];
            p2 += xs2[b + ty * RX];
        }
        ys1[v] = p1;
        ys2[v] = p2;
    }
    __syncthreads();
    // numerators: item (tz, oz, oy), RX outputs along x
    const int items = sz * vz * vy;
    const float inv_zy = 1.0f / (float)(vz * vy), inv_vy = 1.0f / (float)vy;
    for (int it = tid; it < items; it += NT) {
        const int tz = qdiv(it, inv_zy), r = it - tz * vz * vy, oz = qdiv(r, inv_vy), oy = r - oz * vy;
        float acc[RX];
#pragma unroll
        for (int q = 0; q < RX; ++q) acc[q] = 0.0f;
        for (int ty = 0; ty < sy; ++ty) {
            const float* wrow = win + ((oz + tz) * wy + oy + ty) * wx;
            const float* trow = tpl + (tz * sy + ty) * sx;
            if constexpr (SX > 0) {
                float w[SX + RX - 1];
#pragma unroll
                for (int i = 0; i < SX + RX - 1; ++i) w[i] = wrow[i];
#pragma unroll
                for (int tx = 0; tx < SX; ++tx) {
                    const float t = trow[tx];
#pragma unroll
                    for (int q = 0; q < RX; ++q) acc[q] += t * w[tx + q];
                }
            } else {
                float w[RX];
#pragma unroll
                for (int q = 0; q < RX - 1; ++q) w[q] = wrow[q];
                for (int tx = 0; tx < sx; ++tx) {
                    w[RX - 1] = wrow[tx + RX - 1];
                    const float t = trow[tx];
#pragma unroll
                    for (int q = 0; q < RX; ++q) acc[q] += t * w[q];
#pragma unroll
                    for (int q = 0; q < RX - 1; ++q) w[q] = w[q + 1];
                }
            }
        }
        float* dst = part + tz * n_off + (oz * vy + oy) * RX;
#pragma unroll
        for (int q = 0; q < RX; ++q) dst[q] = acc[q];
    }
    __syncthreads();
    const float inv_n = 1.0f / (float)n_tpl;
    const float sqrt_ssf = (float)sqrt(ssf);
    for (int off = tid; off < n_off; off += NT) {
        const int oz = qdiv(off, inv_yrow), r = off - oz * vy * RX;  // r = oy * RX + ox
        float num = 0.0f, s1 = 0.0f, s2 = 0.0f;
        for (int tz = 0; tz < sz; ++tz) {
            num += part[tz * n_off + off];
            s1 += ys1[(oz + tz) * vy * RX + r];
            s2 += ys2[(oz + tz) * vy * RX + r];
        }
        const float var = s2 - s1 * s1 * inv_n;
        float val;
        if (var < 1e-20f) {
            val = 0.0f;
        } else {
            val = num / (sqrtf(var) * sqrt_ssf);
            val = val > 1.0f ? 1.0f : (val < -1.0f ? -1.0f : val);
        }
        o[off] = val;
    }
}

// ---------------------------------------------------------------------------- BF16 MFMA maps
// The correlation C[oz][oy][ox] = sum_{tz,ty,tx} T~[tz][ty][tx] (W[oz+tz][oy+ty][ox+tx] - c) is
// a sum over template rows (tz, ty) of 1-D correlations along x, and a 1-D correlation with a
// fixed row is a product with a banded (Toeplitz) matrix: for one template row,
//     D[oy][ox] += sum_k W[z][oy + ty][k] * Band[k][ox],   Band[k][ox] = T~[tz][ty][k - ox],
// a GEMM with M = oy, N = ox, K = window columns. v_mfma_f32_16x16x32_bf16 computes a 16x16
// output tile over a 32-wide K slice; an output x-tile of 16 columns needs K = 15 + sx window
// columns (KS slices). Operands are bfloat16 (window centred by the template mean c before the
// rounding, template centred by its mean), accumulation float32.
//
// Two kernels per call. k_box_var computes the window variance of every offset,
// var = sum (W - c)^2 - (sum (W - c))^2 / n over the template-sized box, with float32 sliding
// sums (only the numerator sees bfloat16 rounding) and writes it into the map buffer.
// k_zncc_map_bf16 then computes the numerator on MFMA and replaces var by the ZNCC in place. The
// MFMA kernel thus only stages bfloat16 planes: one workgroup = 4 wave64 = 4 consecutive output
// planes oz0 .. oz0 + 3 (one per wave) of one node; window plane z feeds wave g with template
// plane tz = z - oz0 - g, so every staged plane serves all four waves, and the next plane's
// global loads are issued before the current plane's MFMAs (two window buffers, five template
// slots) so that they overlap. (Two output planes per wave, sharing each A fragment between two
// MFMAs, needs ~310 registers: one wave per SIMD, and measured 12 % slower.)
//
// MFMA operand layout (16x16x32, wave64): lane l holds A[m = l % 16][k = 8 (l / 16) + i] and
// B[k = 8 (l / 16) + i][n = l % 16] for i < 8, and D[m = 4 (l / 16) + i][n = l % 16] for i < 4.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int MG = 4;     // output planes per workgroup of the MFMA kernel (one per wave)
constexpr int TRING = 5;  // template plane slots: four in use and the one being loaded
constexpr int PRE = 16;   // window values per lane prefetched for the next plane (57^2 / 256 <= 13)

// A fragment rows moved down by one within each 16-lane DPP row: lane m takes lane m + 1's value
// (row_shl:1); lane 15 of each row, which has no source, reads zero (bound_ctrl).
__device__ inline bf16x8_t shift_rows(bf16x8_t frag) {
    int v[4];
    __builtin_memcpy(v, &frag, sizeof(v));
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = __builtin_amdgcn_mov_dpp(v[i], 0x101, 0xF, 0xF, true);
    bf16x8_t r;
    __builtin_memcpy(&r, v, sizeof(r));
    return r;
}

struct Bf16Geom {
    int sx, sy, sz, vx, vy, vz, wx, wy, wz;
    int nyt, nxt;  // 16-row / 16-column output tiles
    int AH, AW;    // staged window plane (rows, columns) incl. zero padding
    int TW;        // staged template row width: x in [-16, 32 KS)
};

__device__ inline float bf16_round_trip(float v) { return (float)(__bf16)v; }

// Shared prologue of the bfloat16 kernels (NT threads): whether the node has a map, and the
// template mean and energy (float64). Returns false when the map is the flat-template value -2.
template <int NT = BLOCK>
__device__ inline bool ncc_node_stats(const NccArgs& a, const Bf16Geom& q, int64_t node, double& mf, double& ssf) {
    constexpr int NW = NT / 64;
    __shared__ double red[NW];
    __shared__ double stat[2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t x0 = a.coords[3 * node], y0 = a.coords[3 * node + 1], z0 = a.coords[3 * node + 2];
    const int64_t xb = a.wx0[node], yb = a.wy0[node], zb = a.wz0[node];
    const bool tpl_in = x0 - a.hx >= 0 && x0 + a.hx < a.nx && y0 - a.hy >= 0 && y0 + a.hy < a.ny && z0 - a.hz >= 0 &&
                        z0 + a.hz < a.nz;
    const bool win_in = xb >= 0 && xb + q.wx <= a.nx && yb >= 0 && yb + q.wy <= a.ny && zb >= 0 && zb + q.wz <= a.nz;
    if (!a.t_ok[node] || !tpl_in || !win_in) return false;  // block-uniform
    const int n_tpl = q.sx * q.sy * q.sz;
    const int64_t tbase = ((z0 - a.hz) * a.ny + (y0 - a.hy)) * (int64_t)a.nx + (x0 - a.hx);
    double sacc = 0.0;
    for (int v = tid; v < n_tpl; v += NT) {
        const int k = v / (q.sy * q.sx), r = v - k * q.sy * q.sx, j = r / q.sx, i = r - j * q.sx;
        sacc += a.f[tbase + ((int64_t)k * a.ny + j) * a.nx + i];
    }
    sacc = wave_sum(sacc);
    if (lane == 0) red[wave] = sacc;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < NW; ++w) t += red[w];
        stat[0] = t / n_tpl;
    }
    __syncthreads();
    mf = stat[0];
    double ss = 0.0;
    for (int v = tid; v < n_tpl; v += NT) {
        const int k = v / (q.sy * q.sx), r = v - k * q.sy * q.sx, j = r / q.sx, i = r - j * q.sx;
        const double d = (double)a.f[tbase + ((int64_t)k * a.ny + j) * a.nx + i] - mf;
        ss += d * d;
    }
    ss = wave_sum(ss);
    __syncthreads();
    if (lane == 0) red[wave] = ss;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < NW; ++w) t += red[w];
        stat[1] = t;
    }
    __syncthreads();
    ssf = stat[1];
    return ssf >= 1e-20;
}

// Window variance of every offset: one workgroup per node walks the window planes once. Per plane,
// 2-D box sums (x then y, sliding in runs of 8 outputs) are added to running sums over the last sz
// planes; the plane leaving the box (sz planes back) has its 2-D sums recomputed and subtracted,
// and every plane from sz - 1 on completes one output plane. A thread owns the offsets
// tid + 256 k of the plane. The global loads of the next entering and leaving planes are issued
// into registers before the current planes' sums, so their latency is hidden behind that work.
constexpr int VPT = 9;   // offsets per thread and plane: vy * vx <= 256 * VPT (radii up to 23)
constexpr int VPL = 16;  // window values per thread and plane: wy * wx <= 256 * VPL

// plane z of the node's window, less c, into registers (element tid + 256 u, row-major)
template <int NV>
__device__ inline void box_load(const NccArgs& a, const Bf16Geom& q, int64_t node, int z, float (&v)[NV]) {
    const int64_t xb = a.wx0[node], yb = a.wy0[node], zb = a.wz0[node];
    const float* gp = a.g + ((zb + z) * a.ny + yb) * (int64_t)a.nx + xb;
    const int wpl = q.wy * q.wx;
    const float inv_wx = 1.0f / (float)q.wx;  // i < 2^22: the float floor of (i + 0.5) / wx is exact
#pragma unroll
    for (int u = 0; u < NV; ++u) {
        const int i = threadIdx.x + BLOCK * u;
        if (i < wpl) {
            const int y = (int)(((float)i + 0.5f) * inv_wx), x = i - y * q.wx;
            v[u] = gp[(int64_t)y * a.nx + x];
        }
    }
}

// Window variance of every offset, z first: each thread keeps, for its elements of the window plane,
// the running sums over the last sz planes of W - c and of its square (registers; the plane that
// leaves is re-read and subtracted), and each output plane is then the 2-D box sum of those two
// z-summed planes in LDS -- one 2-D pass per output plane rather than one per entering and per leaving
// window plane. The 2-D pass: row sums over sx columns (running, 8 outputs per item), then column
// sums over sy rows, written out as S2 - S1^2 / n.
// NV: window elements per thread and plane (wy * wx <= BLOCK * NV), sized to the window so that the
// six per-thread arrays stay in registers.
template <int NV>
__global__ __launch_bounds__(BLOCK, 3) void k_box_var(NccArgs a, Bf16Geom q) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int tid = threadIdx.x;
    const int64_t node = blockIdx.x;
    double mf, ssf;
    if (!ncc_node_stats(a, q, node, mf, ssf)) return;  // the MFMA kernel writes the -2 map
    const float c = (float)mf;
    const int wpl = q.wy * q.wx, nplane = q.vy * q.vx;
    float* Z1 = reinterpret_cast<float*>(lds);  // wy x wx: sums over the last sz planes of W - c
    float* Z2 = Z1 + wpl;                       // ... and of (W - c)^2
    float* R1 = Z2 + wpl;                       // wy x vx row sums of Z1
    float* R2 = R1 + q.wy * q.vx;               // ... of Z2
    const float inv_n = 1.0f / (float)(q.sx * q.sy * q.sz);
    const int nseg_x = (q.vx + 7) / 8, nseg_y = (q.vy + 7) / 8;
    float* o = a.out + node * (int64_t)nplane * q.vz;
    float z1[NV], z2[NV];
#pragma unroll
    for (int u = 0; u < NV; ++u) z1[u] = z2[u] = 0.0f;
    float vin[NV], vout[NV], nin[NV], nout[NV];
    box_load(a, q, node, 0, vin);
    for (int z = 0; z < q.wz; ++z) {
        if (z + 1 < q.wz) box_load(a, q, node, z + 1, nin);  // next planes in flight
        if (z + 1 >= q.sz && z + 1 < q.wz) box_load(a, q, node, z + 1 - q.sz, nout);
#pragma unroll
        for (int u = 0; u < NV; ++u) {
            const float w = vin[u] - c;
            z1[u] += w;
            z2[u] += w * w;
        }
        if (z >= q.sz) {  // the plane that left the box
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                const float w = vout[u] - c;
                z1[u] -= w;
                z2[u] -= w * w;
            }
        }
        if (z >= q.sz - 1) {
            // the z-summed planes into LDS (the previous output plane's readers passed the barrier
            // after its row sums, so Z1/Z2 are free)
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                const int i = tid + BLOCK * u;
                if (i < wpl) {
                    Z1[i] = z1[u];
                    Z2[i] = z2[u];
                }
            }
            __syncthreads();
            for (int t = tid; t < q.wy * nseg_x; t += BLOCK) {  // row sums, 8 outputs per item
                const int y = t / nseg_x, ox0 = 8 * (t - y * nseg_x);
                const float* r1p = Z1 + y * q.wx;
                const float* r2p = Z2 + y * q.wx;
                float r1 = 0.0f, r2 = 0.0f;
#pragma unroll 8
                for (int x = 0; x < q.sx; ++x) {
                    r1 += r1p[ox0 + x];
                    r2 += r2p[ox0 + x];
                }
                R1[y * q.vx + ox0] = r1;
                R2[y * q.vx + ox0] = r2;
#pragma unroll
                for (int ox = ox0 + 1; ox < ox0 + 8; ++ox) {
                    if (ox >= q.vx) break;
                    r1 += r1p[ox + q.sx - 1] - r1p[ox - 1];
                    r2 += r2p[ox + q.sx - 1] - r2p[ox - 1];
                    R1[y * q.vx + ox] = r1;
                    R2[y * q.vx + ox] = r2;
                }
            }
            __syncthreads();
            float* op = o + (int64_t)(z - q.sz + 1) * nplane;
            for (int t = tid; t < q.vx * nseg_y; t += BLOCK) {  // column sums, 8 outputs per item
                const int ox = t / nseg_y, oy0 = 8 * (t - ox * nseg_y);
                float r1 = 0.0f, r2 = 0.0f;
#pragma unroll 8
                for (int y = 0; y < q.sy; ++y) {
                    r1 += R1[(oy0 + y) * q.vx + ox];
                    r2 += R2[(oy0 + y) * q.vx + ox];
                }
                op[oy0 * q.vx + ox] = r2 - r1 * r1 * inv_n;
#pragma unroll
                for (int oy = oy0 + 1; oy < oy0 + 8; ++oy) {
                    if (oy >= q.vy) break;
                    r1 += R1[(oy + q.sy - 1) * q.vx + ox] - R1[(oy - 1) * q.vx + ox];
                    r2 += R2[(oy + q.sy - 1) * q.vx + ox] - R2[(oy - 1) * q.vx + ox];
                    op[oy * q.vx + ox] = r2 - r1 * r1 * inv_n;
                }
            }
            // the next output plane's Z1/Z2 stores wait for this plane's row sums at the barrier
            // above; its row sums overwrite R1/R2 only after its own first barrier
        }
#pragma unroll
        for (int u = 0; u < NV; ++u) {
            vin[u] = nin[u];
            vout[u] = nout[u];
        }
    }
}

template <int NT, int KS>
__global__ __launch_bounds__(256) void k_zncc_map_bf16(NccArgs a, Bf16Geom q) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n_zgrp = (q.vz + MG - 1) / MG;
    const int64_t node = blockIdx.x / n_zgrp;
    const int oz0 = (int)(blockIdx.x % n_zgrp) * MG;
    const int64_t n_off = (int64_t)q.vx * q.vy * q.vz;
    float* out = a.out + node * n_off;
    const int my_oz = oz0 + wave;
    auto fill_mine = [&](float v) {
        if (my_oz >= q.vz) return;
        float* o = out + (int64_t)my_oz * q.vy * q.vx;
        for (int i = lane; i < q.vy * q.vx; i += 64) o[i] = v;
    };
    double mf, ssf;
    if (!ncc_node_stats(a, q, node, mf, ssf)) {
        fill_mine(-2.0f);
        return;
    }
    const int64_t x0 = a.coords[3 * node], y0 = a.coords[3 * node + 1], z0 = a.coords[3 * node + 2];
    const int64_t xb = a.wx0[node], yb = a.wy0[node], zb = a.wz0[node];
    const int64_t tbase = ((z0 - a.hz) * a.ny + (y0 - a.hy)) * (int64_t)a.nx + (x0 - a.hx);
    const float c = (float)mf;
    // ---- LDS: two window planes (bf16, zero-padded) and TRING template planes
    const int plane_el = q.AH * q.AW;
    __bf16* Ab = (__bf16*)lds;
    // template slots twice: as staged, and shifted by one element (Tr1[e] = Tr[e - 1]) so that every
    // lane's B fragment starts on a 4-byte boundary in one of them
    __bf16* Tr = Ab + 2 * plane_el;
    __bf16* Tr1 = Tr + TRING * q.sy * q.TW;
    for (int i = tid; i < 2 * plane_el; i += BLOCK) Ab[i] = (__bf16)0.0f;
    for (int i = tid; i < 2 * TRING * q.sy * q.TW; i += BLOCK) Tr[i] = (__bf16)0.0f;
    __syncthreads();
    // per-lane prefetch of one window plane (and one template plane)
    const int wpl = q.wy * q.wx, tpl_n = q.sy * q.sx;
    float pre[PRE], tpre[4];
    auto load_plane = [&](int z) {
        const float* gp = a.g + ((zb + z) * a.ny + yb) * (int64_t)a.nx + xb;
#pragma unroll
        for (int u = 0; u < PRE; ++u) {
            const int i = tid + BLOCK * u;
            if (i < wpl) {
                const int y = i / q.wx, x = i - y * q.wx;
                pre[u] = gp[(int64_t)y * a.nx + x];
            }
        }
        const int tz = z - oz0;
        if (tz < q.sz) {
            const float* fp = a.f + tbase + (int64_t)tz * a.ny * a.nx;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = tid + BLOCK * u;
                if (i < tpl_n) {
                    const int y = i / q.sx, x = i - y * q.sx;
                    tpre[u] = fp[(int64_t)y * a.nx + x];
                }
            }
        }
    };
    auto store_plane = [&](int z) {
        __bf16* Ap = Ab + (z & 1) * plane_el;
#pragma unroll
        for (int u = 0; u < PRE; ++u) {
            const int i = tid + BLOCK * u;
            if (i < wpl) {
                const int y = i / q.wx, x = i - y * q.wx;
                Ap[y * q.AW + x] = (__bf16)(pre[u] - c);
            }
        }
        const int tz = z - oz0;
        if (tz < q.sz) {
            __bf16* tp = Tr + (tz % TRING) * q.sy * q.TW;
            __bf16* tp1 = Tr1 + (tz % TRING) * q.sy * q.TW;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = tid + BLOCK * u;
                if (i < tpl_n) {
                    const int y = i / q.sx, x = i - y * q.sx;
                    const __bf16 v = (__bf16)(float)((double)tpre[u] - mf);
                    tp[y * q.TW + 16 + x] = v;
                    tp1[y * q.TW + 17 + x] = v;
                }
            }
        }
    };
    f32x4_t acc[NT][NT];
#pragma unroll
    for (int mt = 0; mt < NT; ++mt)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[mt][j] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    const int am = lane & 15, akb = 8 * (lane >> 4);  // A row / K offset, also B column / K offset
    const int z_end = min(oz0 + MG + q.sz - 1, q.wz);
    load_plane(oz0);
    store_plane(oz0);
    __syncthreads();
    for (int z = oz0; z < z_end; ++z) {
        const bool more = z + 1 < z_end;
        if (more) load_plane(z + 1);  // in flight during this plane's MFMAs
        const int tz = z - my_oz;
        if (my_oz < q.vz && tz >= 0 && tz < q.sz) {
            const __bf16* Ap = Ab + (z & 1) * plane_el;
            // B fragment of lane (n, g): elements 16 + 32 s + 8 g - n + i of the staged row; odd n read
            // the shifted copy one element further, so the start is even (4-byte aligned) for all lanes
            const int odd = am & 1;
            const __bf16* tpl = (odd ? Tr1 : Tr) + (tz % TRING) * q.sy * q.TW + 16 + akb - am + odd;
            // A fragments of the window rows 16 mt + m + ty at the NC distinct column offsets
            // 16 (j + 2 s), held in registers across the template rows: from row ty to ty + 1 every
            // lane takes its lower neighbour's row (a DPP shift inside the 16 lanes of a K group)
            // and only lane 15 of each group reads a new row from LDS.
            constexpr int NC = NT + 2 * (KS - 1);
            bf16x8_t afr[NT][NC];
#pragma unroll
            for (int mt = 0; mt < NT; ++mt) {
                if (mt >= q.nyt) break;
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    __builtin_memcpy(&afr[mt][c], Ap + (16 * mt + am) * q.AW + 16 * c + akb, sizeof(bf16x8_t));
            }
            // B[k][n] = T~[tz][ty][32 s + k - n]: four aligned dwords per slice
            auto load_b = [&](int ty, bf16x8_t* dst) {
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const uint32_t* w = reinterpret_cast<const uint32_t*>(tpl + ty * q.TW + 32 * s);
                    uint32_t v[4] = {w[0], w[1], w[2], w[3]};
                    __builtin_memcpy(&dst[s], v, sizeof(v));
                }
            };
            bf16x8_t bfr[KS];
            load_b(0, bfr);
            for (int ty = 0; ty < q.sy; ++ty) {
#pragma unroll
                for (int mt = 0; mt < NT; ++mt) {
                    if (mt >= q.nyt) break;
#pragma unroll
                    for (int j = 0; j < NT; ++j) {
                        if (j >= q.nxt) break;
#pragma unroll
                        for (int s = 0; s < KS; ++s)
                            acc[mt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[mt][j + 2 * s], bfr[s], acc[mt][j], 0, 0, 0);
                    }
                }
                if (ty + 1 < q.sy) {
                    // the next row's B fragments and, tile by tile, the A rows one further: their LDS
                    // latency overlaps the row shifts of the later tiles
                    load_b(ty + 1, bfr);
#pragma unroll
                    for (int mt = 0; mt < NT; ++mt) {
                        if (mt >= q.nyt) break;
#pragma unroll
                        for (int c = 0; c < NC; ++c) afr[mt][c] = shift_rows(afr[mt][c]);
                        if (am == 15) {  // lane 15 of each K group: window row 16 mt + 15 + (ty + 1)
#pragma unroll
                            for (int c = 0; c < NC; ++c)
                                __builtin_memcpy(&afr[mt][c], Ap + (16 * mt + 16 + ty) * q.AW + 16 * c + akb, sizeof(bf16x8_t));
                        }
                    }
                }
            }
        }
        if (more) store_plane(z + 1);  // the other window buffer and a free template slot
        __syncthreads();
    }
    if (my_oz >= q.vz) return;
    const float sq_ssf = (float)sqrt(ssf);
    float* o = out + (int64_t)my_oz * q.vy * q.vx;
#pragma unroll
    for (int mt = 0; mt < NT; ++mt)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int ox = 16 * j + (lane & 15);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int oy = 16 * mt + 4 * (lane >> 4) + i;
                if (mt < q.nyt && j < q.nxt && oy < q.vy && ox < q.vx) {
                    const float var = o[oy * q.vx + ox];  // written by k_box_var
                    float val = 0.0f;
                    if (var >= 1e-20f) {
                        val = acc[mt][j][i] / (sqrtf(var) * sq_ssf);
                        val = val > 1.0f ? 1.0f : (val < -1.0f ? -1.0f : val);
                    }
                    o[oy * q.vx + ox] = val;
                }
            }
        }
}

// ---------------------------------------------------------------------------- BF16 MFMA maps, skewed planes
// A second tiling of the same correlation for templates up to 32 voxels wide and up to 48 x
// offsets (25^3 templates and 33^3 offsets at radius 16) that wastes less of each MFMA than the
// banded one above (whose K slices are 64 wide for 25 taps):
//   M = 16 x offsets ox (a column tile c), K = 32 template columns tx (zero from sx on),
//   N = 16 = 8 output planes oz0 + j x 2 output rows 2G + d (d = 0, 1).
// Window row y of plane z, A[m][k] = W~[z][y][16c + m + k], meets output row oy through template
// row y - oy and output plane oz0 + j through template plane z - oz0 - j, so for the accumulator
// tile (G, c) (rows 2G and 2G + 1) it is multiplied by
//   B_t[k][(j, d)] = T~[z - oz0 - j][t - d][k],   t = y - 2G in [0, SY],
// zero where that template plane or row does not exist. A wave holds the SY + 1 fragments B_t of
// the current window plane and its accumulator tiles in registers, and reads each window row's A
// fragment from LDS once for every tile of its column that the row reaches. The row and tile loops
// are unrolled (SY, VY and the tile split are template parameters), so every register index is
// static and only the (y, G) pairs with 0 <= y - 2G <= SY are issued. Work split over the NC + 1
// waves of a workgroup: wave c < NC takes column tile c for the tiles G < GA, wave NC takes every
// column tile for G >= GA, with GA chosen to even the MFMA counts out (13 / 13 / 13 / 12 tiles at
// VY = 33). Window planes are double-buffered in LDS as bf16 W - c, twice (the copy shifted by
// one element makes every lane's 8-element row start 4-byte aligned); template planes sit in a
// ring of SK_MZ + 1 slots (the one being loaded included). Output planes in workgroups of SK_MZ.
constexpr int SK_MZ = 8;        // output planes per workgroup
constexpr int SK_RING = SK_MZ + 1;  // template plane slots
constexpr int SK_PITCH = 80;    // staged window row (bf16): three column tiles + 32 taps + shift
constexpr int SK_TP = 32;       // staged template row (bf16), zero from sx on

// bf16 elements of one template slot of `rows` rows plus a zero row on either side (rows -1 and
// `rows`, read by the B fragments that reach past the template): 4 mod 32 dwords, so the 16-byte B
// reads of the lanes of one column group (eight consecutive template planes, two rows) fall on
// different banks
__host__ __device__ constexpr int sk_slot_elems(int rows) {
    const int dw = (rows + 2) * SK_TP / 2;
    return 2 * (dw + ((4 - dw % 32) + 32) % 32);
}
constexpr int SK_NT = 256;      // up to three column waves and the remainder wave
constexpr int SK_PRE = 13;      // window values per lane prefetched for the next plane
constexpr int SK_TPRE = 3;      // template values per lane prefetched for the next plane

// bf16 elements of one copy of a staged window plane of `rows` rows: 16 mod 32 dwords, so the even-
// and odd-column lanes of a half-wave, which read the two copies, use different banks (a half-wave
// of one A read covers 12 consecutive dwords of each copy)
__host__ __device__ constexpr int sk_copy_elems(int rows) {
    const int dw = rows * SK_PITCH / 2;
    return 2 * (dw + ((16 - dw % 32) + 32) % 32);
}

template <int SY, int VY, int NC>
struct SkewShape {
    static constexpr int NG = (VY + 1) / 2;               // row tiles (output rows 2G, 2G + 1)
    static constexpr int GA = (NC * NG + NC) / (NC + 1);  // tiles of the column waves
    static constexpr int NB = SY + 1;                     // B fragments per window plane
    static constexpr int WROWS = 2 * (NG - 1) + SY + 1;   // window rows reached (rows from wy on are zero)
    static constexpr int COPY = sk_copy_elems(WROWS);
};

// The MFMAs of one window plane for the tiles G0 .. G0 + NT - 1 (of NG) of NCOL column tiles.
template <int SY, int NG, int G0, int NT, int NCOL>
__device__ inline void skew_plane(f32x4_t (&acc)[NCOL][NT], const bf16x8_t (&bfr)[SY + 1], const uint32_t* ap) {
    constexpr int ROWS = 2 * (NT - 1) + SY + 1;
#pragma unroll
    for (int yr = 0; yr < ROWS; ++yr) {  // window row 2 G0 + yr
        bf16x8_t afr[NCOL];
#pragma unroll
        for (int cc = 0; cc < NCOL; ++cc) {
            const uint32_t* p = ap + (2 * G0 + yr) * (SK_PITCH / 2) + 8 * cc;
            const uint32_t v4[4] = {p[0], p[1], p[2], p[3]};
            __builtin_memcpy(&afr[cc], v4, sizeof(bf16x8_t));
        }
#pragma unroll
        for (int gl = 0; gl < NT; ++gl) {
            const int t = yr - 2 * gl;
            if (t >= 0 && t <= SY && G0 + gl < NG) {
#pragma unroll
                for (int cc = 0; cc < NCOL; ++cc)
                    acc[cc][gl] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[cc], bfr[t], acc[cc][gl], 0, 0, 0);
            }
        }
    }
}

template <int SY, int VY, int NC>
__global__ __launch_bounds__(SK_NT, 2) void k_zncc_map_bf16_skew(NccArgs a, Bf16Geom q, int nzt) {
    using S = SkewShape<SY, VY, NC>;
    constexpr int COPY = S::COPY;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t node = blockIdx.x / nzt;
    const int oz0 = (int)(blockIdx.x % nzt) * SK_MZ;
    const int nzp = min(SK_MZ, q.vz - oz0);  // output planes of this workgroup
    const int64_t plane_out = (int64_t)q.vy * q.vx;
    float* out = a.out + node * plane_out * q.vz + (int64_t)oz0 * plane_out;
    double mf, ssf;
    if (!ncc_node_stats<SK_NT>(a, q, node, mf, ssf)) {
        for (int64_t i = tid; i < nzp * plane_out; i += SK_NT) out[i] = -2.0f;
        return;
    }
    const int64_t x0 = a.coords[3 * node], y0 = a.coords[3 * node + 1], z0 = a.coords[3 * node + 2];
    const int64_t xb = a.wx0[node], yb = a.wy0[node], zb = a.wz0[node];
    const int64_t tbase = ((z0 - a.hz) * a.ny + (y0 - a.hy)) * (int64_t)a.nx + (x0 - a.hx);
    const float cw = (float)mf;
    __bf16* win = reinterpret_cast<__bf16*>(lds);  // [buffer][copy][WROWS][SK_PITCH]
    // [slot][row -1 .. SY][SK_TP], slots SLOT elements apart; slot SK_RING stays zero (the B source
    // of lanes whose template plane does not exist)
    __bf16* tring = win + 4 * COPY;
    constexpr int SLOT = sk_slot_elems(SY);
    for (int i = tid; i < 4 * COPY + (SK_RING + 1) * SLOT; i += SK_NT) win[i] = (__bf16)0.0f;
    __syncthreads();
    const int wpl = q.wy * q.wx, tpl_n = SY * q.sx;
    // row of element i of a window or template plane as a float product: (i + 0.5) / w lies at least
    // 0.5 / w from an integer and i < 2^22, so the rounding cannot change the floor
    const float inv_wx = 1.0f / (float)q.wx, inv_sx = 1.0f / (float)q.sx;
    float pre[SK_PRE], tpre[SK_TPRE];
    auto load_plane = [&](int z) {
        const float* gp = a.g + ((zb + z) * a.ny + yb) * (int64_t)a.nx + xb;
#pragma unroll
        for (int u = 0; u < SK_PRE; ++u) {
            const int i = tid + SK_NT * u;
            if (i < wpl) {
                const int y = (int)(((float)i + 0.5f) * inv_wx), x = i - y * q.wx;
                pre[u] = gp[(int64_t)y * a.nx + x];
            }
        }
        const int tz = z - oz0;
        if (tz < q.sz) {
            const float* fp = a.f + tbase + (int64_t)tz * a.ny * a.nx;
#pragma unroll
            for (int u = 0; u < SK_TPRE; ++u) {
                const int i = tid + SK_NT * u;
                if (i < tpl_n) {
                    const int y = (int)(((float)i + 0.5f) * inv_sx), x = i - y * q.sx;
                    tpre[u] = fp[(int64_t)y * a.nx + x];
                }
            }
        }
    };
    auto store_plane = [&](int z) {
        __bf16* w0 = win + (z & 1) * 2 * COPY;
#pragma unroll
        for (int u = 0; u < SK_PRE; ++u) {
            const int i = tid + SK_NT * u;
            if (i < wpl) {
                const int y = (int)(((float)i + 0.5f) * inv_wx), x = i - y * q.wx;
                const __bf16 v = (__bf16)(pre[u] - cw);
                w0[y * SK_PITCH + x] = v;
                w0[COPY + y * SK_PITCH + x + 1] = v;
            }
        }
        const int tz = z - oz0;
        if (tz < q.sz) {
            __bf16* tp = tring + (tz % SK_RING) * SLOT + SK_TP;  // row 0 of the slot
#pragma unroll
            for (int u = 0; u < SK_TPRE; ++u) {
                const int i = tid + SK_NT * u;
                if (i < tpl_n) {
                    const int y = (int)(((float)i + 0.5f) * inv_sx), x = i - y * q.sx;
                    tp[y * SK_TP + x] = (__bf16)(float)((double)tpre[u] - mf);
                }
            }
        }
    };
    const int n = lane & 15, j = n >> 1, d = n & 1, g8 = 8 * (lane >> 4);
    const int m = lane & 15;
    const int z_end = min(oz0 + nzp - 1 + q.sz, q.wz);
    const float sq_ssf = (float)sqrt(ssf);
    // the whole plane loop for one wave role: NT tiles from G0 over NCOL column tiles from c0
    auto run = [&](auto& acc, auto g0_tag, int c0) {
        constexpr int G0 = decltype(g0_tag)::value;
        constexpr int NCOL = sizeof(acc) / sizeof(acc[0]);
        constexpr int NT = sizeof(acc[0]) / sizeof(f32x4_t);
#pragma unroll
        for (int cc = 0; cc < NCOL; ++cc)
#pragma unroll
            for (int gl = 0; gl < NT; ++gl) acc[cc][gl] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
        // this lane's A source for column tile c0, row 0: columns 16 c0 + m + g8 .. + 8, from the
        // copy that makes the start even
        const int acol = 16 * c0 + m + g8 + (m & 1);
        load_plane(oz0);
        store_plane(oz0);
        __syncthreads();
        for (int z = oz0; z < z_end; ++z) {
            const bool more = z + 1 < z_end;
            if (more) load_plane(z + 1);  // in flight during this plane's MFMAs
            const int tz = z - oz0 - j;
            const bool zok = j < nzp && tz >= 0 && tz < q.sz;
            // row t - d of the lane's template plane (rows -1 and SY are the slot's zero rows)
            const __bf16* tp = tring + (zok ? tz % SK_RING : SK_RING) * SLOT + (1 - d) * SK_TP + g8;
            bf16x8_t bfr[SY + 1];
#pragma unroll
            for (int t = 0; t <= SY; ++t) __builtin_memcpy(&bfr[t], tp + t * SK_TP, sizeof(bf16x8_t));
            const uint32_t* ap =
                reinterpret_cast<const uint32_t*>(win + (z & 1) * 2 * COPY + (m & 1) * COPY + acol);
            skew_plane<SY, S::NG, G0, NT, NCOL>(acc, bfr, ap);
            if (more) store_plane(z + 1);  // the other window buffer and a free template slot
            __syncthreads();
        }
        // epilogue: the variance from k_box_var in place, replaced by the ZNCC
#pragma unroll
        for (int cc = 0; cc < NCOL; ++cc)
#pragma unroll
            for (int gl = 0; gl < NT; ++gl) {
                const int oy = 2 * (G0 + gl) + d;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int ox = 16 * (c0 + cc) + 4 * (lane >> 4) + i;
                    if (j < nzp && oy < q.vy && ox < q.vx) {
                        float* o = out + (int64_t)j * plane_out + (int64_t)oy * q.vx + ox;
                        const float var = *o;
                        float val = 0.0f;
                        if (var >= 1e-20f) {
                            val = acc[cc][gl][i] / (sqrtf(var) * sq_ssf);
                            val = val > 1.0f ? 1.0f : (val < -1.0f ? -1.0f : val);
                        }
                        *o = val;
                    }
                }
            }
    };
    if (wave < NC) {
        f32x4_t acc[1][S::GA];
        run(acc, std::integral_constant<int, 0>{}, wave);
    } else if (wave == NC) {
        f32x4_t acc[NC][S::NG - S::GA];
        run(acc, std::integral_constant<int, S::GA>{}, 0);
    } else {  // no tiles: staging and barriers only
        f32x4_t acc[1][1];
        (void)acc;
        load_plane(oz0);
        store_plane(oz0);
        __syncthreads();
        for (int z = oz0; z < z_end; ++z) {
            const bool more = z + 1 < z_end;
            if (more) {
                load_plane(z + 1);
                store_plane(z + 1);
            }
            __syncthreads();
        }
    }
}

// The same skewed-plane tiling in float32 on v_mfma_f32_16x16x4f32 (K = 4 template columns per
// instruction, ceil(sx / 4) slices): the float32 maps of the large-radius searches, so that the
// bfloat16 speed-up is quoted against the same design in full precision. A and B are one float per
// lane (A[m][k] = W~[z][y][16c + m + 4s + k], B_t[k][(j, d)] = T~[z - oz0 - j][t - d][4s + k]), read
// as aligned dwords, so the window needs no shifted copy; the slice loop is outermost within a plane
// so only one slice's SY + 1 B values are live.
constexpr int SKF_PITCH = 80;  // staged window row (float): three column tiles + 32 taps

template <int SY, int VY, int NC>
struct SkewShapeF32 {
    using S = SkewShape<SY, VY, NC>;
    static constexpr int WIN = S::WROWS * SKF_PITCH + 16;          // floats per window buffer
    static constexpr int SLOT = (SY + 2) * SK_TP + 4;               // floats per template slot (rows -1 .. SY)
};

template <int SY, int NG, int G0, int NT, int NCOL>
__device__ inline void skew_plane_f32(f32x4_t (&acc)[NCOL][NT], const float (&bfr)[SY + 1], const float* ap) {
    constexpr int ROWS = 2 * (NT - 1) + SY + 1;
#pragma unroll
    for (int yr = 0; yr < ROWS; ++yr) {
        float afr[NCOL];
#pragma unroll
        for (int cc = 0; cc < NCOL; ++cc) afr[cc] = ap[(2 * G0 + yr) * SKF_PITCH + 16 * cc];
#pragma unroll
        for (int gl = 0; gl < NT; ++gl) {
            const int t = yr - 2 * gl;
            if (t >= 0 && t <= SY && G0 + gl < NG) {
#pragma unroll
                for (int cc = 0; cc < NCOL; ++cc)
                    acc[cc][gl] = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[cc], bfr[t], acc[cc][gl], 0, 0, 0);
            }
        }
    }
}

template <int SY, int VY, int NC>
__global__ __launch_bounds__(SK_NT, 2) void k_zncc_map_f32_skew(NccArgs a, Bf16Geom q, int nzt) {
    using S = SkewShape<SY, VY, NC>;
    using F = SkewShapeF32<SY, VY, NC>;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t node = blockIdx.x / nzt;
    const int oz0 = (int)(blockIdx.x % nzt) * SK_MZ;
    const int nzp = min(SK_MZ, q.vz - oz0);
    const int64_t plane_out = (int64_t)q.vy * q.vx;
    float* out = a.out + node * plane_out * q.vz + (int64_t)oz0 * plane_out;
    double mf, ssf;
    if (!ncc_node_stats<SK_NT>(a, q, node, mf, ssf)) {
        for (int64_t i = tid; i < nzp * plane_out; i += SK_NT) out[i] = -2.0f;
        return;
    }
    const int64_t x0 = a.coords[3 * node], y0 = a.coords[3 * node + 1], z0 = a.coords[3 * node + 2];
    const int64_t xb = a.wx0[node], yb = a.wy0[node], zb = a.wz0[node];
    const int64_t tbase = ((z0 - a.hz) * a.ny + (y0 - a.hy)) * (int64_t)a.nx + (x0 - a.hx);
    const float cw = (float)mf;
    float* win = reinterpret_cast<float*>(lds);  // [buffer][WROWS][SKF_PITCH]
    float* tring = win + 2 * F::WIN;             // [slot][row -1 .. SY][SK_TP]; slot SK_RING stays zero
    for (int i = tid; i < 2 * F::WIN + (SK_RING + 1) * F::SLOT; i += SK_NT) win[i] = 0.0f;
    __syncthreads();
    const int wpl = q.wy * q.wx, tpl_n = SY * q.sx;
    const float inv_wx = 1.0f / (float)q.wx, inv_sx = 1.0f / (float)q.sx;  // exact floors for i < 2^22
    float pre[SK_PRE], tpre[SK_TPRE];
    auto load_plane = [&](int z) {
        const float* gp = a.g + ((zb + z) * a.ny + yb) * (int64_t)a.nx + xb;
#pragma unroll
        for (int u = 0; u < SK_PRE; ++u) {
            const int i = tid + SK_NT * u;
            if (i < wpl) {
                const int y = (int)(((float)i + 0.5f) * inv_wx), x = i - y * q.wx;
                pre[u] = gp[(int64_t)y * a.nx + x];
            }
        }
        const int tz = z - oz0;
        if (tz < q.sz) {
            const float* fp = a.f + tbase + (int64_t)tz * a.ny * a.nx;
#pragma unroll
            for (int u = 0; u < SK_TPRE; ++u) {
                const int i = tid + SK_NT * u;
                if (i < tpl_n) {
                    const int y = (int)(((float)i + 0.5f) * inv_sx), x = i - y * q.sx;
                    tpre[u] = fp[(int64_t)y * a.nx + x];
                }
            }
        }
    };
    auto store_plane = [&](int z) {
        float* w0 = win + (z & 1) * F::WIN;
#pragma unroll
        for (int u = 0; u < SK_PRE; ++u) {
            const int i = tid + SK_NT * u;
            if (i < wpl) {
                const int y = (int)(((float)i + 0.5f) * inv_wx), x = i - y * q.wx;
                w0[y * SKF_PITCH + x] = pre[u] - cw;
            }
        }
        const int tz = z - oz0;
        if (tz < q.sz) {
            float* tp = tring + (tz % SK_RING) * F::SLOT + SK_TP;  // row 0 of the slot
#pragma unroll
            for (int u = 0; u < SK_TPRE; ++u) {
                const int i = tid + SK_NT * u;
                if (i < tpl_n) {
                    const int y = (int)(((float)i + 0.5f) * inv_sx), x = i - y * q.sx;
                    tp[y * SK_TP + x] = (float)((double)tpre[u] - mf);
                }
            }
        }
    };
    const int n = lane & 15, j = n >> 1, d = n & 1, kk = lane >> 4, m = lane & 15;
    const int z_end = min(oz0 + nzp - 1 + q.sz, q.wz);
    const int nslices = (q.sx + 3) / 4;
    const float sq_ssf = (float)sqrt(ssf);
    auto run = [&](auto& acc, auto g0_tag, int c0) {
        constexpr int G0 = decltype(g0_tag)::value;
        constexpr int NCOL = sizeof(acc) / sizeof(acc[0]);
        constexpr int NT = sizeof(acc[0]) / sizeof(f32x4_t);
#pragma unroll
        for (int cc = 0; cc < NCOL; ++cc)
#pragma unroll
            for (int gl = 0; gl < NT; ++gl) acc[cc][gl] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
        load_plane(oz0);
        store_plane(oz0);
        __syncthreads();
        for (int z = oz0; z < z_end; ++z) {
            const bool more = z + 1 < z_end;
            if (more) load_plane(z + 1);
            const int tz = z - oz0 - j;
            const bool zok = j < nzp && tz >= 0 && tz < q.sz;
            const float* tp = tring + (zok ? tz % SK_RING : SK_RING) * F::SLOT + (1 - d) * SK_TP + kk;
            const float* ap = win + (z & 1) * F::WIN + 16 * c0 + m + kk;
#pragma unroll 1
            for (int sl = 0; sl < nslices; ++sl) {
                float bfr[SY + 1];
#pragma unroll
                for (int t = 0; t <= SY; ++t) bfr[t] = tp[t * SK_TP + 4 * sl];
                skew_plane_f32<SY, S::NG, G0, NT, NCOL>(acc, bfr, ap + 4 * sl);
            }
            if (more) store_plane(z + 1);
            __syncthreads();
        }
#pragma unroll
        for (int cc = 0; cc < NCOL; ++cc)
#pragma unroll
            for (int gl = 0; gl < NT; ++gl) {
                const int oy = 2 * (G0 + gl) + d;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int ox = 16 * (c0 + cc) + 4 * (lane >> 4) + i;
                    if (j < nzp && oy < q.vy && ox < q.vx) {
                        float* o = out + (int64_t)j * plane_out + (int64_t)oy * q.vx + ox;
                        const float var = *o;  // written by k_box_var
                        float val = 0.0f;
                        if (var >= 1e-20f) {
                            val = acc[cc][gl][i] / (sqrtf(var) * sq_ssf);
                            val = val > 1.0f ? 1.0f : (val < -1.0f ? -1.0f : val);
                        }
                        *o = val;
                    }
                }
            }
    };
    if (wave < NC) {
        f32x4_t acc[1][S::GA];
        run(acc, std::integral_constant<int, 0>{}, wave);
    } else if (wave == NC) {
        f32x4_t acc[NC][S::NG - S::GA];
        run(acc, std::integral_constant<int, S::GA>{}, 0);
    } else {
        load_plane(oz0);
        store_plane(oz0);
        __syncthreads();
        for (int z = oz0; z < z_end; ++z) {
            if (z + 1 < z_end) {
                load_plane(z + 1);
                store_plane(z + 1);
            }
            __syncthreads();
        }
    }
}

template <int SY, int VY, int NC>
size_t skew_f32_lds_bytes() {
    using F = SkewShapeF32<SY, VY, NC>;
    return sizeof(float) * (2 * (size_t)F::WIN + (size_t)(SK_RING + 1) * F::SLOT);
}

template <int SY, int VY, int NC>
size_t skew_lds_bytes() {
    return sizeof(__bf16) * (4 * (size_t)SkewShape<SY, VY, NC>::COPY + (size_t)(SK_RING + 1) * sk_slot_elems(SY));
}

size_t bf16_lds_bytes(const Bf16Geom& q) {
    return sizeof(__bf16) * (2 * (size_t)q.AH * q.AW + 2 * (size_t)TRING * q.sy * q.TW);
}

size_t box_lds_bytes(const Bf16Geom& q) {
    return sizeof(float) * (2 * (size_t)q.wy * q.wx + 2 * (size_t)q.wy * q.vx);
}

// k_box_var over n nodes with the smallest instance whose per-thread arrays hold a window plane
// (the caller has checked wy * wx <= BLOCK * VPL).
int launch_box_var(const NccArgs& a, const Bf16Geom& q, int64_t n, hipStream_t stream) {
    const size_t blds = box_lds_bytes(q);
    const int64_t wpl = (int64_t)q.wy * q.wx;
    auto go = [&](auto kern) -> int {
        if (blds > 64 * 1024 &&
            hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)blds) != hipSuccess)
            return fail(-5, "could not raise the LDS limit");
        hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(BLOCK), blds, stream, a, q);
        return 0;
    };
    if (wpl <= (int64_t)BLOCK * 4) return go(k_box_var<4>);
    if (wpl <= (int64_t)BLOCK * 8) return go(k_box_var<8>);
    if (wpl <= (int64_t)BLOCK * 13) return go(k_box_var<13>);
    return go(k_box_var<VPL>);
}

Bf16Geom bf16_geom(int hx, int hy, int hz, int rx, int ry, int rz, int& ks) {
    Bf16Geom q;
    q.sx = 2 * hx + 1;
    q.sy = 2 * hy + 1;
    q.sz = 2 * hz + 1;
    q.vx = 2 * rx + 1;
    q.vy = 2 * ry + 1;
    q.vz = 2 * rz + 1;
    q.wx = q.sx + q.vx - 1;
    q.wy = q.sy + q.vy - 1;
    q.wz = q.sz + q.vz - 1;
    q.nyt = (q.vy + 15) / 16;
    q.nxt = (q.vx + 15) / 16;
    ks = (q.sx + 15 + 31) / 32;
    q.AH = 16 * q.nyt + q.sy - 1;
    q.AW = 16 * (q.nxt - 1) + 32 * ks;
    // row stride = 8 mod 16 elements: the 16 rows an A fragment reads start on 16 different
    // 4-bank groups
    q.AW = (q.AW + 7) / 8 * 8;
    if (q.AW % 16 == 0) q.AW += 8;
    q.TW = 32 * ks + 16;
    return q;
}

// ---------------------------------------------------------------------------- peak epilogue
// Per node (integer_search.py:444-477): the first index of the maximum in C order (a NaN wins,
// as in np.argmax), the minimum, PCE = (peak - min)^2 / mean((C - min)^2) (0 when the energy is
// not above 1e-30), and, for a peak off the search border with every radius > 0, the quadratic
// fit of its 27-neighbourhood (pinv of the fixed 27 x 10 design, integer_search.py:210-227,
// :297-321). Map value -2 (flat template) reads as 0, as in ncc_search (:420-421).
__device__ inline float map_val(float v) { return v < -1.0f ? 0.0f : v; }

__device__ inline bool better(float a, int64_t ia, float b, int64_t ib) {
    const bool na = a != a, nb = b != b;
    if (na || nb) return na && (!nb || ia < ib);
    return a > b || (a == b && ia < ib);
}

__global__ __launch_bounds__(BLOCK) void k_ncc_peaks(const float* __restrict__ maps, int64_t n, int rx, int ry, int rz,
                                                     const double* __restrict__ pinv, int64_t* peak_idx,
                                                     double* peak_val, double* pce, double* sub) {
    const int64_t node = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int vx = 2 * rx + 1, vy = 2 * ry + 1, vz = 2 * rz + 1;
    const int64_t n_off = (int64_t)vx * vy * vz;
    const float* m = maps + node * n_off;
    __shared__ float wbest[NWAVES], wmin[NWAVES];
    __shared__ int64_t wbi[NWAVES];
    __shared__ double wen[NWAVES];
    __shared__ float sbest, smin;
    __shared__ int64_t sbi;
    float best = -INFINITY, mn = INFINITY;
    int64_t bi = INT64_MAX;
    bool any_nan = false;
    for (int64_t i = tid; i < n_off; i += BLOCK) {
        const float v = map_val(m[i]);
        if (bi == INT64_MAX || better(v, i, best, bi)) {
            best = v;
            bi = i;
        }
        if (v != v) any_nan = true;
        mn = v < mn ? v : mn;
    }
    if (any_nan) mn = NAN;
    for (int off = 32; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off, 64);
        const int64_t oi = __shfl_xor(bi, off, 64);
        const float om = __shfl_xor(mn, off, 64);
        if (oi != INT64_MAX && (bi == INT64_MAX || better(ob, oi, best, bi))) {
            best = ob;
            bi = oi;
        }
        mn = (om != om || mn != mn) ? NAN : (om < mn ? om : mn);
    }
    if (lane == 0) {
        wbest[wave] = best;
        wbi[wave] = bi;
        wmin[wave] = mn;
    }
    __syncthreads();
    if (tid == 0) {
        float b = wbest[0], mm = wmin[0];
        int64_t ii = wbi[0];
        for (int w = 1; w < NWAVES; ++w) {
            if (wbi[w] != INT64_MAX && (ii == INT64_MAX || better(wbest[w], wbi[w], b, ii))) {
                b = wbest[w];
                ii = wbi[w];
            }
            mm = (wmin[w] != wmin[w] || mm != mm) ? NAN : (wmin[w] < mm ? wmin[w] : mm);
        }
        sbest = b;
        sbi = ii;
        smin = mm;
    }
    __syncthreads();
    const double cmin = (double)smin;
    double en = 0.0;
    for (int64_t i = tid; i < n_off; i += BLOCK) {
        const double d = (double)map_val(m[i]) - cmin;
        en += d * d;
    }
    en = wave_sum(en);
    if (lane == 0) wen[wave] = en;
    __syncthreads();
    if (tid != 0) return;
    double energy = 0.0;
    for (int w = 0; w < NWAVES; ++w) energy += wen[w];
    energy /= (double)n_off;
    const double pk = (double)sbest;
    const double pkm = pk - cmin;
    pce[node] = energy > 1e-30 ? pkm * pkm / energy : 0.0;
    peak_idx[node] = sbi;
    peak_val[node] = pk;
    const int64_t b = sbi;
    const int pz = (int)(b / ((int64_t)vy * vx)), rem = (int)(b - (int64_t)pz * vy * vx);
    const int py = rem / vx, px = rem - py * vx;
    double sx_ = 0.0, sy_ = 0.0, sz_ = 0.0;
    const bool border = px == 0 || px == 2 * rx || py == 0 || py == 2 * ry || pz == 0 || pz == 2 * rz;
    if (!border && rx > 0 && ry > 0 && rz > 0) {
        double coef[10];
        for (int q = 0; q < 10; ++q) coef[q] = 0.0;
        int t = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx, ++t) {
                    const double v = (double)map_val(m[((int64_t)(pz + dz) * vy + (py + dy)) * vx + (px + dx)]);
                    for (int q = 0; q < 10; ++q) coef[q] += pinv[q * 27 + t] * v;
                }
        // H x = -a with H = [[2c7, c4, c5], [c4, 2c8, c6], [c5, c6, 2c9]], a = (c1, c2, c3)
        const double h00 = 2.0 * coef[7], h11 = 2.0 * coef[8], h22 = 2.0 * coef[9];
        const double h01 = coef[4], h02 = coef[5], h12 = coef[6];
        const double det = h00 * (h11 * h22 - h12 * h12) - h01 * (h01 * h22 - h12 * h02) + h02 * (h01 * h12 - h11 * h02);
        if (fabs(det) > 1e-12) {
            const double r0 = -coef[1], r1 = -coef[2], r2 = -coef[3];
            const double x = (r0 * (h11 * h22 - h12 * h12) - h01 * (r1 * h22 - h12 * r2) + h02 * (r1 * h12 - h11 * r2)) / det;
            const double y = (h00 * (r1 * h22 - h12 * r2) - r0 * (h01 * h22 - h12 * h02) + h02 * (h01 * r2 - r1 * h02)) / det;
            const double z = (h00 * (h11 * r2 - r1 * h12) - h01 * (h01 * r2 - r1 * h02) + r0 * (h01 * h12 - h11 * h02)) / det;
            if (isfinite(x) && isfinite(y) && isfinite(z) && fabs(x) <= 1.0 && fabs(y) <= 1.0 && fabs(z) <= 1.0) {
                sx_ = x;
                sy_ = y;
                sz_ = z;
            }
        }
    }
    sub[3 * node] = sx_;
    sub[3 * node + 1] = sy_;
    sub[3 * node + 2] = sz_;
}

}  // namespace

extern "C" {

size_t aldvc_zncc_map_workspace_bytes(int64_t n, int hx, int hy, int hz, int rx, int ry, int rz, int precision) {
    (void)n; (void)hx; (void)hy; (void)hz; (void)rx; (void)ry; (void)rz; (void)precision;
    return 0;
}

int aldvc_zncc_map(const float* f, const float* g, int64_t nz, int64_t ny, int64_t nx, const int64_t* coords,
                   int64_t n, int hx, int hy, int hz, int rx, int ry, int rz, const int64_t* wx0, const int64_t* wy0,
                   const int64_t* wz0, const uint8_t* t_ok, float* out, int precision, void* workspace,
                   size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    (void)workspace;
    (void)workspace_bytes;
    if (n < 0) return fail(-2, "negative node count");
    if (n == 0) return 0;
    if (!f || !g || !coords || !wx0 || !wy0 || !wz0 || !t_ok || !out) return fail(-2, "null pointer");
    if (hx < 0 || hy < 0 || hz < 0 || rx < 0 || ry < 0 || rz < 0) return fail(-2, "negative half-width or radius");
    const int64_t sx = 2 * hx + 1, sy = 2 * hy + 1, sz = 2 * hz + 1;
    if (sx + 2 * rx > nx || sy + 2 * ry > ny || sz + 2 * rz > nz) return fail(-2, "search window larger than the volume");
    const int64_t n_tpl = sx * sy * sz;
    if (n_tpl > LDS_FLOATS_MAX) return fail(-2, "template does not fit in LDS (at most ~40k voxels)");
    if (n > INT32_MAX) return fail(-2, "too many nodes for one launch");
    if (precision != 0 && precision != 1) return fail(-2, "precision must be 0 (float32) or 1 (bfloat16 MFMA)");
    NccArgs a{f, g, (int)nz, (int)ny, (int)nx, coords, n, hx, hy, hz, rx, ry, rz, wx0, wy0, wz0, t_ok, out};
    {
        // the skewed-plane tilings (both precisions) where their geometry is instantiated: templates
        // up to 32 voxels wide, up to 48 x offsets, planes within the per-lane prefetch and box sums;
        // the window variance comes first, from k_box_var
        int ks = 0;
        const Bf16Geom q = bf16_geom(hx, hy, hz, rx, ry, rz, ks);
        const bool fits = q.sx <= SK_TP && (int64_t)q.wy * q.wx <= (int64_t)SK_NT * SK_PRE &&
                          (int64_t)q.sy * q.sx <= (int64_t)SK_NT * SK_TPRE && (int64_t)q.vy * q.vx <= (int64_t)BLOCK * VPT &&
                          (int64_t)q.wy * q.wx <= (int64_t)BLOCK * VPL;
        const bool known = (q.sy == 25 && q.vy == 33 && q.nxt == 3) || (q.sy == 13 && q.vy == 17 && q.nxt == 2);
        if (fits && known) {
            const int nzt = (q.vz + SK_MZ - 1) / SK_MZ;
            if (n * nzt > INT32_MAX) return fail(-2, "too many workgroups for one launch");
            if (int e = launch_box_var(a, q, n, stream)) return e;
#define ALDVC_SKEW(KERN, LDSF, SYV, VYV, NCV)                                                               \
    if (q.sy == SYV && q.vy == VYV && q.nxt == NCV) {                                                      \
        const size_t sl = LDSF<SYV, VYV, NCV>();                                                            \
        const void* sf = (const void*)KERN<SYV, VYV, NCV>;                                                  \
        if (sl > 64 * 1024 && hipFuncSetAttribute(sf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sl) != hipSuccess) \
            return fail(-5, "could not raise the LDS limit");                                              \
        hipLaunchKernelGGL((KERN<SYV, VYV, NCV>), dim3((unsigned)(n * nzt)), dim3(SK_NT), sl, stream, a, q, nzt); \
        return launch_status();                                                                            \
    }
            if (precision == 1) {
                ALDVC_SKEW(k_zncc_map_bf16_skew, skew_lds_bytes, 25, 33, 3)
                ALDVC_SKEW(k_zncc_map_bf16_skew, skew_lds_bytes, 13, 17, 2)
            } else {
                ALDVC_SKEW(k_zncc_map_f32_skew, skew_f32_lds_bytes, 25, 33, 3)
                ALDVC_SKEW(k_zncc_map_f32_skew, skew_f32_lds_bytes, 13, 17, 2)
            }
#undef ALDVC_SKEW
            return fail(-2, "skewed-plane shape not instantiated");
        }
    }
    if (precision == 1) {
        int ks = 0;
        const Bf16Geom q = bf16_geom(hx, hy, hz, rx, ry, rz, ks);
        const int nt = std::max(q.nyt, q.nxt);
        if (nt > 3) return fail(-2, "bfloat16 maps support search radii up to 23 voxels in x and y");
        if (ks > 2) return fail(-2, "bfloat16 maps support templates up to 49 voxels along x");
        const size_t lds = bf16_lds_bytes(q);
        if (lds > 160 * 1024 - 256) return fail(-2, "template and window planes do not fit in LDS");
        if ((int64_t)q.wy * q.wx > (int64_t)BLOCK * PRE) return fail(-2, "window plane too large for the prefetch");
        if ((int64_t)q.sy * q.sx > (int64_t)BLOCK * 4) return fail(-2, "template plane too large for the prefetch");
        if ((int64_t)q.vy * q.vx > (int64_t)BLOCK * VPT) return fail(-2, "search plane too large for the box sums");
        if ((int64_t)q.wy * q.wx > (int64_t)BLOCK * VPL) return fail(-2, "window plane too large for the box sums");
        const int64_t nblk = n * ((q.vz + MG - 1) / MG);
        const int64_t nbox = n;
        if (nblk > INT32_MAX) return fail(-2, "too many workgroups for one launch");
        if (int e = launch_box_var(a, q, nbox, stream)) return e;
        const void* fn = nullptr;
#define ALDVC_BF16(NTV, KSV)                                                                                \
    if (nt <= NTV && ks == KSV) {                                                                          \
        fn = (const void*)k_zncc_map_bf16<NTV, KSV>;                                                       \
        if (lds > 64 * 1024 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) \
            return fail(-5, "could not raise the LDS limit");                                              \
        hipLaunchKernelGGL((k_zncc_map_bf16<NTV, KSV>), dim3((unsigned)nblk), dim3(256), lds, stream, a, q); \
        return launch_status();                                                                            \
    }
        ALDVC_BF16(1, 1)
        ALDVC_BF16(1, 2)
        ALDVC_BF16(2, 1)
        ALDVC_BF16(2, 2)
        ALDVC_BF16(3, 1)
        ALDVC_BF16(3, 2)
#undef ALDVC_BF16
        return fail(-2, "unsupported bfloat16 geometry");
    }
    {
        const int vx = 2 * rx + 1;
        const size_t slds = sizeof(float) * (size_t)small_lds_floats((int)sx, (int)sy, (int)sz, vx, 2 * ry + 1, 2 * rz + 1);
        // two workgroups per CU at the most (LDS): 512 lanes each when the (template plane, output
        // row) items fit one round of them, 256 otherwise
        const bool wide = sz * (2 * rz + 1) * (2 * ry + 1) <= 512 && sz * (2 * rz + 1) * (2 * ry + 1) > 256;
#define ALDVC_SMALL_NT(RXV, NTV, SXV)                                                                     \
        {                                                                                                    \
            const void* fn = (const void*)k_zncc_map_small<RXV, NTV, SXV>;                                   \
            if (slds > 64 * 1024 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)slds) != hipSuccess) \
                return fail(-5, "could not raise the LDS limit for the window");                             \
            hipLaunchKernelGGL((k_zncc_map_small<RXV, NTV, SXV>), dim3((unsigned)n), dim3(NTV), slds, stream, a); \
            return launch_status();                                                                          \
        }
#define ALDVC_SMALL(RXV)                                                                                  \
        if (vx == RXV && slds <= SMALL_LDS_MAX) {                                                             \
            if (wide) ALDVC_SMALL_NT(RXV, 512, 0)                                                             \
            ALDVC_SMALL_NT(RXV, 256, 0)                                                                       \
        }
        // the pyramid's templates (16^3 at full resolution, 8^3 on the coarse levels: 17 and 9 wide)
        // with its radii (2, doubled to 4)
#define ALDVC_SMALL_SX(RXV, SXV)                                                                          \
        if (vx == RXV && sx == SXV && slds <= SMALL_LDS_MAX) {                                                \
            if (wide) ALDVC_SMALL_NT(RXV, 512, SXV)                                                           \
            ALDVC_SMALL_NT(RXV, 256, SXV)                                                                     \
        }
        ALDVC_SMALL_SX(5, 17)
        ALDVC_SMALL_SX(5, 9)
        ALDVC_SMALL_SX(9, 17)
        ALDVC_SMALL_SX(9, 9)
        ALDVC_SMALL(3)
        ALDVC_SMALL(5)
        ALDVC_SMALL(7)
        ALDVC_SMALL(9)
#undef ALDVC_SMALL_SX
#undef ALDVC_SMALL
#undef ALDVC_SMALL_NT
    }
    const size_t lds = sizeof(float) * (size_t)n_tpl;
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void*)k_zncc_map_f32, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return fail(-5, "could not raise the LDS limit for the template");
    hipLaunchKernelGGL(k_zncc_map_f32, dim3((unsigned)n), dim3(BLOCK), sizeof(float) * n_tpl, stream, a);
    return launch_status();
}

int aldvc_ncc_peaks(const float* maps, int64_t n, int rx, int ry, int rz, const double* pinv, int64_t* peak_idx,
                    double* peak_val, double* pce, double* sub, hipStream_t stream) {
    clear_error();
    if (n < 0) return fail(-2, "negative node count");
    if (n == 0) return 0;
    if (!maps || !pinv || !peak_idx || !peak_val || !pce || !sub) return fail(-2, "null pointer");
    if (rx < 0 || ry < 0 || rz < 0) return fail(-2, "negative radius");
    if (n > INT32_MAX) return fail(-2, "too many nodes for one launch");
    hipLaunchKernelGGL(k_ncc_peaks, dim3((unsigned)n), dim3(BLOCK), 0, stream, maps, n, rx, ry, rz, pinv, peak_idx,
                       peak_val, pce, sub);
    return launch_status();
}

}  // extern "C"
