// Device-side building blocks of the local DVC step for gfx950 (CDNA4, wave64).
//
// Conventions (reference al_dvc 1.2.1, core/data_structures.py:1-11, numba_kernels.py:9-16):
//   volumes (nz, ny, nx) C order float32; coords (N, 3) int64 [x, y, z];
//   P (N, 12) float64 = [F row-major (9), u, v, w], x_def = x0 + P[9:12] + (I + F)(x - x0);
//   SD = [gx(X,Y,Z), gy(X,Y,Z), gz(X,Y,Z), gx, gy, gz].
// Subset voxels are visited in the reference's sampling order (dz, dy, dx with the stride),
// voxel index v = (iz * sy + iy) * sx + ix, which is also the bit order of split keep rows.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aldvc {

// status codes: numba_kernels.py:29-36
enum Status : int {
    ST_CONVERGED = 0,
    ST_MAX_ITER = 1,
    ST_OUT_OF_BOUNDS = 2,
    ST_INVALID_SUBSET = 3,
    ST_SINGULAR = 4,
    ST_NAN = 5,
    ST_SKIPPED = 6,
    ST_STALLED = 7,
};

// numba_kernels.py:38-44, :81-85
constexpr double ABS_TOL = 1e-5;
constexpr double STALL_ZNCC_EPS = 1e-4;
constexpr double STALL_STEP_DECAY = 0.999;
constexpr int MIN_SUBSET_VOXELS = 27;
constexpr double MIN_VALID_FRACTION = 0.5;
constexpr double LM_DAMPING_3DOF = 1e-3;
constexpr double MIN_CORRECTED_FRACTION = 0.5;
constexpr double NOISE_CORR_STEP = 0.5;
constexpr double PREDICT_CONTRACTION = 0.5;

// interp_kernels.py:23-31
enum InterpMode : int { INTERP_LINEAR = 0, INTERP_CUBIC = 1, INTERP_BSPLINE = 2 };
constexpr float SAMPLE_LO = 1.0f;
constexpr float SAMPLE_HI_MARGIN = 2.0f;

// io/volume_ops.py:25 (STENCIL7), numba_kernels.py:194-196
constexpr float C1 = 0.75f, C2 = -0.15f, C3 = 1.0f / 60.0f;

struct Volumes {
    const float* f;   // reference
    const float* gx;  // stored gradients, or nullptr for the on-the-fly stencil
    const float* gy;
    const float* gz;
    const float* g;   // deformed volume as prepared for the interpolation mode (NaN = masked)
    const uint8_t* mask;  // reference mask, or nullptr when every voxel counts
    int nz, ny, nx;
};

struct Subset {
    int hx, hy, hz, stride;
    int sx, sy, sz;  // sampled voxels per axis
    int S;           // sampled voxels per subset
};

__host__ __device__ inline Subset make_subset(int hx, int hy, int hz, int stride) {
    Subset s;
    s.hx = hx; s.hy = hy; s.hz = hz; s.stride = stride;
    s.sx = (2 * hx) / stride + 1;
    s.sy = (2 * hy) / stride + 1;
    s.sz = (2 * hz) / stride + 1;
    s.S = s.sx * s.sy * s.sz;
    return s;
}

__device__ inline int64_t lin(const Volumes& V, int64_t z, int64_t y, int64_t x) {
    return (z * V.ny + y) * V.nx + x;
}

// ---------------------------------------------------------------------------- interpolation
// Keys cubic convolution a = -0.5 (interp_kernels.py:34-41), cubic B-spline basis (:44-53).
__device__ inline void keys4(float t, float& w0, float& w1, float& w2, float& w3) {
    w0 = ((-0.5f * t + 1.0f) * t - 0.5f) * t;
    w1 = (1.5f * t - 2.5f) * t * t + 1.0f;
    w2 = ((-1.5f * t + 2.0f) * t + 0.5f) * t;
    w3 = (0.5f * t - 0.5f) * t * t;
}

// Keys weights times two (W_k = 2 w_k) from t^2 and a = t^3 - t^2: eight operations instead of
// twelve; the factor 2 per axis is taken out of the tricubic sum once (a power of two, exact).
__device__ inline void keys4x2(float t, float& w0, float& w1, float& w2, float& w3) {
    const float t2 = t * t;
    const float a = __fmaf_rn(t2, t, -t2);                // t^3 - t^2
    w3 = a;                                               // 2 w3 = t^3 - t^2
    w0 = (t2 - t) - a;                                    // 2 w0 = -t^3 + 2 t^2 - t
    w1 = __fmaf_rn(3.0f, a, __fmaf_rn(-2.0f, t2, 2.0f));  // 2 w1 = 3 t^3 - 5 t^2 + 2
    w2 = __fmaf_rn(-3.0f, a, t2 + t);                     // 2 w2 = -3 t^3 + 4 t^2 + t
}

__device__ inline void bspline4(float t, float& w0, float& w1, float& w2, float& w3) {
    const float t2 = t * t, t3 = t2 * t;
    const float s = 1.0f / 6.0f;
    w0 = (1.0f - 3.0f * t + 3.0f * t2 - t3) * s;
    w1 = (4.0f - 6.0f * t2 + 3.0f * t3) * s;
    w2 = (1.0f + 3.0f * t + 3.0f * t2 - 3.0f * t3) * s;
    w3 = t3 * s;
}

// Four consecutive floats from a 4-byte aligned address (one dwordx4 load on gfx9 global memory).
__device__ inline float4 load4(const float* p) {
    float4 v;
    __builtin_memcpy(&v, p, sizeof(float4));
    return v;
}

// Sample the volume at a position given relative to an integer anchor (ax, ay, az), with the
// relative coordinates (rx, ry, rz) small enough that float32 keeps ~1e-6 voxel resolution.
// `base` points at voxel (az, ay, ax). The caller has checked the admissible domain
// 1 <= c <= n - 3 (interp_kernels.py:127-136); the stencil base is clamped to n - 3 (:89-94),
// expressed here as the relative limits (kx, ky, kz) = (nx - 3 - ax, ...).
template <int MODE>
__device__ inline float sample_rel(const float* __restrict__ base, int rowp, int planep, float rx, float ry,
                                   float rz, int kx, int ky, int kz) {
    int ix = (int)floorf(rx), iy = (int)floorf(ry), iz = (int)floorf(rz);
    ix = ix > kx ? kx : ix;
    iy = iy > ky ? ky : iy;
    iz = iz > kz ? kz : iz;
    const float fx = rx - (float)ix, fy = ry - (float)iy, fz = rz - (float)iz;
    if constexpr (MODE == INTERP_LINEAR) {
        const float* p = base + (iz * planep + iy * rowp + ix);
        const float c00 = p[0] * (1.0f - fx) + p[1] * fx;
        const float c10 = p[rowp] * (1.0f - fx) + p[rowp + 1] * fx;
        const float c01 = p[planep] * (1.0f - fx) + p[planep + 1] * fx;
        const float c11 = p[planep + rowp] * (1.0f - fx) + p[planep + rowp + 1] * fx;
        const float c0 = c00 * (1.0f - fy) + c10 * fy;
        const float c1 = c01 * (1.0f - fy) + c11 * fy;
        return c0 * (1.0f - fz) + c1 * fz;
    } else {
        float wx0, wx1, wx2, wx3, wy0, wy1, wy2, wy3, wz0, wz1, wz2, wz3;
        if constexpr (MODE == INTERP_CUBIC) {
            keys4(fx, wx0, wx1, wx2, wx3);
            keys4(fy, wy0, wy1, wy2, wy3);
            keys4(fz, wz0, wz1, wz2, wz3);
        } else {
            bspline4(fx, wx0, wx1, wx2, wx3);
            bspline4(fy, wy0, wy1, wy2, wy3);
            bspline4(fz, wz0, wz1, wz2, wz3);
        }
        const float* p = base + ((iz - 1) * planep + (iy - 1) * rowp + (ix - 1));
        // all 16 row loads are issued before any of them is consumed: a gather that waits for
        // each row in turn is bound by the load latency, not by the memory pipeline
        float4 r[16];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) r[4 * k + j] = load4(p + k * planep + j * rowp);
        const float wz[4] = {wz0, wz1, wz2, wz3};
        const float wy[4] = {wy0, wy1, wy2, wy3};
        float val = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float plane = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float4 q = r[4 * k + j];
                plane += wy[j] * (q.x * wx0 + q.y * wx1 + q.z * wx2 + q.w * wx3);
            }
            val += wz[k] * plane;
        }
        return val;
    }
}

// Buffer-resource sampler: the volume is addressed through a buffer descriptor, so one 32-bit
// VGPR byte offset serves all 16 row loads of a sample and the row offsets (k * plane + j * row)
// are wave-uniform SGPR offsets. Requires the volume to span fewer than 2^32 bytes. `anchor` is
// the linear index of the anchor voxel; everything else as sample_rel.
constexpr unsigned BUF_DWORD3 = 0x00020000u;  // gfx9 raw-buffer descriptor word 3 (32-bit data)

struct BufVol {
    __amdgpu_buffer_rsrc_t rs;
    int rowb, planeb;  // row / plane strides in bytes
};

__device__ inline BufVol make_bufvol(const float* vol, int64_t nvox, int rowp, int planep) {
    BufVol b;
    b.rs = __builtin_amdgcn_make_buffer_rsrc((void*)vol, (short)0, (int)(uint32_t)(nvox * 4), (int)BUF_DWORD3);
    b.rowb = 4 * rowp;
    b.planeb = 4 * planep;
    return b;
}

__device__ inline float4 buf_load4(const BufVol& b, uint32_t voff, int soff) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(b.rs, voff, soff, 0);
    float4 f;
    __builtin_memcpy(&f, &v, sizeof(f));
    return f;
}

template <int MODE>
__device__ inline float sample_buf(const BufVol& b, int anchor, int rowp, int planep, float rx, float ry, float rz,
                                   int kx, int ky, int kz) {
    int ix = (int)floorf(rx), iy = (int)floorf(ry), iz = (int)floorf(rz);
    ix = ix > kx ? kx : ix;
    iy = iy > ky ? ky : iy;
    iz = iz > kz ? kz : iz;
    const float fx = rx - (float)ix, fy = ry - (float)iy, fz = rz - (float)iz;
    if constexpr (MODE == INTERP_LINEAR) {
        const uint32_t voff = 4u * (uint32_t)(anchor + iz * planep + iy * rowp + ix);
        const float2 r00 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(b.rs, voff, 0, 0));
        const float2 r10 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(b.rs, voff, b.rowb, 0));
        const float2 r01 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(b.rs, voff, b.planeb, 0));
        const float2 r11 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(b.rs, voff, b.planeb + b.rowb, 0));
        const float c00 = r00.x * (1.0f - fx) + r00.y * fx;
        const float c10 = r10.x * (1.0f - fx) + r10.y * fx;
        const float c01 = r01.x * (1.0f - fx) + r01.y * fx;
        const float c11 = r11.x * (1.0f - fx) + r11.y * fx;
        const float c0 = c00 * (1.0f - fy) + c10 * fy;
        const float c1 = c01 * (1.0f - fy) + c11 * fy;
        return c0 * (1.0f - fz) + c1 * fz;
    } else {
        float wx0, wx1, wx2, wx3, wy0, wy1, wy2, wy3, wz0, wz1, wz2, wz3;
        if constexpr (MODE == INTERP_CUBIC) {
            keys4(fx, wx0, wx1, wx2, wx3);
            keys4(fy, wy0, wy1, wy2, wy3);
            keys4(fz, wz0, wz1, wz2, wz3);
        } else {
            bspline4(fx, wx0, wx1, wx2, wx3);
            bspline4(fy, wy0, wy1, wy2, wy3);
            bspline4(fz, wz0, wz1, wz2, wz3);
        }
        const uint32_t voff = 4u * (uint32_t)(anchor + (iz - 1) * planep + (iy - 1) * rowp + (ix - 1));
        // the 16 rows in two batches of 8, each batch issued before any row is consumed: one
        // latency per batch, and 32 rather than 64 registers held, which leaves room for a third
        // workgroup per CU
        const float wz[4] = {wz0, wz1, wz2, wz3};
        const float wy[4] = {wy0, wy1, wy2, wy3};
        float val = 0.0f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float4 r[8];
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int j = 0; j < 4; ++j) r[4 * k + j] = buf_load4(b, voff, (2 * h + k) * b.planeb + j * b.rowb);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                float plane = 0.0f;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float4 q = r[4 * k + j];
                    plane += wy[j] * (q.x * wx0 + q.y * wx1 + q.z * wx2 + q.w * wx3);
                }
                val += wz[2 * h + k] * plane;
            }
        }
        return val;
    }
}

// LDS brick sampler. A brick is a block of the deformed volume staged in LDS with row stride BRK
// floats and plane stride `plane` floats (a run-time value: the compiler then forms one base per
// stencil plane, and every tap of a plane is a small immediate offset from it, which lets it pair
// the taps into 2-dword LDS reads). (rx, ry, rz) are anchor-relative positions as in sample_rel and (bx, by, bz) the
// anchor-relative brick origin. The caller guarantees that the position is inside the sampling
// domain (so the n - 3 clamp of the stencil base never applies) and that the brick holds its stencil.
template <int MODE, int BRK>
__device__ inline float sample_brick(const float* brick, int plane, int bx, int by, int bz, float rx, float ry,
                                     float rz) {
    constexpr int ROW = BRK;
    const int PLANE = plane;
    const float flx = floorf(rx), fly = floorf(ry), flz = floorf(rz);
    const float fx = rx - flx, fy = ry - fly, fz = rz - flz;
    // brick index of the first tap, formed in float from the floored coordinates (exact: brick
    // indices stay far below 2^24) and converted once; the origin term is uniform per chunk
    constexpr int LO = MODE == INTERP_LINEAR ? 0 : 1;
    const float org = -(float)((bz + LO) * PLANE + (by + LO) * ROW + (bx + LO));
    const int first = (int)__fmaf_rn(flz, (float)PLANE, __fmaf_rn(fly, (float)ROW, flx + org));
    if constexpr (MODE == INTERP_LINEAR) {
        const float* p = brick + first;
        const float c00 = p[0] * (1.0f - fx) + p[1] * fx;
        const float c10 = p[ROW] * (1.0f - fx) + p[ROW + 1] * fx;
        const float c01 = p[PLANE] * (1.0f - fx) + p[PLANE + 1] * fx;
        const float c11 = p[PLANE + ROW] * (1.0f - fx) + p[PLANE + ROW + 1] * fx;
        const float c0 = c00 * (1.0f - fy) + c10 * fy;
        const float c1 = c01 * (1.0f - fy) + c11 * fy;
        return c0 * (1.0f - fz) + c1 * fz;
    } else {
        float wx0, wx1, wx2, wx3, wy0, wy1, wy2, wy3, wz0, wz1, wz2, wz3;
        if constexpr (MODE == INTERP_CUBIC) {
            keys4x2(fx, wx0, wx1, wx2, wx3);
            keys4x2(fy, wy0, wy1, wy2, wy3);
            keys4x2(fz, wz0, wz1, wz2, wz3);
        } else {
            bspline4(fx, wx0, wx1, wx2, wx3);
            bspline4(fy, wy0, wy1, wy2, wy3);
            bspline4(fz, wz0, wz1, wz2, wz3);
        }
        const float* p = brick + first;
        const float* pk[4] = {p, p + PLANE, p + 2 * PLANE, p + 3 * PLANE};
        // the 64 taps in two batches of 32, each read before any of it is consumed (as sample_buf)
        const float wz[4] = {wz0, wz1, wz2, wz3};
        const float wy[4] = {wy0, wy1, wy2, wy3};
        float val = 0.0f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float t[32];
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i) t[16 * k + 4 * j + i] = pk[2 * h + k][j * ROW + i];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                float plane = 0.0f;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float* r = t + 16 * k + 4 * j;
                    plane += wy[j] * (r[0] * wx0 + r[1] * wx1 + r[2] * wx2 + r[3] * wx3);
                }
                val += wz[2 * h + k] * plane;
            }
        }
        if constexpr (MODE == INTERP_CUBIC) val *= 0.125f;  // the three doubled weight sets
        return val;
    }
}

// Anchor-relative warped position with a pinned evaluation order (explicit fused multiply-adds,
// X innermost, Z outermost): evaluated the same way everywhere, it is monotone in each of X, Y, Z,
// so over a box of subset voxels its extremes are taken at the box corners -- the bound the LDS
// bricks rely on. warp_xy is the part a voxel column of fixed (X, Y) shares; warp_coord =
// warp_z(a2, Z, warp_xy(a0, a1, t, X, Y)) to the bit.
__device__ inline float warp_xy(float a0, float a1, float t, float X, float Y) {
    return __fmaf_rn(a1, Y, __fmaf_rn(a0, X, t));
}
__device__ inline float warp_z(float a2, float Z, float xy) { return __fmaf_rn(a2, Z, xy); }
__device__ inline float warp_coord(float a0, float a1, float a2, float t, float X, float Y, float Z) {
    return warp_z(a2, Z, warp_xy(a0, a1, t, X, Y));
}

// Admissible sampling domain (interp_kernels.py:127-136) in anchor-relative coordinates:
// lo = 1 - a, hi = n - 3 - a per axis.
struct Domain {
    float lox, hix, loy, hiy, loz, hiz;
    int kx, ky, kz;
};

__device__ inline Domain make_domain(const Volumes& V, int ax, int ay, int az) {
    Domain d;
    d.lox = SAMPLE_LO - (float)ax;
    d.hix = (float)(V.nx - 1) - SAMPLE_HI_MARGIN - (float)ax;
    d.loy = SAMPLE_LO - (float)ay;
    d.hiy = (float)(V.ny - 1) - SAMPLE_HI_MARGIN - (float)ay;
    d.loz = SAMPLE_LO - (float)az;
    d.hiz = (float)(V.nz - 1) - SAMPLE_HI_MARGIN - (float)az;
    d.kx = V.nx - 3 - ax;
    d.ky = V.ny - 3 - ay;
    d.kz = V.nz - 3 - az;
    return d;
}

__device__ inline bool inside(const Domain& d, float rx, float ry, float rz) {
    return rx >= d.lox && rx <= d.hix && ry >= d.loy && ry <= d.hiy && rz >= d.loz && rz <= d.hiz;
}

// Reference gradient at a voxel: stored volumes, or the 7-point stencil on f, zero within 3
// voxels of a face (numba_kernels.py:188-212; cuda_kernels.py:343-379).
template <bool STORED>
__device__ inline void grad_at(const Volumes& V, int64_t o, int zz, int yy, int xx, float& gxv, float& gyv,
                               float& gzv) {
    if constexpr (STORED) {
        gxv = V.gx[o];
        gyv = V.gy[o];
        gzv = V.gz[o];
    } else {
        const float* f = V.f;
        const int64_t r = V.nx, p = (int64_t)V.nx * V.ny;
        gxv = gyv = gzv = 0.0f;
        if (xx >= 3 && xx <= V.nx - 4)
            gxv = C1 * (f[o + 1] - f[o - 1]) + C2 * (f[o + 2] - f[o - 2]) + C3 * (f[o + 3] - f[o - 3]);
        if (yy >= 3 && yy <= V.ny - 4)
            gyv = C1 * (f[o + r] - f[o - r]) + C2 * (f[o + 2 * r] - f[o - 2 * r]) + C3 * (f[o + 3 * r] - f[o - 3 * r]);
        if (zz >= 3 && zz <= V.nz - 4)
            gzv = C1 * (f[o + p] - f[o - p]) + C2 * (f[o + 2 * p] - f[o - 2 * p]) + C3 * (f[o + 3 * p] - f[o - 3 * p]);
    }
}

__device__ inline bool kept_bit(const uint8_t* row, int v) { return ((row[v >> 3] >> (v & 7)) & 1) != 0; }

// ---------------------------------------------------------------------------- reductions
// value of lane k (wave-uniform k) on every lane
__device__ inline double lane_bcast(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline float lane_bcast(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// v of lane (l ^ X) within each half-wave of 32 lanes: ds_swizzle in its bit-mask mode, the pattern
// an immediate (no lane-index register, unlike a general shuffle)
template <int X>
__device__ inline int swizzle_xor(int v) {
    static_assert(X > 0 && X < 32, "xor within 32 lanes");
    return __builtin_amdgcn_ds_swizzle(v, 0x1F | (X << 10));
}
template <int X>
__device__ inline double swizzle_xor(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = swizzle_xor<X>((int)b), hi = swizzle_xor<X>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int X>
__device__ inline float swizzle_xor(float v) {
    return __int_as_float(swizzle_xor<X>(__float_as_int(v)));
}

// Sum over the 64 lanes of a wave, the same value on every lane: a butterfly within each half-wave,
// then the two halves' sums (lanes 0 and 32) added.
template <typename T>
__device__ inline T wave_sum(T v) {
    v += swizzle_xor<16>(v);
    v += swizzle_xor<8>(v);
    v += swizzle_xor<4>(v);
    v += swizzle_xor<2>(v);
    v += swizzle_xor<1>(v);
    return lane_bcast(v, 0) + lane_bcast(v, 32);
}

// Block-wide sums of NQ per-thread float partials, accumulated in float64 across lanes and
// waves, written to the shared array `out` (NQ doubles) and visible to every thread on return.
// `scratch` is a shared array of NQ * NWAVES doubles; both may be reused after the call.
template <int NQ, int NWAVES>
__device__ inline void block_sum(const float (&part)[NQ], double* out, double* scratch) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double s = wave_sum((double)part[q]);
        if (lane == 0) scratch[wave * NQ + q] = s;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NQ; q += NWAVES * 64) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NWAVES; ++w) s += scratch[w * NQ + q];
        out[q] = s;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------- float64 algebra
// numba_kernels.py:58-78
__device__ inline bool cholesky12(const double* H, double* L) {
#pragma unroll 1
    for (int i = 0; i < 144; ++i) L[i] = 0.0;
#pragma unroll 1
    for (int j = 0; j < 12; ++j) {
        double s = H[j * 12 + j];
#pragma unroll 1
        for (int k = 0; k < j; ++k) s -= L[j * 12 + k] * L[j * 12 + k];
        if (!(s > 0.0) || !isfinite(s)) return false;
        const double ljj = sqrt(s);
        L[j * 12 + j] = ljj;
#pragma unroll 1
        for (int i = j + 1; i < 12; ++i) {
            double t = H[i * 12 + j];
#pragma unroll 1
            for (int k = 0; k < j; ++k) t -= L[i * 12 + k] * L[j * 12 + k];
            L[i * 12 + j] = t / ljj;
        }
    }
    return true;
}

// numba_kernels.py:88-103
__device__ inline bool corrected_cholesky12(const double* H, const double* pattern, double corr, double* H0,
                                            double* L0) {
    double limit = 1.0;
    for (int i = 9; i < 12; ++i) {
        const double c = corr * pattern[i * 12 + i];
        if (c > 0.0) {
            const double lim = (1.0 - MIN_CORRECTED_FRACTION) * H[i * 12 + i] / c;
            if (lim < limit) limit = lim;
        }
    }
    if (limit <= 0.0) return false;
#pragma unroll 1
    for (int i = 0; i < 144; ++i) H0[i] = H[i] - corr * limit * pattern[i];
    return cholesky12(H0, L0);
}

// numba_kernels.py:122-142
__device__ inline double inv3(const double* a, double* out) {
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (fabs(det) < 1e-300) return 0.0;
    const double inv = 1.0 / det;
    out[0] = (a[4] * a[8] - a[5] * a[7]) * inv;
    out[1] = (a[2] * a[7] - a[1] * a[8]) * inv;
    out[2] = (a[1] * a[5] - a[2] * a[4]) * inv;
    out[3] = (a[5] * a[6] - a[3] * a[8]) * inv;
    out[4] = (a[0] * a[8] - a[2] * a[6]) * inv;
    out[5] = (a[2] * a[3] - a[0] * a[5]) * inv;
    out[6] = (a[3] * a[7] - a[4] * a[6]) * inv;
    out[7] = (a[1] * a[6] - a[0] * a[7]) * inv;
    out[8] = (a[0] * a[4] - a[1] * a[3]) * inv;
    return det;
}

// W(P) <- W(P) W(dP)^-1 (numba_kernels.py:145-180); `work` holds 36 doubles (LDS for the
// control lane, so the 3x3 temporaries do not count against the sampling lanes' registers).
__device__ inline bool compose_warp(double* P, const double* dP, double* work) {
    double* A = work;
    double* dA = work + 9;
    double* dAi = work + 18;
    double* An = work + 27;
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            A[3 * i + j] = P[3 * i + j];
            dA[3 * i + j] = dP[3 * i + j];
        }
        A[3 * i + i] += 1.0;
        dA[3 * i + i] += 1.0;
    }
    if (inv3(dA, dAi) == 0.0) return false;
#pragma unroll 1
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += A[3 * i + k] * dAi[3 * k + j];
            An[3 * i + j] = s;
        }
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += An[3 * i + k] * dP[9 + k];
        P[9 + i] -= s;
    }
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) P[3 * i + j] = An[3 * i + j];
        P[3 * i + i] -= 1.0;
    }
    return true;
}

// numba_kernels.py:910-928: Hinv = (2 (H_tt - corr I) / bf2 + mu I + LM damping)^-1
__device__ inline bool build_h3(const double* H, double bf2, double mu, double corr, double* Hinv) {
    double H3[9], Hd[9];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) H3[3 * i + j] = H[(9 + i) * 12 + 9 + j] * 2.0 / bf2;
        const double limit_i = (1.0 - MIN_CORRECTED_FRACTION) * H[(9 + i) * 12 + 9 + i];
        const double c = corr < limit_i ? corr : limit_i;
        H3[3 * i + i] -= c * 2.0 / bf2;
        H3[3 * i + i] += mu;
    }
    const double md = H3[4] > H3[8] ? H3[4] : H3[8];
    const double max_diag = H3[0] > md ? H3[0] : md;
    if (!(max_diag > 0.0)) return false;
    for (int i = 0; i < 9; ++i) Hd[i] = H3[i];
    for (int i = 0; i < 3; ++i) Hd[3 * i + i] += LM_DAMPING_3DOF * max_diag;
    return inv3(Hd, Hinv) != 0.0;
}

}  // namespace aldvc
