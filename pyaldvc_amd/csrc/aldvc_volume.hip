// Volume preparation on the GPU: VOI statistics and normalisation, the cubic B-spline prefilter,
// and the deformed-frame NaN mask.
//
// Reference: io/volume_ops.py:53-121 (_voi_moments, _normalize_into, voi_mean_std,
// normalize_volume), :215-244 (prefilter_bspline = scipy.ndimage.spline_filter(order=3,
// mode="mirror", output=float32), prepare_deformed).
//
// Statistics: float64 sums of (v - pilot) and (v - pilot)^2 over the VOI, pilot = the VOI's centre
// voxel as in the reference; one partial per (slice, block of rows) written to the workspace and
// summed in a fixed order by one block, so the result does not depend on scheduling.
//
// B-spline prefilter: the cubic B-spline coefficients c of samples s solve
// (c[k-1] + 4 c[k] + c[k+1]) / 6 = s[k]. On an infinite line the inverse filter is the symmetric
// exponential c[k] = sqrt(3) * sum_j z^|j| s[k+j] with z = sqrt(3) - 2 (M. Unser, "Splines: a
// perfect fit for signal and image processing", IEEE Signal Processing Magazine 16(6), 1999), and
// with whole-sample mirror boundaries it is the same sum over the mirrored extension of the line
// (period 2n - 2). |z|^33 < 1e-18, so 32 taps per side reproduce the recursive formulation to
// float64 round-off. Each output voxel is one thread (no serial recursion along a line); axes are
// filtered one after the other (x, y, z) in float64, with the volume stored as float32 between
// axes, as a float32 output array of the library routine is.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "aldvc_abi_util.h"

using namespace aldvc;

namespace {

constexpr int STAT_BLOCK = 256;
constexpr int ROWS_PER_PART = 16;  // rows of one slice per partial sum

enum VolType : int { VT_F32 = 0, VT_U16 = 1, VT_U8 = 2, VT_I16 = 3, VT_F64 = 4 };

// partial sums over (slice, block of ROWS_PER_PART rows) of the VOI
template <typename T>
__global__ __launch_bounds__(STAT_BLOCK) void k_voi_partials(const T* __restrict__ vol, int64_t ny, int64_t nx, int z0,
                                                            int y0, int y1, int x0, int x1, int row_blocks, double pilot,
                                                            double* __restrict__ part) {
    const int zb = blockIdx.x / row_blocks, rb = blockIdx.x - zb * row_blocks;
    const int ya = y0 + rb * ROWS_PER_PART, yb = min(y1, ya + ROWS_PER_PART);
    const int w = x1 - x0;
    double s = 0.0, ss = 0.0;
    for (int y = ya; y < yb; ++y) {
        const T* row = vol + ((int64_t)(z0 + zb) * ny + y) * nx + x0;
        for (int x = threadIdx.x; x < w; x += STAT_BLOCK) {
            const double v = (double)row[x] - pilot;
            s += v;
            ss += v * v;
        }
    }
    __shared__ double rs[STAT_BLOCK], rss[STAT_BLOCK];
    rs[threadIdx.x] = s;
    rss[threadIdx.x] = ss;
    __syncthreads();
    for (int off = STAT_BLOCK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            rs[threadIdx.x] += rs[threadIdx.x + off];
            rss[threadIdx.x] += rss[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = rs[0];
        part[2 * blockIdx.x + 1] = rss[0];
    }
}

__global__ __launch_bounds__(STAT_BLOCK) void k_sum_partials(const double* __restrict__ part, int64_t n,
                                                            double* __restrict__ out) {
    double s = 0.0, ss = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += STAT_BLOCK) {
        s += part[2 * i];
        ss += part[2 * i + 1];
    }
    __shared__ double rs[STAT_BLOCK], rss[STAT_BLOCK];
    rs[threadIdx.x] = s;
    rss[threadIdx.x] = ss;
    __syncthreads();
    for (int off = STAT_BLOCK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            rs[threadIdx.x] += rs[threadIdx.x + off];
            rss[threadIdx.x] += rss[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = rs[0];
        out[1] = rss[0];
    }
}

// out = float32((v - mean) * inv_std), float64 arithmetic (volume_ops.py:75-81)
template <typename T>
__global__ void k_normalize(const T* __restrict__ vol, int64_t n, double mean, double inv_std, float* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)(((double)vol[i] - mean) * inv_std);
}

// ---------------------------------------------------------------------------- B-spline prefilter
constexpr int BSPLINE_TAPS = 32;  // per side

// index of position k of the whole-sample mirror extension of a line of n samples
__device__ inline int64_t mirror(int64_t k, int64_t n) {
    if (n == 1) return 0;
    const int64_t P = 2 * n - 2;
    k %= P;
    if (k < 0) k += P;
    return k < n ? k : P - k;
}

// One reflection suffices when the line is longer than the filter's reach.
__device__ inline int64_t mirror_once(int64_t k, int64_t n) {
    return k < 0 ? -k : (k >= n ? 2 * n - 2 - k : k);
}

// axis 0 = x, 1 = y, 2 = z; grid (ceil(nx / 256), ny, nz), one output voxel per thread
template <int AXIS>
__global__ void k_bspline_axis(const float* __restrict__ in, float* __restrict__ out, int64_t nz, int64_t ny,
                               int64_t nx) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, zz = blockIdx.z;
    if (x >= nx) return;
    const double z = -0.26794919243112270647;   // sqrt(3) - 2
    const double gain = 1.7320508075688772935;  // sqrt(3)
    const int64_t n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    const int64_t k = AXIS == 0 ? x : (AXIS == 1 ? y : zz);
    const int64_t step = AXIS == 0 ? 1 : (AXIS == 1 ? nx : nx * ny);
    const int64_t i = (zz * ny + y) * nx + x;
    const float* line = in + (i - k * step);
    double acc = (double)line[k * step];
    double w = 1.0;
    if (k >= BSPLINE_TAPS && k + BSPLINE_TAPS < n) {  // interior: no reflection
        const float* c = line + k * step;
#pragma unroll 8
        for (int j = 1; j <= BSPLINE_TAPS; ++j) {
            w *= z;
            acc += w * ((double)c[j * step] + (double)c[-j * step]);
        }
    } else if (n > BSPLINE_TAPS) {
#pragma unroll 4
        for (int j = 1; j <= BSPLINE_TAPS; ++j) {
            w *= z;
            acc += w * ((double)line[mirror_once(k + j, n) * step] + (double)line[mirror_once(k - j, n) * step]);
        }
    } else {
        for (int j = 1; j <= BSPLINE_TAPS; ++j) {
            w *= z;
            acc += w * ((double)line[mirror(k + j, n) * step] + (double)line[mirror(k - j, n) * step]);
        }
    }
    out[i] = (float)(gain * acc);
}

// prepare_deformed's mask: NaN where the deformed-frame mask is 0 (volume_ops.py:240-244)
__global__ void k_mask_nan(float* __restrict__ vol, const uint8_t* __restrict__ mask, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (mask[i] == 0) vol[i] = __builtin_nanf("");
}

// Block mean by an integer factor (trailing voxels dropped): one output voxel per thread, its f^3
// inputs summed in float64 (z, y, x order) and the mean rounded once to float32
// (integer_search.py:235-243, block_downsample).
__global__ void k_block_downsample(const float* __restrict__ vol, int64_t ny, int64_t nx, int fac, int64_t oz,
                                   int64_t oy, int64_t ox, float* __restrict__ out) {
    const int64_t n = oz * oy * ox;
    const double inv = 1.0 / ((double)fac * fac * fac);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = i / (oy * ox), r = i - z * oy * ox, y = r / ox, x = r - y * ox;
        const float* b = vol + ((z * fac) * ny + y * fac) * nx + x * fac;
        double acc = 0.0;
        for (int dz = 0; dz < fac; ++dz)
            for (int dy = 0; dy < fac; ++dy) {
                const float* row = b + ((int64_t)dz * ny + dy) * nx;
                for (int dx = 0; dx < fac; ++dx) acc += (double)row[dx];
            }
        out[i] = (float)(acc * inv);
    }
}

unsigned grid_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

template <typename T>
void launch_partials(const void* vol, int64_t ny, int64_t nx, int z0, int z1, int y0, int y1, int x0, int x1,
                     int row_blocks, double pilot, double* part, hipStream_t st) {
    const unsigned nb = (unsigned)((z1 - z0) * row_blocks);
    hipLaunchKernelGGL(k_voi_partials<T>, dim3(nb), dim3(STAT_BLOCK), 0, st, static_cast<const T*>(vol), ny, nx, z0, y0,
                       y1, x0, x1, row_blocks, pilot, part);
}

template <typename T>
void launch_normalize(const void* vol, int64_t n, double mean, double inv_std, float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_normalize<T>, dim3(grid_for(n)), dim3(256), 0, st, static_cast<const T*>(vol), n, mean,
                       inv_std, out);
}

int64_t row_blocks_of(int y0, int y1) { return (y1 - y0 + ROWS_PER_PART - 1) / ROWS_PER_PART; }

}  // namespace

extern "C" {

size_t aldvc_voi_moments_workspace_bytes(int64_t nz, int64_t ny) {
    return sizeof(double) * 2 * ((size_t)nz * (size_t)((ny + ROWS_PER_PART - 1) / ROWS_PER_PART) + 1);
}

int aldvc_voi_moments(const void* vol, int dtype, int64_t nz, int64_t ny, int64_t nx, int z0, int z1, int y0, int y1,
                      int x0, int x1, double pilot, double* sums, void* workspace, size_t workspace_bytes,
                      hipStream_t stream) {
    clear_error();
    if (!vol || !sums || !workspace) return fail(-2, "null pointer");
    if (nz < 1 || ny < 1 || nx < 1) return fail(-2, "bad volume shape");
    if (z0 < 0 || y0 < 0 || x0 < 0 || z1 > nz || y1 > ny || x1 > nx || z0 >= z1 || y0 >= y1 || x0 >= x1)
        return fail(-2, "empty or out-of-range VOI");
    if (workspace_bytes < aldvc_voi_moments_workspace_bytes(nz, ny)) return fail(-3, "workspace too small");
    const int rb = (int)row_blocks_of(y0, y1);
    double* part = static_cast<double*>(workspace);
    switch (dtype) {
        case VT_F32: launch_partials<float>(vol, ny, nx, z0, z1, y0, y1, x0, x1, rb, pilot, part, stream); break;
        case VT_U16: launch_partials<uint16_t>(vol, ny, nx, z0, z1, y0, y1, x0, x1, rb, pilot, part, stream); break;
        case VT_U8: launch_partials<uint8_t>(vol, ny, nx, z0, z1, y0, y1, x0, x1, rb, pilot, part, stream); break;
        case VT_I16: launch_partials<int16_t>(vol, ny, nx, z0, z1, y0, y1, x0, x1, rb, pilot, part, stream); break;
        case VT_F64: launch_partials<double>(vol, ny, nx, z0, z1, y0, y1, x0, x1, rb, pilot, part, stream); break;
        default: return fail(-2, "dtype must be 0 f32, 1 u16, 2 u8, 3 i16 or 4 f64");
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(STAT_BLOCK), 0, stream, (const double*)part,
                       (int64_t)(z1 - z0) * rb, sums);
    return launch_status();
}

int aldvc_normalize(const void* vol, int dtype, int64_t n, double mean, double inv_std, float* out,
                    hipStream_t stream) {
    clear_error();
    if (!vol || !out) return fail(-2, "null pointer");
    if (n <= 0) return 0;
    switch (dtype) {
        case VT_F32: launch_normalize<float>(vol, n, mean, inv_std, out, stream); break;
        case VT_U16: launch_normalize<uint16_t>(vol, n, mean, inv_std, out, stream); break;
        case VT_U8: launch_normalize<uint8_t>(vol, n, mean, inv_std, out, stream); break;
        case VT_I16: launch_normalize<int16_t>(vol, n, mean, inv_std, out, stream); break;
        case VT_F64: launch_normalize<double>(vol, n, mean, inv_std, out, stream); break;
        default: return fail(-2, "dtype must be 0 f32, 1 u16, 2 u8, 3 i16 or 4 f64");
    }
    return launch_status();
}

int aldvc_bspline_prefilter(const float* vol, int64_t nz, int64_t ny, int64_t nx, float* out, float* scratch,
                            hipStream_t stream) {
    clear_error();
    if (!vol || !out || !scratch) return fail(-2, "null pointer");
    if (nz < 1 || ny < 1 || nx < 1) return fail(-2, "bad volume shape");
    if (out == vol || scratch == vol || scratch == out) return fail(-2, "vol, out and scratch must be distinct");
    if (ny > 65535 || nz > 65535) return fail(-2, "volumes beyond 65535 rows or planes are not supported");
    const dim3 g((unsigned)((nx + 255) / 256), (unsigned)ny, (unsigned)nz);
    hipLaunchKernelGGL(k_bspline_axis<0>, g, dim3(256), 0, stream, vol, out, nz, ny, nx);
    hipLaunchKernelGGL(k_bspline_axis<1>, g, dim3(256), 0, stream, (const float*)out, scratch, nz, ny, nx);
    hipLaunchKernelGGL(k_bspline_axis<2>, g, dim3(256), 0, stream, (const float*)scratch, out, nz, ny, nx);
    return launch_status();
}

int aldvc_block_downsample(const float* vol, int64_t nz, int64_t ny, int64_t nx, int factor, float* out,
                           hipStream_t stream) {
    clear_error();
    if (!vol || !out) return fail(-2, "null pointer");
    if (factor < 1) return fail(-2, "factor must be at least 1");
    if (nz < 0 || ny < 0 || nx < 0) return fail(-2, "negative volume shape");
    const int64_t oz = nz / factor, oy = ny / factor, ox = nx / factor;
    if (oz * oy * ox == 0) return 0;
    hipLaunchKernelGGL(k_block_downsample, dim3(grid_for(oz * oy * ox)), dim3(256), 0, stream, vol, ny, nx, factor, oz,
                       oy, ox, out);
    return launch_status();
}

int aldvc_mask_nan(float* vol, const uint8_t* mask, int64_t n, hipStream_t stream) {
    clear_error();
    if (!vol || !mask) return fail(-2, "null pointer");
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_mask_nan, dim3(grid_for(n)), dim3(256), 0, stream, vol, mask, n);
    return launch_status();
}

}  // extern "C"
