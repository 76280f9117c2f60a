// Subset splitting on the GPU: the in-mask fraction of every node's window and the packed keep
// rows of the 6-connected in-mask component around each cut subset's centre.
//
// Reference: mesh/grid_mesh.py:135-188 (subset_valid_fraction), solver/numba_kernels.py:223-300
// (build_split_rows), utils/flood_fill.py:12-72 (flood_fill_centre).
//
// The fraction comes from a summed-volume table of the mask (int32, (nz+1)(ny+1)(nx+1)), built by
// three line-scan passes, so each node is 8 table reads. The split rows are one 256-lane workgroup
// per candidate subset: the window is held in LDS as one 64-bit word per (z, y) row (windows up to
// 64 voxels wide), and the centre component is grown by repeated 6-neighbour dilation -- a row
// takes the bits of its y and z neighbours, then fills along x within its own runs -- until a
// sweep changes nothing. Every update only sets bits of the final component, so sweeps may read
// rows that other lanes are updating.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "aldvc_abi_util.h"

using namespace aldvc;

namespace {

constexpr int SPLIT_BLOCK = 256;
constexpr int MAX_WINDOW_ROWS = 4096;  // Sz * Sy: two 32 KB word arrays in LDS

// ---------------------------------------------------------------------------- summed-volume table
// T[z][y][x] = number of mask voxels with z' < z, y' < y, x' < x; T is (nz+1, ny+1, nx+1).
__global__ void k_sat_x(const uint8_t* __restrict__ mask, int64_t nz, int64_t ny, int64_t nx, int32_t* T) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (z, y) row of the mask
    const int64_t PX = nx + 1, PY = ny + 1;
    if (r >= (nz + 1) * (ny + 1)) return;
    const int64_t z = r / PY, y = r - z * PY;
    int32_t* t = T + (z * PY + y) * PX;
    t[0] = 0;
    if (z == 0 || y == 0) {
        for (int64_t x = 1; x <= nx; ++x) t[x] = 0;
        return;
    }
    const uint8_t* m = mask + ((z - 1) * ny + (y - 1)) * nx;
    int32_t run = 0;
    for (int64_t x = 0; x < nx; ++x) {
        run += m[x] != 0;
        t[x + 1] = run;
    }
}

__global__ void k_sat_y(int64_t nz, int64_t ny, int64_t nx, int32_t* T) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (z, x) column, x fastest
    const int64_t PX = nx + 1, PY = ny + 1;
    if (c >= (nz + 1) * PX) return;
    const int64_t z = c / PX, x = c - z * PX;
    int32_t* t = T + z * PY * PX + x;
    int32_t run = 0;
    for (int64_t y = 0; y < PY; ++y) {
        run += t[y * PX];
        t[y * PX] = run;
    }
}

__global__ void k_sat_z(int64_t nz, int64_t ny, int64_t nx, int32_t* T) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (y, x) column, x fastest
    const int64_t PX = nx + 1, PY = ny + 1;
    if (c >= PY * PX) return;
    int32_t* t = T + c;
    int32_t run = 0;
    for (int64_t z = 0; z <= nz; ++z) {
        run += t[z * PY * PX];
        t[z * PY * PX] = run;
    }
}

// In-mask voxels of the window [c - h, c + h] per axis, clipped to the volume, over the nominal
// window volume (grid_mesh.py:158-188).
__global__ void k_subset_valid_fraction(const int32_t* __restrict__ T, int64_t nz, int64_t ny, int64_t nx,
                                        const int64_t* __restrict__ coords, int64_t N, int hx, int hy, int hz,
                                        double* __restrict__ frac) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int64_t PX = nx + 1, PY = ny + 1;
    auto clip = [](int64_t v, int64_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
    const int64_t cx = coords[3 * n], cy = coords[3 * n + 1], cz = coords[3 * n + 2];
    const int64_t x0 = clip(cx - hx, nx), x1 = clip(cx + hx + 1, nx);
    const int64_t y0 = clip(cy - hy, ny), y1 = clip(cy + hy + 1, ny);
    const int64_t z0 = clip(cz - hz, nz), z1 = clip(cz + hz + 1, nz);
    int64_t count = 0;
    if (x1 > x0 && y1 > y0 && z1 > z0) {
        auto at = [&](int64_t z, int64_t y, int64_t x) { return (int64_t)T[(z * PY + y) * PX + x]; };
        count = at(z1, y1, x1) - at(z0, y1, x1) - at(z1, y0, x1) - at(z1, y1, x0) + at(z0, y0, x1) + at(z0, y1, x0) +
                at(z1, y0, x0) - at(z0, y0, x0);
    }
    frac[n] = (double)count / (double)((2 * hx + 1) * (2 * hy + 1) * (2 * hz + 1));
}

// ---------------------------------------------------------------------------- split rows
// Fill along x within the runs of `s` that hold a bit of `n` (n is a subset of s).
__device__ inline uint64_t fill_runs(uint64_t n, uint64_t s) {
    for (;;) {
        const uint64_t t = (n | (n << 1) | (n >> 1)) & s;
        if (t == n) return n;
        n = t;
    }
}

__global__ __launch_bounds__(SPLIT_BLOCK) void k_split_rows(const int64_t* __restrict__ coords,
                                                            const int64_t* __restrict__ cand, int hx, int hy,
                                                            int hz, int stride, const uint8_t* __restrict__ mask,
                                                            int64_t nz, int64_t ny, int64_t nx, int margin,
                                                            int64_t keep_bytes, uint8_t* __restrict__ rows,
                                                            int64_t* __restrict__ n_keep,
                                                            int64_t* __restrict__ n_inmask) {
    extern __shared__ __align__(16) uint64_t words[];
    const int Sz = 2 * hz + 1, Sy = 2 * hy + 1, Sx = 2 * hx + 1, R = Sz * Sy;
    uint64_t* sub = words;
    uint64_t* out = words + R;
    __shared__ int changed;
    __shared__ unsigned long long cnt[2];
    const int tid = threadIdx.x;
    const int64_t m = blockIdx.x;
    const int64_t node = cand[m];
    const int64_t x0 = coords[3 * node], y0 = coords[3 * node + 1], z0 = coords[3 * node + 2];

    // the full-resolution window; voxels outside the volume or within `margin` of a face count as
    // masked (numba_kernels.py:242-258)
    for (int r = tid; r < R; r += SPLIT_BLOCK) {
        const int iz = r / Sy, iy = r - iz * Sy;
        const int64_t zz = z0 - hz + iz, yy = y0 - hy + iy;
        uint64_t w = 0;
        if (zz >= margin && zz < nz - margin && yy >= margin && yy < ny - margin) {
            const uint8_t* row = mask + (zz * ny + yy) * nx;
            for (int ix = 0; ix < Sx; ++ix) {
                const int64_t xx = x0 - hx + ix;
                if (xx >= margin && xx < nx - margin && row[xx] != 0) w |= (uint64_t)1 << ix;
            }
        }
        sub[r] = w;
        out[r] = 0;
    }
    if (tid == 0) {
        changed = 0;
        cnt[0] = cnt[1] = 0;
    }
    __syncthreads();
    const int cr = hz * Sy + hy;
    if (tid == 0 && ((sub[cr] >> hx) & 1)) {
        out[cr] = fill_runs((uint64_t)1 << hx, sub[cr]);
        changed = 1;
    }
    __syncthreads();
    // grow until a sweep sets no bit; every sweep either adds a voxel or ends the loop
    while (changed) {
        __syncthreads();
        if (tid == 0) changed = 0;
        __syncthreads();
        for (int r = tid; r < R; r += SPLIT_BLOCK) {
            const uint64_t s = sub[r];
            if (!s) continue;
            const int iz = r / Sy, iy = r - iz * Sy;
            const uint64_t o = out[r];
            uint64_t nb = o;
            if (iy > 0) nb |= out[r - 1];
            if (iy < Sy - 1) nb |= out[r + 1];
            if (iz > 0) nb |= out[r - Sy];
            if (iz < Sz - 1) nb |= out[r + Sy];
            nb &= s;
            if (nb & ~o) {
                out[r] = fill_runs(nb, s);
                changed = 1;
            }
        }
        __syncthreads();
    }
    // pack in the kernels' sampling order (dz, dy, dx with the stride), bit i of byte i / 8
    const int sx = (2 * hx) / stride + 1, sy = (2 * hy) / stride + 1, sz = (2 * hz) / stride + 1;
    const int64_t S = (int64_t)sx * sy * sz;
    unsigned long long keep = 0, inm = 0;
    uint8_t* dst = rows + m * keep_bytes;
    for (int64_t b = tid; b < keep_bytes; b += SPLIT_BLOCK) {
        unsigned byte = 0;
        for (int k = 0; k < 8; ++k) {
            const int64_t v = 8 * b + k;
            if (v >= S) break;
            const int kz = (int)(v / ((int64_t)sx * sy));
            const int rem = (int)(v - (int64_t)kz * sx * sy);
            const int ky = rem / sx, kx = rem - ky * sx;
            const int r = (kz * stride) * Sy + ky * stride, ix = kx * stride;
            if ((sub[r] >> ix) & 1) ++inm;
            if ((out[r] >> ix) & 1) {
                byte |= 1u << k;
                ++keep;
            }
        }
        dst[b] = (uint8_t)byte;
    }
    atomicAdd(&cnt[0], keep);
    atomicAdd(&cnt[1], inm);
    __syncthreads();
    if (tid == 0) {
        n_keep[m] = (int64_t)cnt[0];
        n_inmask[m] = (int64_t)cnt[1];
    }
}

// ---------------------------------------------------------------------------- bridging elements
// An element of the node grid spans a masked boundary when its eight corners (voxels, rounded) do
// not all lie in the 6-connected in-mask component of its first corner within the corners'
// bounding box (mesh/mesh_cut.py:44-95); a first corner on masked material makes it bridging at
// once. One wave per element: the box in LDS as one 64-bit word per (z, y) row, the component
// grown from the first corner by the dilation sweeps of the split rows. out[e] = 1 bridging, 0
// not, 2 when the box exceeds `max_extent` voxels along an axis (the caller raises).
constexpr int BRIDGE_BLOCK = 64;

__global__ __launch_bounds__(BRIDGE_BLOCK) void k_bridging(const int64_t* __restrict__ corners, int64_t E,
                                                           const uint8_t* __restrict__ mask, int64_t nz, int64_t ny,
                                                           int64_t nx, int max_extent, uint8_t* __restrict__ out) {
    extern __shared__ __align__(16) uint64_t words[];
    __shared__ int changed;
    const int tid = threadIdx.x;
    const int64_t e = blockIdx.x;
    const int64_t* c = corners + e * 24;
    int64_t lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};  // x, y, z
    for (int k = 1; k < 8; ++k)
        for (int d = 0; d < 3; ++d) {
            lo[d] = min(lo[d], c[3 * k + d]);
            hi[d] = max(hi[d], c[3 * k + d]);
        }
    const int64_t dims[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d) {
        lo[d] = max(lo[d], (int64_t)0);
        hi[d] = min(hi[d], dims[d] - 1);
    }
    const int64_t Sx = hi[0] - lo[0] + 1, Sy = hi[1] - lo[1] + 1, Sz = hi[2] - lo[2] + 1;
    if (Sx < 1 || Sy < 1 || Sz < 1 || Sx > max_extent || Sy > max_extent || Sz > max_extent) {
        if (tid == 0) out[e] = Sx < 1 || Sy < 1 || Sz < 1 ? 1 : 2;  // an empty box: no in-mask corner
        return;
    }
    const int R = (int)(Sz * Sy), sx = (int)Sx, sy = (int)Sy;
    uint64_t* sub = words;
    uint64_t* comp = words + (int64_t)max_extent * max_extent;
    for (int r = tid; r < R; r += BRIDGE_BLOCK) {
        const int iz = r / sy, iy = r - iz * sy;
        const uint8_t* row = mask + ((lo[2] + iz) * ny + (lo[1] + iy)) * nx + lo[0];
        uint64_t w = 0;
        for (int ix = 0; ix < sx; ++ix)
            if (row[ix] != 0) w |= (uint64_t)1 << ix;
        sub[r] = w;
        comp[r] = 0;
    }
    auto local = [&](int k, int& rz, int& ry, int& rx) {
        rx = (int)min(max(c[3 * k] - lo[0], (int64_t)0), Sx - 1);
        ry = (int)min(max(c[3 * k + 1] - lo[1], (int64_t)0), Sy - 1);
        rz = (int)min(max(c[3 * k + 2] - lo[2], (int64_t)0), Sz - 1);
    };
    int z0, y0, x0;
    local(0, z0, y0, x0);
    if (tid == 0) changed = 0;
    __syncthreads();
    const int r0 = z0 * sy + y0;
    const bool seeded = (sub[r0] >> x0) & 1;
    if (!seeded) {
        if (tid == 0) out[e] = 1;
        return;
    }
    if (tid == 0) {
        comp[r0] = fill_runs((uint64_t)1 << x0, sub[r0]);
        changed = 1;
    }
    __syncthreads();
    while (changed) {
        __syncthreads();
        if (tid == 0) changed = 0;
        __syncthreads();
        for (int r = tid; r < R; r += BRIDGE_BLOCK) {
            const uint64_t sm = sub[r];
            if (!sm) continue;
            const int iz = r / sy, iy = r - iz * sy;
            const uint64_t o = comp[r];
            uint64_t nb = o;
            if (iy > 0) nb |= comp[r - 1];
            if (iy < sy - 1) nb |= comp[r + 1];
            if (iz > 0) nb |= comp[r - sy];
            if (iz < Sz - 1) nb |= comp[r + sy];
            nb &= sm;
            if (nb & ~o) {
                comp[r] = fill_runs(nb, sm);
                changed = 1;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        bool bridging = false;
        for (int k = 1; k < 8; ++k) {
            int rz, ry, rx;
            local(k, rz, ry, rx);
            if (!((comp[rz * sy + ry] >> rx) & 1)) bridging = true;
        }
        out[e] = bridging ? 1 : 0;
    }
}

}  // namespace

extern "C" {

int aldvc_bridging_elements(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, const int64_t* corners, int64_t E,
                            int max_extent, uint8_t* out, hipStream_t stream) {
    clear_error();
    if (E < 0) return fail(-2, "negative element count");
    if (E == 0) return 0;
    if (!mask || !corners || !out) return fail(-2, "null pointer");
    if (nz < 1 || ny < 1 || nx < 1) return fail(-2, "bad volume shape");
    if (max_extent < 1 || max_extent > 64) return fail(-2, "element boxes must be 1 to 64 voxels along every axis");
    if (E > INT32_MAX) return fail(-2, "too many elements for one launch");
    const size_t lds = 2 * sizeof(uint64_t) * (size_t)max_extent * max_extent;
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void*)k_bridging, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return fail(-5, "could not raise the LDS limit");
    hipLaunchKernelGGL(k_bridging, dim3((unsigned)E), dim3(BRIDGE_BLOCK), lds, stream, corners, E, mask, nz, ny, nx,
                       max_extent, out);
    return launch_status();
}


size_t aldvc_mask_table_bytes(int64_t nz, int64_t ny, int64_t nx) {
    return sizeof(int32_t) * (size_t)(nz + 1) * (size_t)(ny + 1) * (size_t)(nx + 1);
}

int aldvc_subset_valid_fraction(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, const int64_t* coords,
                                int64_t N, int hx, int hy, int hz, double* frac, void* workspace,
                                size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    if (!mask || !coords || !frac || !workspace) return fail(-2, "null pointer");
    if (nz < 1 || ny < 1 || nx < 1 || nz * ny * nx > INT32_MAX) return fail(-2, "bad volume shape");
    if (hx < 0 || hy < 0 || hz < 0) return fail(-2, "bad subset half-widths");
    if (workspace_bytes < aldvc_mask_table_bytes(nz, ny, nx)) return fail(-3, "workspace too small");
    if (N <= 0) return 0;
    int32_t* T = static_cast<int32_t*>(workspace);
    const int64_t rows = (nz + 1) * (ny + 1), cols_y = (nz + 1) * (nx + 1), cols_z = (ny + 1) * (nx + 1);
    hipLaunchKernelGGL(k_sat_x, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, stream, mask, nz, ny, nx, T);
    hipLaunchKernelGGL(k_sat_y, dim3((unsigned)((cols_y + 255) / 256)), dim3(256), 0, stream, nz, ny, nx, T);
    hipLaunchKernelGGL(k_sat_z, dim3((unsigned)((cols_z + 255) / 256)), dim3(256), 0, stream, nz, ny, nx, T);
    hipLaunchKernelGGL(k_subset_valid_fraction, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, stream,
                       (const int32_t*)T, nz, ny, nx, coords, N, hx, hy, hz, frac);
    return launch_status();
}

int aldvc_split_rows(const int64_t* coords, const int64_t* cand, int64_t M, int hx, int hy, int hz, int stride,
                     const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, int margin, int64_t keep_bytes,
                     uint8_t* rows, int64_t* n_keep, int64_t* n_inmask, hipStream_t stream) {
    clear_error();
    if (!coords || !cand || !mask || !rows || !n_keep || !n_inmask) return fail(-2, "null pointer");
    if (hx < 0 || hy < 0 || hz < 0 || stride < 1 || margin < 0) return fail(-2, "bad subset half-widths or stride");
    if (2 * hx + 1 > 64) return fail(-2, "split rows support windows up to 64 voxels along x");
    const int64_t R = (int64_t)(2 * hz + 1) * (2 * hy + 1);
    if (R > MAX_WINDOW_ROWS) return fail(-2, "split rows support windows up to 4096 (y, z) rows");
    const int64_t S = (int64_t)((2 * hx) / stride + 1) * ((2 * hy) / stride + 1) * ((2 * hz) / stride + 1);
    if (keep_bytes < (S + 7) / 8) return fail(-2, "keep_bytes smaller than ceil(S / 8)");
    if (nz < 1 || ny < 1 || nx < 1) return fail(-2, "bad volume shape");
    if (M <= 0) return 0;
    if (M > INT32_MAX) return fail(-2, "too many candidate subsets for one launch");
    const size_t lds = sizeof(uint64_t) * 2 * (size_t)R;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_split_rows), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
        return fail(-5, "cannot raise the LDS limit of the split-row kernel");
    hipLaunchKernelGGL(k_split_rows, dim3((unsigned)M), dim3(SPLIT_BLOCK), lds, stream, coords, cand, hx, hy, hz, stride,
                       mask, nz, ny, nx, margin, keep_bytes, rows, n_keep, n_inmask);
    return launch_status();
}

}  // extern "C"
