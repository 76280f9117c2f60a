/*
 * ORACLE — test infrastructure only, never product code.
 *
 * A plain C99 + OpenMP restatement, in float64, of the reference's CPU kernels of the local
 * DVC step (pyALDVC / al_dvc 1.2.1):
 *
 *   solver/interp_kernels.py   sample_volume, interp_margin_ok, sample_points
 *   solver/numba_kernels.py    _cholesky12, _corrected_cholesky12, _chol_solve12, _inv3,
 *                              compose_warp_inplace, _grad_at, _precompute_one,
 *                              _precompute_nodes_jit, _warp_and_sample, _zncc_from_buffer,
 *                              _icgn_12dof_single, _icgn_12dof_parallel_jit, _build_h3,
 *                              _icgn_3dof_single, _icgn_3dof_parallel_jit, evaluate_zncc_parallel,
 *                              _build_split_rows_jit
 *   solver/integer_search.py   _zncc_direct
 *   io/volume_ops.py           _gradient_stencil7
 *   utils/flood_fill.py        flood_fill_centre
 *
 * Only tests/, the smoke check and the CPU-baseline leg of the benchmark load this file's
 * library, and only as the checker (or, for the baseline, as the timed CPU restatement).
 * The arithmetic order of every accumulation follows the reference loops and the library
 * is compiled with -ffp-contract=off, so the float64 results reproduce the reference's
 * Numba kernels (which do not contract either) up to libm's sqrt/floor, i.e. exactly.
 *
 * Layouts (reference core/data_structures.py:1-11, numba_kernels.py:9-16):
 *   volumes (nz, ny, nx) C order, float32;  coords (N, 3) int64 [x, y, z];
 *   P (N, 12) float64 = [F row-major (9), u, v, w];  H, L (N, 12, 12) float64.
 * "No gradient volumes" (the on-the-fly stencil) is gx == NULL; "no mask" is mask == NULL.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* numba_kernels.py:29-44, :81-85 */
#define ST_CONVERGED 0
#define ST_MAX_ITER 1
#define ST_OUT_OF_BOUNDS 2
#define ST_INVALID_SUBSET 3
#define ST_SINGULAR 4
#define ST_NAN 5
#define ST_SKIPPED 6
#define ST_STALLED 7

#define ABS_TOL 1e-5
#define STALL_ZNCC_EPS 1e-4
#define STALL_STEP_DECAY 0.999
#define SCHEDULE_LANES 64
#define MIN_SUBSET_VOXELS 27
#define MIN_VALID_FRACTION 0.5
#define LM_DAMPING_3DOF 1e-3
#define MIN_CORRECTED_FRACTION 0.5
#define NOISE_CORR_STEP 0.5
#define PREDICT_CONTRACTION 0.5

/* interp_kernels.py:23-31 */
#define INTERP_LINEAR 0
#define INTERP_CUBIC 1
#define INTERP_BSPLINE 2
#define SAMPLE_LO 1.0
#define SAMPLE_HI_MARGIN 2.0

typedef struct {
    const float *f, *gx, *gy, *gz, *g;
    const uint8_t *mask;
    int64_t nz, ny, nx;
} vols_t;

#define AT(v, z, y, x) ((v)[((int64_t)(z) * V->ny + (int64_t)(y)) * V->nx + (int64_t)(x)])

int orc_abi_version(void) { return 1; }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ interpolation */

/* interp_kernels.py:34-41 */
static inline void keys_w(double t, double *w0, double *w1, double *w2, double *w3)
{
    *w0 = ((-0.5 * t + 1.0) * t - 0.5) * t;
    *w1 = (1.5 * t - 2.5) * t * t + 1.0;
    *w2 = ((-1.5 * t + 2.0) * t + 0.5) * t;
    *w3 = (0.5 * t - 0.5) * t * t;
}

/* interp_kernels.py:44-53 */
static inline void bspl_w(double t, double *w0, double *w1, double *w2, double *w3)
{
    double t2 = t * t, t3 = t2 * t;
    *w0 = (1.0 - 3.0 * t + 3.0 * t2 - t3) / 6.0;
    *w1 = (4.0 - 6.0 * t2 + 3.0 * t3) / 6.0;
    *w2 = (1.0 + 3.0 * t + 3.0 * t2 - 3.0 * t3) / 6.0;
    *w3 = t3 / 6.0;
}

/* interp_kernels.py:56-58 (_row4) */
static inline double row4(const float *vol, const vols_t *V, int64_t zz, int64_t yy, int64_t xb,
                          double w0, double w1, double w2, double w3)
{
    const float *p = vol + (zz * V->ny + yy) * V->nx + xb;
    return (double)p[0] * w0 + (double)p[1] * w1 + (double)p[2] * w2 + (double)p[3] * w3;
}

/* interp_kernels.py:61-68 (_plane4) */
static inline double plane4(const float *vol, const vols_t *V, int64_t zz, int64_t yb, int64_t xb,
                            double wx0, double wx1, double wx2, double wx3,
                            double wy0, double wy1, double wy2, double wy3)
{
    return row4(vol, V, zz, yb, xb, wx0, wx1, wx2, wx3) * wy0
         + row4(vol, V, zz, yb + 1, xb, wx0, wx1, wx2, wx3) * wy1
         + row4(vol, V, zz, yb + 2, xb, wx0, wx1, wx2, wx3) * wy2
         + row4(vol, V, zz, yb + 3, xb, wx0, wx1, wx2, wx3) * wy3;
}

/* interp_kernels.py:71-124 (sample_volume); the caller has checked the domain */
static double sample_volume(const float *vol, const vols_t *V, double z, double y, double x, int mode)
{
    int64_t ix = (int64_t)floor(x), iy = (int64_t)floor(y), iz = (int64_t)floor(z);
    if (ix > V->nx - 3) ix = V->nx - 3;
    if (iy > V->ny - 3) iy = V->ny - 3;
    if (iz > V->nz - 3) iz = V->nz - 3;
    double fx = x - (double)ix, fy = y - (double)iy, fz = z - (double)iz;
    if (mode == INTERP_LINEAR) {
        double c00 = (double)AT(vol, iz, iy, ix) * (1.0 - fx) + (double)AT(vol, iz, iy, ix + 1) * fx;
        double c10 = (double)AT(vol, iz, iy + 1, ix) * (1.0 - fx) + (double)AT(vol, iz, iy + 1, ix + 1) * fx;
        double c01 = (double)AT(vol, iz + 1, iy, ix) * (1.0 - fx) + (double)AT(vol, iz + 1, iy, ix + 1) * fx;
        double c11 = (double)AT(vol, iz + 1, iy + 1, ix) * (1.0 - fx) + (double)AT(vol, iz + 1, iy + 1, ix + 1) * fx;
        double c0 = c00 * (1.0 - fy) + c10 * fy;
        double c1 = c01 * (1.0 - fy) + c11 * fy;
        return c0 * (1.0 - fz) + c1 * fz;
    }
    double wx0, wx1, wx2, wx3, wy0, wy1, wy2, wy3, wz0, wz1, wz2, wz3;
    if (mode == INTERP_CUBIC) {
        keys_w(fx, &wx0, &wx1, &wx2, &wx3);
        keys_w(fy, &wy0, &wy1, &wy2, &wy3);
        keys_w(fz, &wz0, &wz1, &wz2, &wz3);
    } else {
        bspl_w(fx, &wx0, &wx1, &wx2, &wx3);
        bspl_w(fy, &wy0, &wy1, &wy2, &wy3);
        bspl_w(fz, &wz0, &wz1, &wz2, &wz3);
    }
    int64_t xb = ix - 1, yb = iy - 1, zb = iz - 1;
    return plane4(vol, V, zb, yb, xb, wx0, wx1, wx2, wx3, wy0, wy1, wy2, wy3) * wz0
         + plane4(vol, V, zb + 1, yb, xb, wx0, wx1, wx2, wx3, wy0, wy1, wy2, wy3) * wz1
         + plane4(vol, V, zb + 2, yb, xb, wx0, wx1, wx2, wx3, wy0, wy1, wy2, wy3) * wz2
         + plane4(vol, V, zb + 3, yb, xb, wx0, wx1, wx2, wx3, wy0, wy1, wy2, wy3) * wz3;
}

/* interp_kernels.py:127-136 */
static inline int margin_ok(double z, double y, double x, const vols_t *V)
{
    if (x < SAMPLE_LO || x > (double)(V->nx - 1) - SAMPLE_HI_MARGIN) return 0;
    if (y < SAMPLE_LO || y > (double)(V->ny - 1) - SAMPLE_HI_MARGIN) return 0;
    if (z < SAMPLE_LO || z > (double)(V->nz - 1) - SAMPLE_HI_MARGIN) return 0;
    return 1;
}

/* interp_kernels.py:139-152 (sample_points): NaN outside the admissible domain */
void orc_sample_points(const float *vol, int64_t nz, int64_t ny, int64_t nx, const double *zs,
                       const double *ys, const double *xs, int64_t n, int mode, double *out)
{
    vols_t Vs = {0};
    Vs.nz = nz; Vs.ny = ny; Vs.nx = nx;
    const vols_t *V = &Vs;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        out[i] = margin_ok(zs[i], ys[i], xs[i], V) ? sample_volume(vol, V, zs[i], ys[i], xs[i], mode) : NAN;
}

/* ------------------------------------------------------------------ gradients */

/* io/volume_ops.py:124-155 (_gradient_stencil7): float64 arithmetic, float32 storage,
 * zero within `border` voxels of any face */
void orc_gradients(const float *f, int64_t nz, int64_t ny, int64_t nx, int border, float *gx,
                   float *gy, float *gz)
{
    vols_t Vs = {0};
    Vs.nz = nz; Vs.ny = ny; Vs.nx = nx;
    const vols_t *V = &Vs;
    const double c1 = 3.0 / 4.0, c2 = -3.0 / 20.0, c3 = 1.0 / 60.0;
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < nz; ++z) {
        int z_in = z >= border && z < nz - border;
        for (int64_t y = 0; y < ny; ++y) {
            int y_in = y >= border && y < ny - border;
            for (int64_t x = 0; x < nx; ++x) {
                int64_t o = (z * ny + y) * nx + x;
                if (z_in && y_in && x >= border && x < nx - border) {
                    gx[o] = (float)(c1 * ((double)AT(f, z, y, x + 1) - (double)AT(f, z, y, x - 1))
                                    + c2 * ((double)AT(f, z, y, x + 2) - (double)AT(f, z, y, x - 2))
                                    + c3 * ((double)AT(f, z, y, x + 3) - (double)AT(f, z, y, x - 3)));
                    gy[o] = (float)(c1 * ((double)AT(f, z, y + 1, x) - (double)AT(f, z, y - 1, x))
                                    + c2 * ((double)AT(f, z, y + 2, x) - (double)AT(f, z, y - 2, x))
                                    + c3 * ((double)AT(f, z, y + 3, x) - (double)AT(f, z, y - 3, x)));
                    gz[o] = (float)(c1 * ((double)AT(f, z + 1, y, x) - (double)AT(f, z - 1, y, x))
                                    + c2 * ((double)AT(f, z + 2, y, x) - (double)AT(f, z - 2, y, x))
                                    + c3 * ((double)AT(f, z + 3, y, x) - (double)AT(f, z - 3, y, x)));
                } else {
                    gx[o] = 0.0f;
                    gy[o] = 0.0f;
                    gz[o] = 0.0f;
                }
            }
        }
    }
}

/* numba_kernels.py:188-212 (_grad_at). The on-the-fly stencil is zero within 3 voxels of a
 * face, as the reference GPU kernel does (cuda_kernels.py:343-379); the reference CPU kernel
 * never reaches those voxels because such subsets are refused (local_icgn.py:216-217). */
static inline void grad_at(const vols_t *V, int64_t zz, int64_t yy, int64_t xx, double *gxv,
                           double *gyv, double *gzv)
{
    if (V->gx) {
        *gxv = (double)AT(V->gx, zz, yy, xx);
        *gyv = (double)AT(V->gy, zz, yy, xx);
        *gzv = (double)AT(V->gz, zz, yy, xx);
        return;
    }
    const float *f = V->f;
    const double c1 = 0.75, c2 = -0.15, c3 = 1.0 / 60.0;
    *gxv = *gyv = *gzv = 0.0;
    if (xx >= 3 && xx <= V->nx - 4)
        *gxv = c1 * ((double)AT(f, zz, yy, xx + 1) - (double)AT(f, zz, yy, xx - 1))
             + c2 * ((double)AT(f, zz, yy, xx + 2) - (double)AT(f, zz, yy, xx - 2))
             + c3 * ((double)AT(f, zz, yy, xx + 3) - (double)AT(f, zz, yy, xx - 3));
    if (yy >= 3 && yy <= V->ny - 4)
        *gyv = c1 * ((double)AT(f, zz, yy + 1, xx) - (double)AT(f, zz, yy - 1, xx))
             + c2 * ((double)AT(f, zz, yy + 2, xx) - (double)AT(f, zz, yy - 2, xx))
             + c3 * ((double)AT(f, zz, yy + 3, xx) - (double)AT(f, zz, yy - 3, xx));
    if (zz >= 3 && zz <= V->nz - 4)
        *gzv = c1 * ((double)AT(f, zz + 1, yy, xx) - (double)AT(f, zz - 1, yy, xx))
             + c2 * ((double)AT(f, zz + 2, yy, xx) - (double)AT(f, zz - 2, yy, xx))
             + c3 * ((double)AT(f, zz + 3, yy, xx) - (double)AT(f, zz - 3, yy, xx));
}

/* ------------------------------------------------------------------ small dense algebra */

/* numba_kernels.py:58-78 */
int orc_cholesky12(const double *H, double *L)
{
    for (int i = 0; i < 144; ++i) L[i] = 0.0;
    for (int j = 0; j < 12; ++j) {
        double s = H[j * 12 + j];
        for (int k = 0; k < j; ++k) s -= L[j * 12 + k] * L[j * 12 + k];
        if (s <= 0.0 || !isfinite(s)) return 0;
        double ljj = sqrt(s);
        L[j * 12 + j] = ljj;
        for (int i = j + 1; i < 12; ++i) {
            double t = H[i * 12 + j];
            for (int k = 0; k < j; ++k) t -= L[i * 12 + k] * L[j * 12 + k];
            L[i * 12 + j] = t / ljj;
        }
    }
    return 1;
}

/* numba_kernels.py:88-103 */
static int corrected_cholesky12(const double *H, const double *pattern, double corr, double *H0,
                                double *L0)
{
    double limit = 1.0;
    for (int i = 9; i < 12; ++i) {
        double c = corr * pattern[i * 12 + i];
        if (c > 0.0) {
            double lim = (1.0 - MIN_CORRECTED_FRACTION) * H[i * 12 + i] / c;
            if (lim < limit) limit = lim;
        }
    }
    if (limit <= 0.0) return 0;
    for (int i = 0; i < 144; ++i) H0[i] = H[i] - corr * limit * pattern[i];
    return orc_cholesky12(H0, L0);
}

/* numba_kernels.py:106-119 */
static void chol_solve12(const double *L, const double *b, double *x)
{
    double y[12];
    for (int i = 0; i < 12; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i * 12 + k] * y[k];
        y[i] = s / L[i * 12 + i];
    }
    for (int i = 11; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 12; ++k) s -= L[k * 12 + i] * x[k];
        x[i] = s / L[i * 12 + i];
    }
}

/* numba_kernels.py:122-142 */
static double inv3(const double a[9], double out[9])
{
    double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6])
               + a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (fabs(det) < 1e-300) return 0.0;
    double inv = 1.0 / det;
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

/* numba_kernels.py:145-180: W(P) <- W(P) W(dP)^-1 */
int orc_compose_warp(double *P, const double *dP)
{
    double A[9], dA[9], dAi[9], An[9];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            A[3 * i + j] = P[3 * i + j];
            dA[3 * i + j] = dP[3 * i + j];
        }
        A[3 * i + i] += 1.0;
        dA[3 * i + i] += 1.0;
    }
    if (inv3(dA, dAi) == 0.0) return 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += A[3 * i + k] * dAi[3 * k + j];
            An[3 * i + j] = s;
        }
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += An[3 * i + k] * dP[9 + k];
        P[9 + i] = P[9 + i] - s;
    }
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) P[3 * i + j] = An[3 * i + j];
        P[3 * i + i] -= 1.0;
    }
    return 1;
}

/* ------------------------------------------------------------------ subset helpers */

/* numba_kernels.py:52-55 */
static inline int64_t subset_count(int hx, int hy, int hz, int stride)
{
    return (int64_t)((2 * hx) / stride + 1) * ((2 * hy) / stride + 1) * ((2 * hz) / stride + 1);
}

/* numba_kernels.py:320-323 */
static inline int kept(const uint8_t *keep, int64_t idx) { return ((keep[idx >> 3] >> (idx & 7)) & 1) != 0; }

/* numba_kernels.py:326-418 (_precompute_one). Returns ok; writes meanf, bottomf, n_valid. */
static int precompute_one(const vols_t *V, int64_t x0, int64_t y0, int64_t z0, int hx, int hy, int hz,
                          int stride, double min_valid_ratio, double cond_max, double *H, double *L,
                          const uint8_t *keep, double *meanf_o, double *bottomf_o, int64_t *nv_o)
{
    for (int i = 0; i < 144; ++i) H[i] = 0.0;
    *meanf_o = 0.0; *bottomf_o = 1.0; *nv_o = 0;
    if (!keep && (x0 - hx < 0 || x0 + hx >= V->nx || y0 - hy < 0 || y0 + hy >= V->ny || z0 - hz < 0
                  || z0 + hz >= V->nz))
        return 0;
    double sd[12];
    int64_t idx = 0, n_valid = 0;
    double sum_f = 0.0, sum_f2 = 0.0;
    for (int dz = -hz; dz <= hz; dz += stride) {
        int64_t zz = z0 + dz;
        double Z = (double)dz;
        for (int dy = -hy; dy <= hy; dy += stride) {
            int64_t yy = y0 + dy;
            double Y = (double)dy;
            for (int dx = -hx; dx <= hx; dx += stride) {
                int64_t xx = x0 + dx;
                int k;
                if (keep) k = kept(keep, idx);
                else k = !V->mask || AT(V->mask, zz, yy, xx) != 0;
                idx += 1;
                if (!k) continue;
                double X = (double)dx;
                double fv = (double)AT(V->f, zz, yy, xx);
                double gxv, gyv, gzv;
                grad_at(V, zz, yy, xx, &gxv, &gyv, &gzv);
                n_valid += 1;
                sum_f += fv;
                sum_f2 += fv * fv;
                sd[0] = gxv * X; sd[1] = gxv * Y; sd[2] = gxv * Z;
                sd[3] = gyv * X; sd[4] = gyv * Y; sd[5] = gyv * Z;
                sd[6] = gzv * X; sd[7] = gzv * Y; sd[8] = gzv * Z;
                sd[9] = gxv; sd[10] = gyv; sd[11] = gzv;
                for (int a = 0; a < 12; ++a) {
                    double sa = sd[a];
                    for (int b = a; b < 12; ++b) H[a * 12 + b] += sa * sd[b];
                }
            }
        }
    }
    for (int a = 0; a < 12; ++a)
        for (int b = 0; b < a; ++b) H[a * 12 + b] = H[b * 12 + a];
    *nv_o = n_valid;
    int64_t total = subset_count(hx, hy, hz, stride);
    if (n_valid < 27 || (double)n_valid < min_valid_ratio * (double)total) return 0;
    double meanf = sum_f / (double)n_valid;
    double ssf = sum_f2 - (double)n_valid * meanf * meanf;
    if (ssf < 0.0) ssf = 0.0;
    double bottomf = sqrt(ssf > 1e-30 ? ssf : 1e-30);
    *meanf_o = meanf;
    *bottomf_o = bottomf;
    if (bottomf < 1e-10) return 0;
    if (!orc_cholesky12(H, L)) return 0;
    double dmin = L[0], dmax = L[0];
    for (int i = 1; i < 12; ++i) {
        if (L[i * 13] < dmin) dmin = L[i * 13];
        if (L[i * 13] > dmax) dmax = L[i * 13];
    }
    if (dmin <= 0.0 || (dmax / dmin) * (dmax / dmin) > cond_max) return 0;
    return 1;
}

static inline const uint8_t *keep_row(const int64_t *split_index, const uint8_t *split_keep,
                                      int64_t keep_bytes, int64_t n)
{
    if (!split_index || !split_keep) return NULL;
    int64_t r = split_index[n];
    return r >= 0 ? split_keep + r * keep_bytes : NULL;
}

/* numba_kernels.py:446-509 (_precompute_nodes_jit / precompute_nodes). The caller zeroes
 * H_all/L_all/meanf/nvalid/valid and sets bottomf to 1 (the reference's np.zeros/np.ones). */
void orc_precompute(const int64_t *coords, int64_t N, int hx, int hy, int hz, int stride,
                    const float *f, const float *gx, const float *gy, const float *gz,
                    const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, double min_valid_ratio,
                    double cond_max, const int64_t *split_index, const uint8_t *split_keep,
                    int64_t keep_bytes, double *H_all, double *L_all, double *meanf,
                    double *bottomf, int64_t *nvalid, uint8_t *valid)
{
    vols_t Vs = {f, gx, gy, gz, NULL, mask, nz, ny, nx};
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
        double mf, bf;
        int64_t nv;
        int ok = precompute_one(&Vs, coords[3 * n], coords[3 * n + 1], coords[3 * n + 2], hx, hy, hz,
                                stride, min_valid_ratio, cond_max, H_all + 144 * n, L_all + 144 * n,
                                keep_row(split_index, split_keep, keep_bytes, n), &mf, &bf, &nv);
        meanf[n] = mf;
        bottomf[n] = bf;
        nvalid[n] = nv;
        valid[n] = (uint8_t)ok;
    }
}

/* ------------------------------------------------------------------ warp + sample */

typedef struct {
    int ok;
    int64_t n_valid;
    double meanf, bottomf, meang, bottomg;
} wstat_t;

/* numba_kernels.py:517-598 (_warp_and_sample) */
static wstat_t warp_and_sample(const double *P, int64_t x0, int64_t y0, int64_t z0, int hx, int hy,
                               int hz, int stride, const vols_t *V, int mode, double *gbuf,
                               const uint8_t *keep)
{
    wstat_t r = {0, 0, 0.0, 1.0, 0.0, 1.0};
    double a00 = 1.0 + P[0], a01 = P[1], a02 = P[2];
    double a10 = P[3], a11 = 1.0 + P[4], a12 = P[5];
    double a20 = P[6], a21 = P[7], a22 = 1.0 + P[8];
    double tx = (double)x0 + P[9], ty = (double)y0 + P[10], tz = (double)z0 + P[11];
    int64_t n_valid = 0, n_ref = 0, idx = 0;
    double s1 = 0.0, s2 = 0.0, s1f = 0.0, s2f = 0.0;
    for (int dz = -hz; dz <= hz; dz += stride) {
        double Z = (double)dz;
        int64_t zz = z0 + dz;
        for (int dy = -hy; dy <= hy; dy += stride) {
            double Y = (double)dy;
            int64_t yy = y0 + dy;
            for (int dx = -hx; dx <= hx; dx += stride) {
                int skip;
                if (keep) skip = !kept(keep, idx);
                else skip = V->mask && AT(V->mask, zz, yy, x0 + dx) == 0;
                if (skip) {
                    gbuf[idx] = NAN;
                    idx += 1;
                    continue;
                }
                n_ref += 1;
                double X = (double)dx;
                double xw = a00 * X + a01 * Y + a02 * Z + tx;
                double yw = a10 * X + a11 * Y + a12 * Z + ty;
                double zw = a20 * X + a21 * Y + a22 * Z + tz;
                if (!margin_ok(zw, yw, xw, V)) return r; /* ok=0, n_valid=0: out of bounds */
                double val = sample_volume(V->g, V, zw, yw, xw, mode);
                gbuf[idx] = val;
                idx += 1;
                if (val != val) continue;
                double fv = (double)AT(V->f, zz, yy, x0 + dx);
                n_valid += 1;
                s1 += val;
                s2 += val * val;
                s1f += fv;
                s2f += fv * fv;
            }
        }
    }
    r.n_valid = n_valid;
    if (n_valid < MIN_SUBSET_VOXELS || (double)n_valid < MIN_VALID_FRACTION * (double)n_ref) return r;
    double meang = s1 / (double)n_valid;
    double ssg = s2 - (double)n_valid * meang * meang;
    if (ssg < 0.0) ssg = 0.0;
    double meanf = s1f / (double)n_valid;
    double ssf = s2f - (double)n_valid * meanf * meanf;
    if (ssf < 0.0) ssf = 0.0;
    r.ok = 1;
    r.meang = meang;
    r.bottomg = sqrt(ssg > 1e-30 ? ssg : 1e-30);
    r.meanf = meanf;
    r.bottomf = sqrt(ssf > 1e-30 ? ssf : 1e-30);
    return r;
}

/* numba_kernels.py:601-616 (_zncc_from_buffer) */
static double zncc_from_buffer(int64_t x0, int64_t y0, int64_t z0, int hx, int hy, int hz, int stride,
                               const vols_t *V, const double *gbuf, const wstat_t *w)
{
    double s = 0.0;
    int64_t idx = 0;
    for (int dz = -hz; dz <= hz; dz += stride)
        for (int dy = -hy; dy <= hy; dy += stride)
            for (int dx = -hx; dx <= hx; dx += stride) {
                double gv = gbuf[idx];
                if (gv == gv)
                    s += ((double)AT(V->f, z0 + dz, y0 + dy, x0 + dx) - w->meanf) * (gv - w->meang);
                idx += 1;
            }
    return s / (w->bottomf * w->bottomg);
}

/* ------------------------------------------------------------------ 12-DOF IC-GN */

/* numba_kernels.py:624-792 (_icgn_12dof_single). Returns status; n_iter, zncc through pointers. */
static int icgn12_single(double *P, int64_t x0, int64_t y0, int64_t z0, int hx, int hy, int hz,
                         int stride, const vols_t *V, int mode, const double *L, double tol,
                         double dp_tol, int max_iter, int patience, double *gbuf, const double *H,
                         const double *pattern, double noise_gain, int predictive,
                         const uint8_t *keep, int *n_iter, double *zncc_o)
{
    double b[12], dP[12], negb[12], H0[144], L0[144], P_best[12];
    double n_full = pattern ? pattern[9 * 12 + 9] : 0.0;
    double last_dp = 1e300, dp_prev = -1.0, norm_init = -1.0;
    int hm = hy > hz ? hy : hz;
    double half_scale = (double)(hx > hm ? hx : hm);
    double zncc = NAN, best_zncc = -2.0, best_dp = 1e300;
    int stall = 0;
    for (int k = 0; k < 12; ++k) P_best[k] = P[k];
    for (int it = 1; it <= max_iter; ++it) {
        *n_iter = it;
        wstat_t w = warp_and_sample(P, x0, y0, z0, hx, hy, hz, stride, V, mode, gbuf, keep);
        if (!w.ok) {
            *zncc_o = NAN;
            return w.n_valid == 0 ? ST_OUT_OF_BOUNDS : ST_INVALID_SUBSET;
        }
        for (int k = 0; k < 12; ++k) b[k] = 0.0;
        double inv_bf = 1.0 / w.bottomf, inv_bg = 1.0 / w.bottomg, s_cc = 0.0;
        int64_t idx = 0;
        for (int dz = -hz; dz <= hz; dz += stride) {
            double Z = (double)dz;
            int64_t zz = z0 + dz;
            for (int dy = -hy; dy <= hy; dy += stride) {
                double Y = (double)dy;
                int64_t yy = y0 + dy;
                for (int dx = -hx; dx <= hx; dx += stride) {
                    int64_t xx = x0 + dx;
                    double gv = gbuf[idx];
                    if (gv != gv) {
                        idx += 1;
                        continue;
                    }
                    double X = (double)dx;
                    double fd = (double)AT(V->f, zz, yy, xx) - w.meanf;
                    double gd = gv - w.meang;
                    s_cc += fd * gd;
                    double res = fd * inv_bf - gd * inv_bg;
                    idx += 1;
                    double dfx, dfy, dfz;
                    grad_at(V, zz, yy, xx, &dfx, &dfy, &dfz);
                    double gxv = dfx * res, gyv = dfy * res, gzv = dfz * res;
                    b[0] += gxv * X; b[1] += gxv * Y; b[2] += gxv * Z;
                    b[3] += gyv * X; b[4] += gyv * Y; b[5] += gyv * Z;
                    b[6] += gzv * X; b[7] += gzv * Y; b[8] += gzv * Z;
                    b[9] += gxv; b[10] += gyv; b[11] += gzv;
                }
            }
        }
        double norm_abs = 0.0;
        for (int k = 0; k < 12; ++k) {
            b[k] *= w.bottomf;
            norm_abs += b[k] * b[k];
        }
        norm_abs = sqrt(norm_abs);
        if (!isfinite(norm_abs)) {
            *zncc_o = NAN;
            return ST_NAN;
        }
        if (norm_init < 0.0) norm_init = norm_abs;
        double norm_rel = norm_init > 1e-300 ? norm_abs / norm_init : 0.0;
        zncc = s_cc * inv_bf * inv_bg;
        *zncc_o = zncc;
        if (norm_rel < tol || norm_abs < ABS_TOL) return ST_CONVERGED;
        for (int k = 0; k < 12; ++k) negb[k] = -b[k];
        int use_corrected = 0;
        if (noise_gain > 0.0 && n_full > 0.0 && last_dp < NOISE_CORR_STEP && H && pattern) {
            double om = 1.0 - zncc;
            double s2 = (om > 0.0 ? om : 0.0) * w.bottomf * w.bottomf / (double)w.n_valid;
            double corr = noise_gain * s2 * ((double)w.n_valid / n_full);
            use_corrected = corrected_cholesky12(H, pattern, corr, H0, L0);
        }
        chol_solve12(use_corrected ? L0 : L, negb, dP);
        double dp_norm = 0.0;
        for (int k = 0; k < 9; ++k) {
            double v = dP[k] * half_scale;
            dp_norm += v * v;
        }
        for (int k = 9; k < 12; ++k) dp_norm += dP[k] * dP[k];
        dp_norm = sqrt(dp_norm);
        if (!isfinite(dp_norm)) {
            *zncc_o = NAN;
            return ST_NAN;
        }
        last_dp = dp_norm;
        if (dp_norm < dp_tol) return ST_CONVERGED;
        if (predictive && dp_prev > 0.0 && dp_norm < PREDICT_CONTRACTION * dp_prev
            && dp_norm * dp_norm < dp_tol * dp_prev) {
            if (!orc_compose_warp(P, dP)) return ST_SINGULAR;
            return ST_CONVERGED;
        }
        dp_prev = dp_norm;
        int improved = zncc > best_zncc + STALL_ZNCC_EPS;
        if (improved) {
            best_zncc = zncc;
            for (int k = 0; k < 12; ++k) P_best[k] = P[k];
        }
        if (dp_norm < best_dp * STALL_STEP_DECAY) {
            best_dp = dp_norm;
            improved = 1;
        }
        if (improved)
            stall = 0;
        else {
            stall += 1;
            if (patience > 0 && stall >= patience) {
                for (int k = 0; k < 12; ++k) P[k] = P_best[k];
                *zncc_o = best_zncc;
                return ST_STALLED;
            }
        }
        if (!orc_compose_warp(P, dP)) return ST_SINGULAR;
    }
    *n_iter = max_iter;
    return ST_MAX_ITER;
}

/* numba_kernels.py:795-901 (_icgn_12dof_parallel_jit). P (N,12) is updated in place (the
 * caller passes the copy of P0); n_iter/status/zncc are written for every node. H_all and
 * pattern may be NULL when noise_gain == 0. */
void orc_icgn12(const int64_t *coords, int64_t N, double *P, int hx, int hy, int hz, int stride,
                const float *f, const float *gx, const float *gy, const float *gz,
                const uint8_t *mask, const float *g, int64_t nz, int64_t ny, int64_t nx, int mode,
                const double *L_all, const double *H_all, const double *pattern, double noise_gain,
                const uint8_t *valid, double tol, double dp_tol, int max_iter, int patience,
                int predictive, const int64_t *split_index, const uint8_t *split_keep,
                int64_t keep_bytes, int32_t *n_iter, int8_t *status, double *zncc)
{
    vols_t Vs = {f, gx, gy, gz, g, mask, nz, ny, nx};
    int64_t S = subset_count(hx, hy, hz, stride);
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
    int64_t M = 0;
    for (int64_t n = 0; n < N; ++n) {
        n_iter[n] = 0;
        status[n] = valid[n] ? ST_SKIPPED : ST_INVALID_SUBSET;
        zncc[n] = NAN;
        if (valid[n]) idx[M++] = n;
    }
    if (!(H_all && pattern && noise_gain > 0.0)) noise_gain = 0.0;
    int64_t Lr = (M + SCHEDULE_LANES - 1) / SCHEDULE_LANES;
#pragma omp parallel
    {
        double *gbuf = (double *)malloc(sizeof(double) * S);
#pragma omp for schedule(static)
        for (int64_t kk = 0; kk < SCHEDULE_LANES * Lr; ++kk) {
            /* block-cyclic schedule over 64 stripes: numba_kernels.py:843-851 */
            int64_t p = (kk % SCHEDULE_LANES) * Lr + kk / SCHEDULE_LANES;
            if (p >= M) continue;
            int64_t n = idx[p];
            double *Pn = P + 12 * n;
            int finite = 1;
            for (int k = 0; k < 12; ++k)
                if (!isfinite(Pn[k])) finite = 0;
            if (!finite) {
                status[n] = ST_NAN;
                continue;
            }
            int it = 0;
            double zc = NAN;
            int st = icgn12_single(Pn, coords[3 * n], coords[3 * n + 1], coords[3 * n + 2], hx, hy, hz,
                                   stride, &Vs, mode, L_all + 144 * n, tol, dp_tol, max_iter, patience,
                                   gbuf, noise_gain > 0.0 ? H_all + 144 * n : NULL,
                                   noise_gain > 0.0 ? pattern : NULL, noise_gain, predictive,
                                   keep_row(split_index, split_keep, keep_bytes, n), &it, &zc);
            n_iter[n] = it;
            status[n] = (int8_t)st;
            zncc[n] = zc;
        }
        free(gbuf);
    }
    free(idx);
}

/* ------------------------------------------------------------------ 3-DOF IC-GN */

/* numba_kernels.py:910-928 (_build_h3) */
static int build_h3(const double *H, double bf2, double mu, double corr, double *Hinv)
{
    double H3[9], Hd[9];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) H3[3 * i + j] = H[(9 + i) * 12 + 9 + j] * 2.0 / bf2;
        double limit_i = (1.0 - MIN_CORRECTED_FRACTION) * H[(9 + i) * 12 + 9 + i];
        double c = corr < limit_i ? corr : limit_i;
        H3[3 * i + i] -= c * 2.0 / bf2;
        H3[3 * i + i] += mu;
    }
    double md = H3[4] > H3[8] ? H3[4] : H3[8];
    double max_diag = H3[0] > md ? H3[0] : md;
    if (!(max_diag > 0.0)) return 0;
    for (int i = 0; i < 9; ++i) Hd[i] = H3[i];
    for (int i = 0; i < 3; ++i) Hd[3 * i + i] += LM_DAMPING_3DOF * max_diag;
    return inv3(Hd, Hinv) != 0.0;
}

/* numba_kernels.py:931-1082 (_icgn_3dof_single) */
static int icgn3_single(double *P, int64_t x0, int64_t y0, int64_t z0, int hx, int hy, int hz,
                        int stride, const vols_t *V, int mode, const double *H, double bottomf,
                        const double *u_old, const double *vdual, double mu, double tol,
                        double dp_tol, int max_iter, int patience, double *gbuf, double noise_gain,
                        int predictive, const uint8_t *keep, int *n_iter, double *zncc_o)
{
    double best_dn = 1e300;
    int stall = 0;
    double bf2 = bottomf * bottomf;
    if (bf2 < 1e-30) bf2 = 1e-30;
    double Hinv[9];
    *n_iter = 0;
    *zncc_o = NAN;
    if (!build_h3(H, bf2, mu, 0.0, Hinv)) return ST_SINGULAR;
    double last_dn = 1e300, dn_prev = -1.0, norm_init = -1.0, zncc = NAN;
    double b[3], tb[3], dt[3];
    double a00 = 1.0 + P[0], a01 = P[1], a02 = P[2];
    double a10 = P[3], a11 = 1.0 + P[4], a12 = P[5];
    double a20 = P[6], a21 = P[7], a22 = 1.0 + P[8];
    for (int it = 1; it <= max_iter; ++it) {
        *n_iter = it;
        wstat_t w = warp_and_sample(P, x0, y0, z0, hx, hy, hz, stride, V, mode, gbuf, keep);
        if (!w.ok) {
            *zncc_o = NAN;
            return w.n_valid == 0 ? ST_OUT_OF_BOUNDS : ST_INVALID_SUBSET;
        }
        b[0] = b[1] = b[2] = 0.0;
        double inv_bf = 1.0 / w.bottomf, inv_bg = 1.0 / w.bottomg, s_cc = 0.0;
        int64_t idx = 0;
        for (int dz = -hz; dz <= hz; dz += stride) {
            int64_t zz = z0 + dz;
            for (int dy = -hy; dy <= hy; dy += stride) {
                int64_t yy = y0 + dy;
                for (int dx = -hx; dx <= hx; dx += stride) {
                    int64_t xx = x0 + dx;
                    double gv = gbuf[idx];
                    if (gv != gv) {
                        idx += 1;
                        continue;
                    }
                    double fd = (double)AT(V->f, zz, yy, xx) - w.meanf;
                    double gd = gv - w.meang;
                    s_cc += fd * gd;
                    double res = fd * inv_bf - gd * inv_bg;
                    idx += 1;
                    double dfx, dfy, dfz;
                    grad_at(V, zz, yy, xx, &dfx, &dfy, &dfz);
                    b[0] += dfx * res;
                    b[1] += dfy * res;
                    b[2] += dfz * res;
                }
            }
        }
        double norm_abs = 0.0;
        for (int k = 0; k < 3; ++k) {
            b[k] *= w.bottomf;
            tb[k] = b[k] * 2.0 / bf2 + mu * (P[9 + k] - u_old[k] - vdual[k]);
            norm_abs += tb[k] * tb[k];
        }
        norm_abs = sqrt(norm_abs);
        if (!isfinite(norm_abs)) {
            *zncc_o = NAN;
            return ST_NAN;
        }
        if (norm_init < 0.0) norm_init = norm_abs;
        double norm_rel = norm_init > 1e-300 ? norm_abs / norm_init : 0.0;
        zncc = s_cc * inv_bf * inv_bg;
        *zncc_o = zncc;
        if (norm_rel < tol || norm_abs < mu * 1e-4) return ST_CONVERGED;
        if (noise_gain > 0.0 && last_dn < NOISE_CORR_STEP) {
            double om = 1.0 - zncc;
            double s2 = (om > 0.0 ? om : 0.0) * w.bottomf * w.bottomf / (double)w.n_valid;
            if (!build_h3(H, bf2, mu, noise_gain * s2 * (double)w.n_valid, Hinv))
                build_h3(H, bf2, mu, 0.0, Hinv);
        }
        for (int i = 0; i < 3; ++i) {
            double s = 0.0;
            for (int j = 0; j < 3; ++j) s += Hinv[3 * i + j] * tb[j];
            dt[i] = -s;
        }
        double dn = sqrt(dt[0] * dt[0] + dt[1] * dt[1] + dt[2] * dt[2]);
        if (!isfinite(dn)) {
            *zncc_o = NAN;
            return ST_NAN;
        }
        last_dn = dn;
        if (dn < dp_tol) return ST_CONVERGED;
        if (predictive && dn_prev > 0.0 && dn < PREDICT_CONTRACTION * dn_prev && dn * dn < dp_tol * dn_prev) {
            P[9] -= a00 * dt[0] + a01 * dt[1] + a02 * dt[2];
            P[10] -= a10 * dt[0] + a11 * dt[1] + a12 * dt[2];
            P[11] -= a20 * dt[0] + a21 * dt[1] + a22 * dt[2];
            return ST_CONVERGED;
        }
        dn_prev = dn;
        if (dn < best_dn * STALL_STEP_DECAY) {
            best_dn = dn;
            stall = 0;
        } else {
            stall += 1;
            if (patience > 0 && stall >= patience) return ST_STALLED;
        }
        P[9] -= a00 * dt[0] + a01 * dt[1] + a02 * dt[2];
        P[10] -= a10 * dt[0] + a11 * dt[1] + a12 * dt[2];
        P[11] -= a20 * dt[0] + a21 * dt[1] + a22 * dt[2];
    }
    *n_iter = max_iter;
    return ST_MAX_ITER;
}

/* numba_kernels.py:1085-1208 (_icgn_3dof_parallel_jit). U (N,3) holds U_old on entry and the
 * result on exit (the caller passes the copy). n_full only matters through noise_gain. */
void orc_icgn3(const int64_t *coords, int64_t N, const double *U_old, double *U, const double *F_fixed,
               const double *vdual, int hx, int hy, int hz, int stride, const float *f,
               const float *gx, const float *gy, const float *gz, const uint8_t *mask, const float *g,
               int64_t nz, int64_t ny, int64_t nx, int mode, const double *H_all,
               const double *bottomf_all, const uint8_t *valid, double mu, double tol, double dp_tol,
               int max_iter, int patience, double noise_gain, int predictive,
               const int64_t *split_index, const uint8_t *split_keep, int64_t keep_bytes,
               int32_t *n_iter, int8_t *status, double *zncc)
{
    vols_t Vs = {f, gx, gy, gz, g, mask, nz, ny, nx};
    int64_t S = subset_count(hx, hy, hz, stride);
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
    int64_t M = 0;
    for (int64_t n = 0; n < N; ++n) {
        n_iter[n] = 0;
        status[n] = valid[n] ? ST_SKIPPED : ST_INVALID_SUBSET;
        zncc[n] = NAN;
        if (valid[n]) idx[M++] = n;
    }
    int64_t Lr = (M + SCHEDULE_LANES - 1) / SCHEDULE_LANES;
#pragma omp parallel
    {
        double *gbuf = (double *)malloc(sizeof(double) * S);
#pragma omp for schedule(static)
        for (int64_t kk = 0; kk < SCHEDULE_LANES * Lr; ++kk) {
            int64_t p = (kk % SCHEDULE_LANES) * Lr + kk / SCHEDULE_LANES;
            if (p >= M) continue;
            int64_t n = idx[p];
            double P[12];
            int finite = 1;
            for (int k = 0; k < 9; ++k) {
                P[k] = F_fixed[9 * n + k];
                if (!isfinite(P[k])) finite = 0;
            }
            for (int k = 0; k < 3; ++k) {
                P[9 + k] = U_old[3 * n + k];
                if (!isfinite(P[9 + k])) finite = 0;
            }
            if (!finite) {
                status[n] = ST_NAN;
                continue;
            }
            int it = 0;
            double zc = NAN;
            int st = icgn3_single(P, coords[3 * n], coords[3 * n + 1], coords[3 * n + 2], hx, hy, hz,
                                  stride, &Vs, mode, H_all + 144 * n, bottomf_all[n], U_old + 3 * n,
                                  vdual + 3 * n, mu, tol, dp_tol, max_iter, patience, gbuf, noise_gain,
                                  predictive, keep_row(split_index, split_keep, keep_bytes, n), &it, &zc);
            for (int k = 0; k < 3; ++k) U[3 * n + k] = P[9 + k];
            n_iter[n] = it;
            status[n] = (int8_t)st;
            zncc[n] = zc;
        }
        free(gbuf);
    }
    free(idx);
}

/* numba_kernels.py:1354-1387 (evaluate_zncc_parallel): NaN where invalid or not sampled */
void orc_evaluate_zncc(const int64_t *coords, int64_t N, const double *P_all, int hx, int hy, int hz,
                       int stride, const float *f, const uint8_t *mask, const float *g, int64_t nz,
                       int64_t ny, int64_t nx, int mode, const uint8_t *valid, double *out)
{
    vols_t Vs = {f, NULL, NULL, NULL, g, mask, nz, ny, nx};
    int64_t S = subset_count(hx, hy, hz, stride);
#pragma omp parallel
    {
        double *gbuf = (double *)malloc(sizeof(double) * S);
#pragma omp for schedule(static)
        for (int64_t n = 0; n < N; ++n) {
            out[n] = NAN;
            if (!valid[n]) continue;
            int64_t x0 = coords[3 * n], y0 = coords[3 * n + 1], z0 = coords[3 * n + 2];
            wstat_t w = warp_and_sample(P_all + 12 * n, x0, y0, z0, hx, hy, hz, stride, &Vs, mode, gbuf, NULL);
            if (!w.ok) continue;
            out[n] = zncc_from_buffer(x0, y0, z0, hx, hy, hz, stride, &Vs, gbuf, &w);
        }
        free(gbuf);
    }
}

/* ------------------------------------------------------------------ integer NCC map */

/* integer_search.py:44-151 (_zncc_direct): out (N, vz, vy, vx) float64 */
void orc_zncc_direct(const float *f, const float *g, int64_t nz, int64_t ny, int64_t nx,
                     const int64_t *coords, int64_t N, int hx, int hy, int hz, int rx, int ry, int rz,
                     const int64_t *wx0, const int64_t *wy0, const int64_t *wz0, const uint8_t *t_ok,
                     double *out)
{
    vols_t Vs = {0};
    Vs.nz = nz; Vs.ny = ny; Vs.nx = nx;
    const vols_t *V = &Vs;
    int sx = 2 * hx + 1, sy = 2 * hy + 1, sz = 2 * hz + 1;
    int vx = 2 * rx + 1, vy = 2 * ry + 1, vz = 2 * rz + 1;
    int wx = sx + 2 * rx, wy = sy + 2 * ry, wz = sz + 2 * rz;
    double n_tpl = (double)sx * sy * sz;
    int64_t n_off = (int64_t)vx * vy * vz;
    int64_t tsz = (int64_t)(wz + 1) * (wy + 1) * (wx + 1);
#pragma omp parallel
    {
        double *T = (double *)malloc(sizeof(double) * sx * sy * sz);
        double *S1 = (double *)malloc(sizeof(double) * tsz);
        double *S2 = (double *)malloc(sizeof(double) * tsz);
#pragma omp for schedule(dynamic, 1)
        for (int64_t n = 0; n < N; ++n) {
            double *o = out + n * n_off;
            if (!t_ok[n]) {
                for (int64_t i = 0; i < n_off; ++i) o[i] = -2.0;
                continue;
            }
            int64_t x0 = coords[3 * n], y0 = coords[3 * n + 1], z0 = coords[3 * n + 2];
            double sf = 0.0;
            for (int k = 0; k < sz; ++k)
                for (int j = 0; j < sy; ++j)
                    for (int i = 0; i < sx; ++i) {
                        double v = (double)AT(f, z0 - hz + k, y0 - hy + j, x0 - hx + i);
                        T[(k * sy + j) * sx + i] = v;
                        sf += v;
                    }
            double mf = sf / n_tpl, ssf = 0.0;
            for (int64_t t = 0; t < (int64_t)sx * sy * sz; ++t) {
                T[t] -= mf;
                ssf += T[t] * T[t];
            }
            if (ssf < 1e-20) {
                for (int64_t i = 0; i < n_off; ++i) o[i] = -2.0;
                continue;
            }
            memset(S1, 0, sizeof(double) * tsz);
            memset(S2, 0, sizeof(double) * tsz);
            int64_t zb = wz0[n], yb = wy0[n], xb = wx0[n];
#define SA(A, z, y, x) A[((int64_t)(z) * (wy + 1) + (y)) * (wx + 1) + (x)]
            for (int z = 0; z < wz; ++z)
                for (int y = 0; y < wy; ++y) {
                    double r1 = 0.0, r2 = 0.0;
                    for (int x = 0; x < wx; ++x) {
                        double v = (double)AT(g, zb + z, yb + y, xb + x);
                        r1 += v;
                        r2 += v * v;
                        SA(S1, z + 1, y + 1, x + 1) = SA(S1, z, y + 1, x + 1) + SA(S1, z + 1, y, x + 1) - SA(S1, z, y, x + 1) + r1;
                        SA(S2, z + 1, y + 1, x + 1) = SA(S2, z, y + 1, x + 1) + SA(S2, z + 1, y, x + 1) - SA(S2, z, y, x + 1) + r2;
                    }
                }
            for (int dz = 0; dz < vz; ++dz)
                for (int dy = 0; dy < vy; ++dy)
                    for (int dx = 0; dx < vx; ++dx) {
                        int z1 = dz + sz, y1 = dy + sy, x1 = dx + sx;
                        double sg = SA(S1, z1, y1, x1) - SA(S1, dz, y1, x1) - SA(S1, z1, dy, x1) - SA(S1, z1, y1, dx)
                                  + SA(S1, dz, dy, x1) + SA(S1, dz, y1, dx) + SA(S1, z1, dy, dx) - SA(S1, dz, dy, dx);
                        double sgg = SA(S2, z1, y1, x1) - SA(S2, dz, y1, x1) - SA(S2, z1, dy, x1) - SA(S2, z1, y1, dx)
                                   + SA(S2, dz, dy, x1) + SA(S2, dz, y1, dx) + SA(S2, z1, dy, dx) - SA(S2, dz, dy, dx);
                        double var = sgg - sg * sg / n_tpl;
                        int64_t oi = ((int64_t)dz * vy + dy) * vx + dx;
                        if (var < 1e-20) {
                            o[oi] = 0.0;
                            continue;
                        }
                        double sfg = 0.0;
                        for (int k = 0; k < sz; ++k)
                            for (int j = 0; j < sy; ++j) {
                                const float *gr = g + ((zb + dz + k) * ny + (yb + dy + j)) * nx + (xb + dx);
                                const double *tr = T + ((int64_t)k * sy + j) * sx;
                                for (int i = 0; i < sx; ++i) sfg += tr[i] * (double)gr[i];
                            }
                        double val = sfg / sqrt(var * ssf);
                        if (val > 1.0) val = 1.0;
                        else if (val < -1.0) val = -1.0;
                        o[oi] = val;
                    }
#undef SA
        }
        free(T);
        free(S1);
        free(S2);
    }
}

/* ------------------------------------------------------------------ subset splitting */

/* utils/flood_fill.py:12-72 (flood_fill_centre): the 6-connected component of `sub` that holds
 * the centre voxel, written into `out` (zeroed by the caller); queue holds Sz*Sy*Sx entries. */
static void flood_fill_centre(const uint8_t *sub, uint8_t *out, int64_t *queue, int Sz, int Sy, int Sx)
{
    int cz = Sz / 2, cy = Sy / 2, cx = Sx / 2;
    int64_t c = ((int64_t)cz * Sy + cy) * Sx + cx;
    if (!sub[c]) return;
    int64_t head = 0, tail = 0;
    out[c] = 1;
    queue[tail++] = c;
    while (head < tail) {
        int64_t v = queue[head++];
        int z = (int)(v / ((int64_t)Sy * Sx)), rem = (int)(v - (int64_t)z * Sy * Sx);
        int y = rem / Sx, x = rem - y * Sx;
        int nb[6][3] = {{z - 1, y, x}, {z + 1, y, x}, {z, y - 1, x}, {z, y + 1, x}, {z, y, x - 1}, {z, y, x + 1}};
        for (int q = 0; q < 6; ++q) {
            int zz = nb[q][0], yy = nb[q][1], xx = nb[q][2];
            if (zz < 0 || zz >= Sz || yy < 0 || yy >= Sy || xx < 0 || xx >= Sx) continue;
            int64_t u = ((int64_t)zz * Sy + yy) * Sx + xx;
            if (sub[u] && !out[u]) {
                out[u] = 1;
                queue[tail++] = u;
            }
        }
    }
}

/* numba_kernels.py:223-300 (_build_split_rows_jit): rows (M, keep_bytes) zeroed by the caller */
void orc_split_rows(const int64_t *coords, const int64_t *cand, int64_t M, int hx, int hy, int hz,
                    int stride, const uint8_t *mask, int64_t nz, int64_t ny, int64_t nx, int margin,
                    int64_t keep_bytes, uint8_t *rows, int64_t *n_keep, int64_t *n_inmask)
{
    int Sz = 2 * hz + 1, Sy = 2 * hy + 1, Sx = 2 * hx + 1;
    int64_t W = (int64_t)Sz * Sy * Sx;
#pragma omp parallel
    {
        uint8_t *sub = (uint8_t *)malloc(W);
        uint8_t *out = (uint8_t *)malloc(W);
        int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * W);
#pragma omp for schedule(dynamic, 4)
        for (int64_t m = 0; m < M; ++m) {
            int64_t n = cand[m];
            int64_t x0 = coords[3 * n], y0 = coords[3 * n + 1], z0 = coords[3 * n + 2];
            memset(sub, 0, W);
            memset(out, 0, W);
            for (int iz = 0; iz < Sz; ++iz) {
                int64_t zz = z0 - hz + iz;
                if (zz < margin || zz >= nz - margin) continue;
                for (int iy = 0; iy < Sy; ++iy) {
                    int64_t yy = y0 - hy + iy;
                    if (yy < margin || yy >= ny - margin) continue;
                    for (int ix = 0; ix < Sx; ++ix) {
                        int64_t xx = x0 - hx + ix;
                        if (xx < margin || xx >= nx - margin) continue;
                        if (mask[(zz * ny + yy) * nx + xx] != 0) sub[((int64_t)iz * Sy + iy) * Sx + ix] = 1;
                    }
                }
            }
            flood_fill_centre(sub, out, queue, Sz, Sy, Sx);
            uint8_t *row = rows + m * keep_bytes;
            int64_t i = 0, nk = 0, nm = 0;
            for (int dz = -hz; dz <= hz; dz += stride)
                for (int dy = -hy; dy <= hy; dy += stride)
                    for (int dx = -hx; dx <= hx; dx += stride) {
                        int64_t o = ((int64_t)(dz + hz) * Sy + (dy + hy)) * Sx + (dx + hx);
                        if (sub[o]) nm += 1;
                        if (out[o]) {
                            row[i >> 3] = (uint8_t)(row[i >> 3] | (1u << (i & 7)));
                            nk += 1;
                        }
                        i += 1;
                    }
            n_keep[m] = nk;
            n_inmask[m] = nm;
        }
        free(sub);
        free(out);
        free(queue);
    }
}
