/*
 * CPU port of algotom's FBP path -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only to check results against or as the CPU baseline.
 * The product (algotom_amd) never links it.
 *
 * Restates, with numba's typing (algotom 1.6.0):
 *   orc_backproject  <- algotom/rec/reconstruction.py:307-359 (CPU kernel),
 *                       :156-190 (2D GPU kernel), :229-264 (chunk kernel)
 *   orc_ramp_filter  <- reconstruction.py:118-153 (pad, windowed ramp,
 *                       real part, crop) as the equivalent circular
 *                       convolution, evaluated with a power-of-two FFT
 *   orc_trig_tables  <- reconstruction.py:169-171 / :334-336
 *   orc_forward_project  the transpose of orc_backproject's float64-sum
 *                       form (no reference counterpart: the adjoint that
 *                       differentiable FBP needs), as a scatter over
 *                       pixels per angle
 *
 * Deterministic and race-free: OpenMP runs over image rows, each pixel sums
 * its angles in ascending order (the reference's prange over angles at
 * reconstruction.py:333 races on recon[y, x]).  Build with
 * -ffp-contract=off: numba does not contract a*b+c into an FMA.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_ACC_F64 = 0, ORC_ACC_F32SEQ = 1 };
enum { ORC_PAD_EDGE = 0, ORC_PAD_CONSTANT = 1, ORC_PAD_REFLECT = 2,
       ORC_PAD_SYMMETRIC = 3, ORC_PAD_WRAP = 4 };

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* float32 cos/sin of float32 -angle, evaluated in double and rounded. */
void orc_trig_tables(const float* angles, int n, float* cos_t, float* sin_t) {
    for (int i = 0; i < n; ++i) {
        float th = -angles[i];
        cos_t[i] = (float)cos((double)th);
        sin_t[i] = (float)sin((double)th);
    }
}

/*
 * Back-projection of nslice sinograms stored (nrow, nslice, ncol), output
 * (ncol, nslice, ncol) = recon[y, j, x] (the reference's 3D layout; a 2D
 * call is nslice = 1).  acc = ORC_ACC_F64: float64 accumulator rounded once
 * (2D GPU kernel); ORC_ACC_F32SEQ: float32 storage after every angle (CPU and
 * chunk kernels).
 */
void orc_backproject(const float* sino, int nrow, int nslice, int ncol,
                     const float* cos_t, const float* sin_t, float center,
                     int edge_pad, int acc, float* recon, int nthreads) {
    set_threads(nthreads);
    const int w1 = ncol - 1;
    const double icenter = 0.5 * (double)w1;
    const double cen = (double)center;
    const long long sstride = (long long)nslice * ncol;
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < ncol; ++y) {
        const double yc = (double)y - icenter;
        double* a64 = (double*)malloc(sizeof(double) * (size_t)ncol * nslice);
        float* a32 = (float*)malloc(sizeof(float) * (size_t)ncol * nslice);
        for (long long k = 0; k < (long long)ncol * nslice; ++k) {
            a64[k] = 0.0;
            a32[k] = 0.0f;
        }
        for (int i = 0; i < nrow; ++i) {
            const double c = (double)cos_t[i];
            const double s = (double)sin_t[i];
            const double ys = yc * s;
            const float* srow = sino + (long long)i * sstride;
            for (int x = 0; x < ncol; ++x) {
                const double xc = (double)x - icenter;
                const double f = (xc * c + ys) + cen;
                int inside = (f >= 0.0) && (f <= (double)w1);
                if (!inside && !edge_pad) continue;
                int d = 0;
                double t = 0.0;
                int interp = 0;
                if (inside) {
                    const double fl = floor(f);
                    d = (int)fl;
                    interp = (ceil(f) != fl);
                    t = f - fl;
                }
                for (int j = 0; j < nslice; ++j) {
                    const float* r = srow + (long long)j * ncol;
                    double val;
                    if (inside) {
                        if (interp) {
                            const float yd = r[d];
                            const float dl = r[d + 1] - yd;  /* float32 */
                            val = (double)yd + (double)dl * t;
                        } else {
                            val = (double)r[d];
                        }
                    } else {
                        val = (double)(f < 0.0 ? r[0] : r[w1]);
                    }
                    const long long k = (long long)j * ncol + x;
                    if (acc == ORC_ACC_F64) a64[k] += val;
                    else a32[k] = (float)((double)a32[k] + val);
                }
            }
        }
        for (int j = 0; j < nslice; ++j)
            for (int x = 0; x < ncol; ++x) {
                const long long k = (long long)j * ncol + x;
                recon[((long long)y * nslice + j) * ncol + x] =
                    (acc == ORC_ACC_F64) ? (float)a64[k] : a32[k];
            }
        free(a64);
        free(a32);
    }
}

/*
 * Transpose of the back-projection: for every angle i, pixel (x, y) with
 * f = (x - ic) c + (y - ic) s + center in [0, ncol-1], d = floor f,
 * t = f - d, the masked pixel value v adds (1 - t) v to bin d and t v to
 * bin d + 1 (when d + 1 < ncol) of row i -- the weights the
 * back-projector gives s[d] and s[d+1] (reconstruction.py:170-181).
 * image (ncol, nslice, ncol) float32, sino out (nrow, nslice, ncol)
 * float32 = scale * sum (float64 sums over the pixels in row-major order).
 * Parallel over angles: each angle owns its output row.
 */
void orc_forward_project(const float* image, int nrow, int nslice, int ncol,
                         const float* cos_t, const float* sin_t, float center,
                         double mask_radius, double scale, float* sino,
                         int nthreads) {
    set_threads(nthreads);
    const int w1 = ncol - 1;
    const double icenter = 0.5 * (double)w1;
    const double cen = (double)center;
    const int mc = ncol / 2;
    const double r2 = mask_radius * mask_radius;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nrow; ++i) {
        double* acc = (double*)calloc((size_t)nslice * ncol, sizeof(double));
        const double c = (double)cos_t[i];
        const double s = (double)sin_t[i];
        for (int y = 0; y < ncol; ++y) {
            const double ys = ((double)y - icenter) * s;
            for (int x = 0; x < ncol; ++x) {
                if (mask_radius >= 0.0) {
                    const long long dx = x - mc, dy = y - mc;
                    if (!((double)(dx * dx + dy * dy) <= r2)) continue;
                }
                const double f = (((double)x - icenter) * c + ys) + cen;
                if (!(f >= 0.0 && f <= (double)w1)) continue;
                const double fl = floor(f);
                const int d = (int)fl;
                const double t = f - fl;
                for (int j = 0; j < nslice; ++j) {
                    const double v =
                        (double)image[((long long)y * nslice + j) * ncol + x];
                    acc[(long long)j * ncol + d] += (1.0 - t) * v;
                    if (d + 1 <= w1) acc[(long long)j * ncol + d + 1] += t * v;
                }
            }
        }
        for (int j = 0; j < nslice; ++j)
            for (int k = 0; k < ncol; ++k)
                sino[((long long)i * nslice + j) * ncol + k] =
                    (float)(acc[(long long)j * ncol + k] * scale);
        free(acc);
    }
}

/* All-float64 back-projection (no float32 rounding of data, trig or sums):
 * the reference's formula in exact-as-possible arithmetic, for reporting
 * errors.  sino (nrow, nslice, ncol) double; recon (ncol, nslice, ncol). */
void orc_backproject_f64(const double* sino, int nrow, int nslice, int ncol,
                         const double* cos_t, const double* sin_t,
                         double center, double* recon, int nthreads) {
    set_threads(nthreads);
    const int w1 = ncol - 1;
    const double icenter = 0.5 * (double)w1;
    const long long sstride = (long long)nslice * ncol;
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < ncol; ++y) {
        const double yc = (double)y - icenter;
        double* acc = (double*)calloc((size_t)ncol * nslice, sizeof(double));
        for (int i = 0; i < nrow; ++i) {
            const double c = cos_t[i], s = sin_t[i];
            const double* srow = sino + (long long)i * sstride;
            for (int x = 0; x < ncol; ++x) {
                const double f = ((double)x - icenter) * c + yc * s + center;
                if (!(f >= 0.0 && f <= (double)w1)) continue;
                const double fl = floor(f);
                const int d = (int)fl;
                const double t = f - fl;
                for (int j = 0; j < nslice; ++j) {
                    const double* r = srow + (long long)j * ncol;
                    const double v = t > 0.0 ? r[d] + (r[d + 1] - r[d]) * t
                                             : r[d];
                    acc[(long long)j * ncol + x] += v;
                }
            }
        }
        for (int j = 0; j < nslice; ++j)
            for (int x = 0; x < ncol; ++x)
                recon[((long long)y * nslice + j) * ncol + x] =
                    acc[(long long)j * ncol + x];
        free(acc);
    }
}

/* ------------------------------------------------------------------------
 * Filter: out = crop(real(pad(row) (*)_W g)), g the real kernel.
 * Circular convolution of length W evaluated as a linear convolution with
 * the doubled kernel [g, g] on a power-of-two FFT of length L >= 2W - pad.
 * --------------------------------------------------------------------- */
static void fft_inplace(double* re, double* im, int n, int inverse,
                        const double* twr, const double* twi) {
    /* iterative radix-2, bit-reversed input permutation */
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        const int step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; ++k) {
                const double wr = twr[k * step];
                const double wi = inverse ? -twi[k * step] : twi[k * step];
                const int a = i + k, b = i + k + half;
                const double xr = re[b] * wr - im[b] * wi;
                const double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
        }
    }
}

static int pad_src(int i, int n, int mode) {
    /* i = padded index - pad; returns source column or -1 for zero */
    if (i >= 0 && i < n) return i;
    switch (mode) {
    case ORC_PAD_EDGE: return i < 0 ? 0 : n - 1;
    case ORC_PAD_CONSTANT: return -1;
    case ORC_PAD_WRAP: { int j = i % n; return j < 0 ? j + n : j; }
    case ORC_PAD_REFLECT: {
        if (n == 1) return 0;
        int p = 2 * (n - 1); int j = i % p; if (j < 0) j += p;
        return j < n ? j : p - j; }
    case ORC_PAD_SYMMETRIC: {
        int p = 2 * n; int j = i % p; if (j < 0) j += p;
        return j < n ? j : p - 1 - j; }
    }
    return -1;
}

/*
 * rows: nrows x ncol (float32 if in_f64 == 0, else float64), contiguous.
 * g: real kernel, length W = ncol + 2 pad.  out: nrows x ncol float64.
 * apply_log: x -> -log(x) evaluated in double and rounded to the input
 * type (the reference takes -np.log of the input array at :436).
 */
void orc_ramp_filter(const void* rows, int in_f64, int nrows, int ncol,
                     int pad, int pad_mode, int apply_log, const double* g,
                     double* out, int nthreads) {
    set_threads(nthreads);
    const int width = ncol + 2 * pad;
    int L = 1;
    while (L < 2 * width - pad || L < 2) L <<= 1;
    double* twr = (double*)malloc(sizeof(double) * L);
    double* twi = (double*)malloc(sizeof(double) * L);
    for (int k = 0; k < L; ++k) {
        twr[k] = cos(-2.0 * M_PI * (double)k / (double)L);
        twi[k] = sin(-2.0 * M_PI * (double)k / (double)L);
    }
    /* spectrum of the doubled kernel */
    double* gr = (double*)calloc(L, sizeof(double));
    double* gi = (double*)calloc(L, sizeof(double));
    for (int k = 0; k < width; ++k) { gr[k] = g[k]; gr[k + width] = g[k]; }
    fft_inplace(gr, gi, L, 0, twr, twi);
#pragma omp parallel
    {
        double* xr = (double*)malloc(sizeof(double) * L);
        double* xi = (double*)malloc(sizeof(double) * L);
#pragma omp for schedule(static)
        for (int r = 0; r < nrows; ++r) {
            for (int k = 0; k < L; ++k) { xr[k] = 0.0; xi[k] = 0.0; }
            for (int k = 0; k < width; ++k) {
                int src = pad_src(k - pad, ncol, pad_mode);
                double v = 0.0;
                if (src >= 0) {
                    if (in_f64) {
                        double x = ((const double*)rows)[(long long)r * ncol + src];
                        v = apply_log ? -log(x) : x;
                    } else {
                        float x = ((const float*)rows)[(long long)r * ncol + src];
                        v = apply_log ? (double)(float)(-log((double)x)) : (double)x;
                    }
                }
                xr[k] = v;
            }
            fft_inplace(xr, xi, L, 0, twr, twi);
            for (int k = 0; k < L; ++k) {
                const double a = xr[k] * gr[k] - xi[k] * gi[k];
                const double b = xr[k] * gi[k] + xi[k] * gr[k];
                xr[k] = a; xi[k] = b;
            }
            fft_inplace(xr, xi, L, 1, twr, twi);
            for (int c = 0; c < ncol; ++c)
                out[(long long)r * ncol + c] = xr[width + pad + c] / (double)L;
        }
        free(xr);
        free(xi);
    }
    free(twr); free(twi); free(gr); free(gi);
}

/* Mask (utility.py:160-182) and final scale (reconstruction.py:458-459,
 * :488-493) applied in float32 as numpy does: (v * mask) * pi32 / (n-1). */
void orc_finalize(float* recon, int ncol, int nslice, int nrow, double ratio,
                  int use_mask) {
    const int c = ncol / 2;
    const double radius = ratio * (double)c;
    const float pi32 = (float)M_PI;
    const float den = (float)(nrow - 1);
    for (int y = 0; y < ncol; ++y)
        for (int j = 0; j < nslice; ++j)
            for (int x = 0; x < ncol; ++x) {
                float* p = recon + ((long long)y * nslice + j) * ncol + x;
                float v = *p;
                if (use_mask) {
                    const long long dx = x - c, dy = y - c;
                    const float m = ((double)(dx * dx + dy * dy) <= radius * radius)
                                        ? 1.0f : 0.0f;
                    v = v * m;
                }
                v = v * pi32;
                *p = v / den;
            }
}
