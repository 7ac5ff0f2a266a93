// Half-acquisition (offset rotation axis) 360 -> 180 degree sinogram
// conversion on the GPU, the step before FBP for such scans.
//
// algotom's convert_sinogram_360_to_180 (algotom/prep/conversion.py:318-364)
// takes the first N//2+1 rows and the left-right flipped last N//2+1 rows of
// a 360-degree sinogram and stitches them with a linear ramp over `overlap`
// columns (stitch_image, conversion.py:76-144; weights from
// make_weight_matrix, :45-73): the second image is first scaled by the ratio
// of the mean intensities of the two overlap regions (norm=True).  For an
// integer overlap no sub-pixel shift is involved.  Here one workgroup per
// slice reduces the two overlap means (float64 sums), then a streaming
// kernel writes the stitched rows; every output sample is one or two
// multiply-adds of input samples, so the pass is HBM-bound.
#include "ath_internal.h"
#include "log32.h"
#include "spline.h"

namespace ath {
namespace {

constexpr int kThreads = 256;

// numpy.linspace(1.0, 0.0, ov)[t]: t * step + start, the last value exactly
// the stop (make_weight_matrix's list_down)
__device__ __forceinline__ double ramp_down(int t, int ov) {
    if (ov == 1) return 1.0;
    if (t == ov - 1) return 0.0;
    const double step = -1.0 / (double)(ov - 1);
    return (double)t * step + 1.0;
}

// The two images of a stitch, slice j row r: image 1 at
// img1 + (r * nslice + j) * ncol; image 2 at img2 + (r * nslice + j) * ncol,
// read left-right flipped when flip2 (the raw bottom rows of a 360-degree
// sinogram) or as stored (rows already flipped and shifted).
struct StitchImages {
    const float* img1;
    const float* img2;
    int flip2;
    __device__ __forceinline__ float at1(size_t row, int col, int ncol) const {
        return img1[row * ncol + col];
    }
    __device__ __forceinline__ float at2(size_t row, int col, int ncol) const {
        return img2[row * ncol + (flip2 ? ncol - 1 - col : col)];
    }
};

// factors[2*j] = mean of the first image's overlap, factors[2*j+1] = mean of
// the second image's overlap (slice j), float64 sums rounded to float32.
__global__ __launch_bounds__(kThreads) void stitch_means_kernel(
    StitchImages im, int nslice, int ncol, int nrow_180, int overlap,
    int side, float* __restrict__ factors) {
    __shared__ double red[2][kThreads];
    const int j = blockIdx.x;
    const int tid = threadIdx.x;
    // side 1: mean(img1[:, -ov:]), mean(img2[:, :ov]); side 0:
    // mean(img1[:, :ov]), mean(img2[:, -ov:]) (image coordinates)
    const int c1 = side == 1 ? ncol - overlap : 0;
    const int c2 = side == 1 ? 0 : ncol - overlap;
    const long long n = (long long)nrow_180 * overlap;
    double s1 = 0.0, s2 = 0.0;
    for (long long e = tid; e < n; e += kThreads) {
        const int r = (int)(e / overlap), k = (int)(e % overlap);
        const size_t row = (size_t)r * nslice + j;
        s1 += (double)im.at1(row, c1 + k, ncol);
        s2 += (double)im.at2(row, c2 + k, ncol);
    }
    red[0][tid] = s1;
    red[1][tid] = s2;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (tid < w) {
            red[0][tid] += red[0][tid + w];
            red[1][tid] += red[1][tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        factors[2 * j] = (float)(red[0][0] / (double)n);
        factors[2 * j + 1] = (float)(red[1][0] / (double)n);
    }
}

// out: (nrow_180, nslice, width), width >= 2*ncol - overlap (edge-extended).
// w1 / w2: the caller's weight matrices (nrow_180, ncol) of type TW shared
// by every slice, or NULL for make_weight_matrix's linear ramp.  The
// product of a sample with its weight is formed as numpy forms mat *
// wei_mat: in float32 for float32 weights, in float64 for float64 ones.
// mat_comb[:, first] = img_first * w_first (rounded to float32), then
// mat_comb[:, second] += img_second * w_second in the product's type.
template <typename TW>
__global__ __launch_bounds__(kThreads) void stitch_kernel(
    StitchImages im, int nslice, int ncol, int nrow_180, int overlap,
    int side, int norm, const float* __restrict__ factors,
    const TW* __restrict__ w1, const TW* __restrict__ w2, int width,
    float* __restrict__ out) {
    using V = TW;
    const int rows = nrow_180 * nslice;
    const int w0 = 2 * ncol - overlap;
    for (int rj = blockIdx.y; rj < rows; rj += gridDim.y) {
      const int j = rj % nslice, r = rj / nslice;  // block-uniform
      const size_t row = (size_t)r * nslice + j;
      // image 2 is scaled by factor1 / factor2 ((mat2 * f1) / f2)
      const float f1 = norm ? factors[2 * j] : 1.0f;
      const float f2 = norm ? factors[2 * j + 1] : 1.0f;
      float* dst = out + (size_t)rj * width;
      for (int kk = blockIdx.x * kThreads + threadIdx.x; kk < width;
           kk += gridDim.x * kThreads) {
        const int k = kk < w0 ? kk : w0 - 1;  // edge extension to width
        // first image (placed at 0) and second image (placed at ncol - ov)
        // side 1: first = img1 (top), second = img2; side 0: swapped
        float v = 0.0f;
        for (int part = 0; part < 2; ++part) {
            const int off = part == 0 ? 0 : ncol - overlap;
            const int col = k - off;
            if (col < 0 || col >= ncol) continue;
            const bool is_img1 = (side == 1) == (part == 0);
            float x;
            if (is_img1) {
                x = im.at1(row, col, ncol);
            } else {
                x = im.at2(row, col, ncol);
                if (norm) x = (x * f1) / f2;
            }
            V prod;
            if (w1 != nullptr) {
                const TW* w = is_img1 ? w1 : w2;
                prod = (V)x * w[(size_t)r * ncol + col];
            } else {
                // linear ramp over the overlap (make_weight_matrix): the
                // image on the left fades out over its last ov columns, the
                // one on the right fades in over its first ov columns
                float wgt = 1.0f;
                if (part == 0 && col >= ncol - overlap)
                    wgt = (float)ramp_down(col - (ncol - overlap), overlap);
                else if (part == 1 && col < overlap)
                    wgt = (float)(1.0 - ramp_down(col, overlap));  // list_up
                prod = (V)(x * wgt);
            }
            // the first image is stored (float32), the second added to it
            v = part == 0 ? (float)prod : (float)((V)v + prod);
        }
        dst[kk] = v;
      }
    }
}

// scipy.ndimage.shift(row, (0, sh), mode="nearest") (order 3) of nrows
// float32 rows of ncol samples (row q at src + q * ncol, read left-right
// flipped when flip, -log applied first when apply_log), float32 out at
// dst + q * ncol: the reference's sub-pixel shift of the two images before
// stitching (conversion.py:109-117) and of the logged sinogram in
// extend_sinogram (:435-453).  One workgroup per row, grid-strided.
__global__ __launch_bounds__(kThreads) void shift_rows_kernel(
    const float* __restrict__ src, long long nrows, int ncol, int flip,
    int apply_log, double sh, float* __restrict__ dst) {
    extern __shared__ double sh_line[];
    __shared__ double hs[kK3 + 1];
    __shared__ Log32Table s_tab;
    const int L1 = ncol + 2 * kNearestPad;
    double* line = sh_line;
    double* coef = sh_line + L1;
    const int tid = threadIdx.x;
    if (tid <= kK3) hs[tid] = cubic_tap(tid);
    if (apply_log) stage_log32_table<kThreads>(&s_tab);
    __syncthreads();
    for (long long q = blockIdx.x; q < nrows; q += gridDim.x) {
        const float* in = src + q * ncol;
        for (int p = tid; p < L1; p += kThreads) {
            const int c = min(max(p - kNearestPad, 0), ncol - 1);
            float x = in[flip ? ncol - 1 - c : c];
            if (apply_log) x = neg_log_f32(x, s_tab);
            line[p] = (double)x;
        }
        __syncthreads();
        float* o = dst + q * ncol;
        shift_line_nearest<kThreads>(line, coef, ncol, sh, hs,
                                     [o](int jx, double v) {
                                         o[jx] = (float)v;
                                     });
    }
}

// numpy.linspace(0.0, 1.0, ov)[t]: t * step, the last value exactly 1
__device__ __forceinline__ double ramp_up(int t, int ov) {
    if (ov == 1) return 0.0;
    if (t == ov - 1) return 1.0;
    return (double)t * (1.0 / (double)(ov - 1)) + 0.0;
}

// extend_sinogram (conversion.py:395-454) for an integer overlap: (log),
// the overlap columns weighted by a linspace ramp (float64 product rounded
// to float32), edge padding to 2*ncol - overlap columns, times 2.
// side 1: ramp 1 -> 0 over the last ov columns, padding on the right;
// side 0: ramp 0 -> 1 over the first ov columns, padding on the left.
__global__ __launch_bounds__(kThreads) void extend_kernel(
    const float* __restrict__ sino, int nrow, int nslice, int ncol,
    int overlap, int side, int apply_log, float* __restrict__ out) {
    const int width = 2 * ncol - overlap;
    const int pad = ncol - overlap;
    const long long rows = (long long)nrow * nslice;
    for (long long rj = blockIdx.y; rj < rows; rj += gridDim.y) {
        const float* src = sino + rj * ncol;
        float* dst = out + rj * width;
        for (int k = blockIdx.x * kThreads + threadIdx.x; k < width;
             k += gridDim.x * kThreads) {
            // source column of the padded array (edge padding)
            const int c = side == 1 ? min(k, ncol - 1) : max(k - pad, 0);
            float v = src[c];
            if (apply_log) v = neg_log_f32(v);
            if (side == 1 && c >= ncol - overlap)
                v = (float)((double)v *
                            ramp_down(c - (ncol - overlap), overlap));
            else if (side == 0 && c < overlap)
                v = (float)((double)v * ramp_up(c, overlap));
            dst[k] = 2.0f * v;
        }
    }
}

// flat_field_correction (correction.py:45-98): out = (float32(proj) - dark)
// / denom, or float32(proj) / denom without dark, in the fields' type T
// (float32 fields: float32 arithmetic; float64 fields: float64, as numpy
// promotes).  dark / denom: (height, width), broadcast over nouter.
template <typename Tin, typename T>
__global__ __launch_bounds__(kThreads) void flat_field_kernel(
    const Tin* __restrict__ proj, long long nouter, int height, int width,
    long long outer_stride, long long row_stride, const T* __restrict__ dark,
    const T* __restrict__ denom, T* __restrict__ out) {
    const long long rows = nouter * height;
    for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
        const long long n = r / height;  // block-uniform
        const int iy = (int)(r - n * height);
        const Tin* src = proj + n * outer_stride + iy * row_stride;
        const T* dk = dark ? dark + (size_t)iy * width : nullptr;
        const T* dn = denom + (size_t)iy * width;
        T* dst = out + r * width;
        for (int jx = blockIdx.x * kThreads + threadIdx.x; jx < width;
             jx += gridDim.x * kThreads) {
            const T v = (T)(float)src[jx];
            dst[jx] = dk ? (v - dk[jx]) / dn[jx] : v / dn[jx];
        }
    }
}

// out = -log(in) elementwise, float32 <- float64 log rounded (the filter
// kernel's convention for the same step).  VEC = 4: float4 loads and stores
// over the first n - n % 4 elements (16-byte aligned pointers), the rest one
// by one.
template <int VEC>
__global__ __launch_bounds__(kThreads) void minus_log_kernel(
    const float* __restrict__ in, float* __restrict__ out, long long n) {
    __shared__ Log32Table s_tab;
    stage_log32_table<kThreads>(&s_tab);
    __syncthreads();
    const long long nv = n / VEC;
    const long long stride = (long long)gridDim.x * kThreads;
    const long long e0 = (long long)blockIdx.x * kThreads + threadIdx.x;
    for (long long e = e0; e < nv; e += stride) {
        if constexpr (VEC == 4) {
            float4 x = reinterpret_cast<const float4*>(in)[e];
            x.x = neg_log_f32(x.x, s_tab);
            x.y = neg_log_f32(x.y, s_tab);
            x.z = neg_log_f32(x.z, s_tab);
            x.w = neg_log_f32(x.w, s_tab);
            reinterpret_cast<float4*>(out)[e] = x;
        } else {
            out[e] = neg_log_f32(in[e], s_tab);
        }
    }
    const long long t = nv * VEC + e0;  // tail (VEC = 4 only)
    if (t < n) out[t] = neg_log_f32(in[t], s_tab);
}

}  // namespace

int minus_log(const float* in, float* out, long long n, hipStream_t s) {
    if (n <= 0) return ATH_OK;
    const bool vec = reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(out) % 16 == 0;
    const long long items = vec ? n / 4 : n;
    // a few workgroups per CU, each striding over many elements (every
    // workgroup stages the 2 KB log table once)
    const long long blocks = std::max<long long>(
        1, std::min<long long>((items + kThreads - 1) / kThreads, 8192));
    if (vec)
        hipLaunchKernelGGL(minus_log_kernel<4>, dim3((unsigned)blocks),
                           dim3(kThreads), 0, s, in, out, n);
    else
        hipLaunchKernelGGL(minus_log_kernel<1>, dim3((unsigned)blocks),
                           dim3(kThreads), 0, s, in, out, n);
    return check_hip(hipGetLastError(), "minus_log_kernel launch");
}

int extend_sinogram(const float* sino, int nrow, int nslice, int ncol,
                    int overlap, int side, int apply_log, float* out,
                    hipStream_t s) {
    const dim3 grid =
        row_grid((long long)nrow * nslice, 2 * ncol - overlap, kThreads);
    hipLaunchKernelGGL(extend_kernel, grid, dim3(kThreads),
                       0, s, sino, nrow, nslice, ncol, overlap, side,
                       apply_log, out);
    return check_hip(hipGetLastError(), "extend_kernel launch");
}

template <typename Tin, typename T>
static int launch_flat(const void* proj, long long nouter, int height,
                       int width, long long outer_stride, long long row_stride,
                       const void* dark, const void* denom, void* out,
                       hipStream_t s) {
    const dim3 grid = row_grid(nouter * height, width, kThreads);
    hipLaunchKernelGGL((flat_field_kernel<Tin, T>), grid,
                       dim3(kThreads), 0, s, static_cast<const Tin*>(proj),
                       nouter, height, width, outer_stride, row_stride,
                       static_cast<const T*>(dark),
                       static_cast<const T*>(denom), static_cast<T*>(out));
    return check_hip(hipGetLastError(), "flat_field_kernel launch");
}

int flat_field_correction(const void* proj, int in_dtype, long long nouter,
                          int height, int width, long long outer_stride,
                          long long row_stride, const void* dark,
                          const void* denom, int field_dtype, void* out,
                          hipStream_t s) {
    const bool f64 = field_dtype == ATH_F64;
#define ATH_FLAT(TIN)                                                         \
    return f64 ? launch_flat<TIN, double>(proj, nouter, height, width,        \
                                          outer_stride, row_stride, dark,     \
                                          denom, out, s)                      \
               : launch_flat<TIN, float>(proj, nouter, height, width,         \
                                         outer_stride, row_stride, dark,      \
                                         denom, out, s);
    if (in_dtype == ATH_U16) { ATH_FLAT(unsigned short) }
    if (in_dtype == ATH_F64) { ATH_FLAT(double) }
    ATH_FLAT(float)
#undef ATH_FLAT
}

int shift_rows(const float* src, long long nrows, int ncol, int flip,
               int apply_log, double sh, float* dst, hipStream_t s) {
    const size_t lds = 2 * (size_t)(ncol + 2 * kNearestPad) * sizeof(double);
    if (lds > 152 * 1024)
        return set_error(ATH_ERR_UNSUPPORTED,
                         "sub-pixel shift: rows above 9700 columns");
    if (lds > 64 * 1024) {
        const int st = allow_dynamic_lds(
            reinterpret_cast<const void*>(shift_rows_kernel), lds);
        if (st != ATH_OK) return st;
    }
    int grid = 0;
    int st = resident_grid(reinterpret_cast<const void*>(shift_rows_kernel),
                           kThreads, lds, &grid);
    if (st != ATH_OK) return st;
    const long long blocks = std::max<long long>(
        1, std::min<long long>(nrows, (long long)grid));
    hipLaunchKernelGGL(shift_rows_kernel, dim3((unsigned)blocks),
                       dim3(kThreads), lds, s, src, nrows, ncol, flip,
                       apply_log, sh, dst);
    return check_hip(hipGetLastError(), "shift_rows_kernel launch");
}

int stitch_images(const float* img1, const float* img2, int flip2,
                  int nrow_180, int nslice, int ncol, int overlap, int side,
                  int norm, int width, const void* w1, const void* w2,
                  int wdtype, float* factors, float* out, hipStream_t s) {
    const StitchImages im{img1, img2, flip2};
    if (norm) {
        hipLaunchKernelGGL(stitch_means_kernel, dim3(nslice), dim3(kThreads),
                           0, s, im, nslice, ncol, nrow_180, overlap, side,
                           factors);
        const int st = check_hip(hipGetLastError(), "stitch_means_kernel");
        if (st != ATH_OK) return st;
    }
    const dim3 grid = row_grid((long long)nrow_180 * nslice, width, kThreads);
    if (wdtype == ATH_F64)
        hipLaunchKernelGGL(stitch_kernel<double>, grid, dim3(kThreads), 0, s,
                           im, nslice, ncol, nrow_180, overlap, side, norm,
                           factors, static_cast<const double*>(w1),
                           static_cast<const double*>(w2), width, out);
    else
        hipLaunchKernelGGL(stitch_kernel<float>, grid, dim3(kThreads), 0, s,
                           im, nslice, ncol, nrow_180, overlap, side, norm,
                           factors, static_cast<const float*>(w1),
                           static_cast<const float*>(w2), width, out);
    return check_hip(hipGetLastError(), "stitch_kernel launch");
}

}  // namespace ath
