// Kernels of the centre-of-rotation search (algotom/rec/reconstruction.py
// :1211-1458): Gaussian smoothing of the sinogram rows, the order-1 zoom of
// the coarse search and the per-slice metrics scored for every candidate
// centre.  Every sum follows the order numpy / scipy use, so the metrics
// equal the reference's (bit-for-bit for the sharpness; the entropy up to
// the last-bit differences of the two libraries' log2).
//
// gaussian_rows_kernel   scipy.ndimage.gaussian_filter1d(x, sigma, axis=-1)
//                        (mode "reflect" = numpy's "symmetric", truncate 4):
//                        float64 weights exp(-x^2 / (2 sigma^2)) / sum and
//                        float64 sums, float32 output.
// zoom_linear_kernel     scipy.ndimage.zoom(x, z, order=1, mode="nearest")
//                        (:1434), bilinear in float64.
// level_histogram_kernel + entropy_finish_kernel
//                        __calculate_histogram_entropy (:1211-1220) on
//                        clip((rec - nmin) / (nmax - nmin), 0, 1): uint8
//                        levels v = trunc(255 r) (float32), 256-bin counts
//                        (ndi.histogram(., 0, 255, 256) puts level v in bin
//                        v), h = 1 + counts, 7-tap boxcar "valid",
//                        metric = -sum h log2 h (numpy's pairwise sum).
//                        Integer LDS counts: exact.
// sharpness_chunk_kernel + sharpness_finish_kernel
//                        __calculate_edge_sharpness (:1223-1228): np.mean of
//                        |np.gradient| (central differences inside, one-sided
//                        at the borders), float32 sums in numpy's order.
#include "ath_internal.h"

namespace ath {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxRadius = 64;
constexpr int kChunk = 8192;  // numpy's default ufunc buffer size

__global__ __launch_bounds__(kThreads) void gaussian_rows_kernel(
    const float* __restrict__ in, float* __restrict__ out, int ncol,
    const double* __restrict__ weights, int radius) {
    __shared__ double w[2 * kMaxRadius + 1];
    const size_t row = blockIdx.x;
    for (int k = threadIdx.x; k <= 2 * radius; k += kThreads)
        w[k] = weights[k];
    __syncthreads();
    const float* src = in + row * (size_t)ncol;
    for (int c = threadIdx.x; c < ncol; c += kThreads) {
        // scipy's symmetric-kernel order: x w0 + sum_j (x[+j] + x[-j]) w_j
        double acc = (double)src[c] * w[radius];
        for (int k = 1; k <= radius; ++k) {
            const int cp = pad_source(c + k, ncol, ATH_PAD_SYMMETRIC);
            const int cm = pad_source(c - k, ncol, ATH_PAD_SYMMETRIC);
            acc += ((double)src[cp] + (double)src[cm]) * w[radius + k];
        }
        out[row * (size_t)ncol + c] = (float)acc;
    }
}

// numpy's pairwise_sum (umath loops_utils) for one block of n <= 128
// values: 8 running partial sums, combined as a tree, then the remainder.
template <typename T>
__device__ T np_pairwise_leaf(const T* a, int n) {
    if (n < 8) {
        T r = T(0);
        for (int i = 0; i < n; ++i) r += a[i];
        return r;
    }
    T r0 = a[0], r1 = a[1], r2 = a[2], r3 = a[3];
    T r4 = a[4], r5 = a[5], r6 = a[6], r7 = a[7];
    int i = 8;
    for (; i < n - (n % 8); i += 8) {
        r0 += a[i];     r1 += a[i + 1]; r2 += a[i + 2]; r3 += a[i + 3];
        r4 += a[i + 4]; r5 += a[i + 5]; r6 += a[i + 6]; r7 += a[i + 7];
    }
    T res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
    for (; i < n; ++i) res += a[i];
    return res;
}

// numpy's pairwise_sum for n <= kChunk values on one thread: blocks of at
// most 128 values, larger ranges split at n/2 rounded down to a multiple of
// 8 and summed left + right.  The recursion runs on an explicit stack (depth
// <= 9 for n <= 8192).
template <typename T>
__device__ T np_pairwise_serial(const T* a, int n) {
    int off[16], len[16], stage[16];
    T val[16];
    int sp = 1, vp = 0;
    off[0] = 0;
    len[0] = n;
    stage[0] = 0;
    while (sp > 0) {
        const int t = sp - 1;
        const int m = len[t];
        if (m <= 128) {
            val[vp++] = np_pairwise_leaf(a + off[t], m);
            --sp;
            continue;
        }
        int m2 = m / 2;
        m2 -= m2 % 8;
        if (stage[t] == 0) {
            stage[t] = 1;
            off[sp] = off[t];
            len[sp] = m2;
            stage[sp++] = 0;
        } else if (stage[t] == 1) {
            stage[t] = 2;
            off[sp] = off[t] + m2;
            len[sp] = m - m2;
            stage[sp++] = 0;
        } else {
            const T r = val[--vp];
            const T l = val[--vp];
            val[vp++] = l + r;
            --sp;
        }
    }
    return val[0];
}

// Sample of an image as the reference's float32 arithmetic sees it
// (np.float32(image) for float64 slices).
template <typename T>
__device__ __forceinline__ float as_f32(T v) { return (float)v; }

// uint8 level of one sample: clip((v - nmin) / (nmax - nmin), 0, 1) * 255,
// truncated, in the image's precision (float32 slices: float32 scalars;
// float64 slices: float64, as numpy evaluates it).
__device__ __forceinline__ int level_of(float v, double nmin, double nmax) {
    const float lo = (float)nmin;
    float r = (v - lo) / ((float)nmax - lo);
    r = fminf(fmaxf(r, 0.0f), 1.0f);
    return (int)(r * 255.0f);
}
__device__ __forceinline__ int level_of(double v, double nmin, double nmax) {
    double r = (v - nmin) / (nmax - nmin);
    r = fmin(fmax(r, 0.0), 1.0);
    return (int)(r * 255.0);
}

// Histogram of the uint8 levels: block (b, k) counts its contiguous share
// of image k into LDS and writes the 256 counts to
// partial[(k * nblk + b) * 256 + bin] (integers: the split changes nothing).
template <typename T>
__global__ __launch_bounds__(kThreads) void level_histogram_kernel(
    const T* __restrict__ images, long long stride, long long n, double nmin,
    double nmax, unsigned int* __restrict__ partial) {
    __shared__ unsigned int hist[256];
    const int tid = threadIdx.x;
    const int nblk = gridDim.x;
    const T* img = images + (size_t)blockIdx.y * (size_t)stride;
    hist[tid] = 0u;
    __syncthreads();
    const long long per = (n + nblk - 1) / nblk;
    const long long e0 = (long long)blockIdx.x * per;
    const long long e1 = min(n, e0 + per);
    for (long long e = e0 + tid; e < e1; e += kThreads)
        atomicAdd(&hist[level_of(img[e], nmin, nmax)], 1u);
    __syncthreads();
    partial[((size_t)blockIdx.y * nblk + blockIdx.x) * 256 + tid] = hist[tid];
}

// One workgroup per image: add the block histograms, h = 1 + counts,
// boxcar(7) "valid" -> 250 sliding sums of integers (exact in float64),
// metric = -sum h log2 h with numpy's pairwise order.
__global__ __launch_bounds__(kThreads) void entropy_finish_kernel(
    const unsigned int* __restrict__ partial, int nblk,
    double* __restrict__ out) {
    __shared__ unsigned int hist[256];
    __shared__ double h7[256];
    const int tid = threadIdx.x;
    const unsigned int* p = partial + (size_t)blockIdx.x * nblk * 256;
    unsigned int c = 0u;
    for (int b = 0; b < nblk; ++b) c += p[(size_t)b * 256 + tid];
    hist[tid] = c;
    __syncthreads();
    if (tid < 250) {
        double s = 0.0;
        for (int k = 0; k < 7; ++k) s += 1.0 + (double)hist[tid + k];
        h7[tid] = s * log2(s);
    }
    __syncthreads();
    if (tid == 0) out[blockIdx.x] = -np_pairwise_serial(h7, 250);
}

// |np.gradient| of pixel e of a (nrows, ncols) image converted to float32:
// axis 0 then axis 1, central differences halved inside, one-sided at the
// borders, float32.
template <typename T>
__device__ __forceinline__ float grad_magnitude(const T* img, int nrows,
                                                int ncols, int e) {
    const int y = e / ncols, x = e - y * ncols;
    float gy, gx;
    const float v = as_f32(img[e]);
    if (y == 0) gy = as_f32(img[e + ncols]) - v;
    else if (y == nrows - 1) gy = v - as_f32(img[e - ncols]);
    else gy = (as_f32(img[e + ncols]) - as_f32(img[e - ncols])) / 2.0f;
    if (x == 0) gx = as_f32(img[e + 1]) - v;
    else if (x == ncols - 1) gx = v - as_f32(img[e - 1]);
    else gx = (as_f32(img[e + 1]) - as_f32(img[e - 1])) / 2.0f;
    return sqrtf(gy * gy + gx * gx);
}

// np.mean of a float32 image: numpy reduces through its 8192-element
// buffer, so the float32 sum is a sequential float32 accumulation (from 0)
// of pairwise sums of consecutive 8192-value chunks of the flattened image;
// mean = float32(float64(sum) / n).
// Block (c, k) forms chunk c of image k: magnitudes to LDS; a full chunk is
// 64 blocks of 128 (64 threads) combined as the perfect binary tree the
// recursion forms, a shorter tail chunk is summed by one thread.  The chunk
// sum goes to partial[k * nchunk + c].
template <typename T>
__global__ __launch_bounds__(kThreads) void sharpness_chunk_kernel(
    const T* __restrict__ images, long long stride, int nrows, int ncols,
    float* __restrict__ partial) {
    __shared__ float mag[kChunk];
    __shared__ float part[64];
    const T* img = images + (size_t)blockIdx.y * (size_t)stride;
    const int tid = threadIdx.x;
    const int n = nrows * ncols;
    const int c0 = blockIdx.x * kChunk;
    const int m = min(kChunk, n - c0);
    for (int k = tid; k < m; k += kThreads)
        mag[k] = grad_magnitude(img, nrows, ncols, c0 + k);
    __syncthreads();
    float chunk;
    if (m == kChunk) {
        if (tid < 64) part[tid] = np_pairwise_leaf(mag + 128 * tid, 128);
        __syncthreads();
        for (int w = 1; w < 64; w <<= 1) {
            if (tid < 64 && (tid % (2 * w)) == 0)
                part[tid] = part[tid] + part[tid + w];
            __syncthreads();
        }
        chunk = part[0];
    } else {
        chunk = tid == 0 ? np_pairwise_serial(mag, m) : 0.0f;
    }
    if (tid == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = chunk;
}

// One workgroup per image: the chunk sums added in order, staged in LDS.
__global__ __launch_bounds__(kThreads) void sharpness_finish_kernel(
    const float* __restrict__ partial, int nchunk, long long n,
    double* __restrict__ out) {
    __shared__ float buf[4096];
    const int tid = threadIdx.x;
    const float* p = partial + (size_t)blockIdx.x * nchunk;
    float total = 0.0f;  // meaningful on thread 0 only
    for (int b0 = 0; b0 < nchunk; b0 += 4096) {
        const int m = min(4096, nchunk - b0);
        for (int k = tid; k < m; k += kThreads) buf[k] = p[b0 + k];
        __syncthreads();
        if (tid == 0)
            for (int k = 0; k < m; ++k) total = total + buf[k];
        __syncthreads();
    }
    if (tid == 0) out[blockIdx.x] = (double)(float)((double)total / (double)n);
}

// vertrec's autocorrelation metric (vertrec.py:1255-1261): the mean over
// rows of max |ifft(|fft(row)|^2)|.  The circular autocorrelation of a real
// row peaks at lag 0, where it equals sum(row^2) (Parseval), so each row
// contributes its float64 sum of squares of the float32 samples.
template <typename T>
__global__ __launch_bounds__(kThreads) void autocorr_kernel(
    const T* __restrict__ images, long long stride, int nrows, int ncols,
    double* __restrict__ out) {
    __shared__ double red[kThreads];
    const T* img = images + (size_t)blockIdx.x * (size_t)stride;
    const int tid = threadIdx.x;
    double acc = 0.0;
    for (int r = tid; r < nrows; r += kThreads) {
        const T* row = img + (size_t)r * ncols;
        double s = 0.0;
        for (int c = 0; c < ncols; ++c) {
            const double v = (double)as_f32(row[c]);
            s += v * v;
        }
        acc += s;
    }
    red[tid] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) out[blockIdx.x] = red[0] / (double)nrows;
}

// scipy.ndimage.gaussian_filter1d along the middle axis of a contiguous
// (nouter, length, ninner) float32 array, ninner > 1: one thread per inner
// index (coalesced across lanes), looping along the axis; same sums and
// boundary as gaussian_rows_kernel.
__global__ __launch_bounds__(kThreads) void gaussian_inner_kernel(
    const float* __restrict__ in, float* __restrict__ out, int length,
    int ninner, const double* __restrict__ weights, int radius) {
    __shared__ double w[2 * kMaxRadius + 1];
    for (int k = threadIdx.x; k <= 2 * radius; k += kThreads)
        w[k] = weights[k];
    __syncthreads();
    const int b = blockIdx.x * kThreads + threadIdx.x;
    if (b >= ninner) return;
    const size_t base = (size_t)blockIdx.y * length * ninner + b;
    const float* src = in + base;
    float* dst = out + base;
    for (int k = 0; k < length; ++k) {
        double acc = (double)src[(size_t)k * ninner] * w[radius];
        for (int j = 1; j <= radius; ++j) {
            const int kp = pad_source(k + j, length, ATH_PAD_SYMMETRIC);
            const int km = pad_source(k - j, length, ATH_PAD_SYMMETRIC);
            acc += ((double)src[(size_t)kp * ninner] +
                    (double)src[(size_t)km * ninner]) * w[radius + j];
        }
        dst[(size_t)k * ninner] = (float)acc;
    }
}

// scipy.ndimage.zoom(x, zoom, order=1, mode="nearest") of a 2D float32
// image (grid_mode=False): output pixel (i, j) samples the input at
// (i * (h-1)/(oh-1), j * (w-1)/(ow-1)) by bilinear interpolation in float64,
// neighbours past the last row/column clamped to it.
__global__ __launch_bounds__(kThreads) void zoom_linear_kernel(
    const float* __restrict__ in, int h, int w, float* __restrict__ out,
    int oh, int ow, double zy, double zx) {
    for (int i = blockIdx.y; i < oh; i += gridDim.y)
      for (int j = blockIdx.x * kThreads + threadIdx.x; j < ow;
           j += gridDim.x * kThreads) {
        const double yy = (double)i * zy, xx = (double)j * zx;
        int y0 = (int)floor(yy), x0 = (int)floor(xx);
        y0 = min(max(y0, 0), h - 1);
        x0 = min(max(x0, 0), w - 1);
        const double ty = yy - (double)y0, tx = xx - (double)x0;
        const int y1 = min(y0 + 1, h - 1), x1 = min(x0 + 1, w - 1);
        const double f00 = in[(size_t)y0 * w + x0], f01 = in[(size_t)y0 * w + x1];
        const double f10 = in[(size_t)y1 * w + x0], f11 = in[(size_t)y1 * w + x1];
        const double top = (1.0 - tx) * f00 + tx * f01;
        const double bot = (1.0 - tx) * f10 + tx * f11;
        out[(size_t)i * ow + j] = (float)((1.0 - ty) * top + ty * bot);
      }
}

}  // namespace

int gaussian_rows(const float* in, float* out, int nrows, int ncol,
                  const double* weights, int radius, hipStream_t s) {
    if (radius > kMaxRadius)
        return set_error(ATH_ERR_UNSUPPORTED, "gaussian radius above 64");
    hipLaunchKernelGGL(gaussian_rows_kernel, dim3(nrows), dim3(kThreads), 0,
                       s, in, out, ncol, weights, radius);
    return check_hip(hipGetLastError(), "gaussian_rows_kernel launch");
}

int zoom_linear(const float* in, int h, int w, float* out, int oh, int ow,
                hipStream_t s) {
    const double zy = oh > 1 ? (double)(h - 1) / (double)(oh - 1) : 1.0;
    const double zx = ow > 1 ? (double)(w - 1) / (double)(ow - 1) : 1.0;
    hipLaunchKernelGGL(zoom_linear_kernel, row_grid(oh, ow, kThreads),
                       dim3(kThreads), 0, s, in, h, w, out, oh, ow, zy, zx);
    return check_hip(hipGetLastError(), "zoom_linear_kernel launch");
}

static int entropy_blocks(long long n) {
    return (int)std::min<long long>(
        64, std::max<long long>(1, (n + 65535) / 65536));
}

size_t slice_metrics_workspace(int nimg, int nrows, int ncols, int metric) {
    const long long n = (long long)nrows * ncols;
    if (metric == 0)
        return (size_t)nimg * entropy_blocks(n) * 256 * sizeof(unsigned);
    if (metric == 1)
        return (size_t)nimg * (size_t)((n + kChunk - 1) / kChunk) *
               sizeof(float);
    return 0;
}

template <typename T>
static int metrics_typed(const T* images, int nimg, int nrows, int ncols,
                         long long stride, double nmin, double nmax,
                         int metric, double* out, void* workspace,
                         hipStream_t s) {
    const long long n = (long long)nrows * ncols;
    if (metric == 0) {
        const int nblk = entropy_blocks(n);
        auto* partial = static_cast<unsigned int*>(workspace);
        hipLaunchKernelGGL(level_histogram_kernel<T>, dim3(nblk, nimg),
                           dim3(kThreads), 0, s, images, stride, n, nmin,
                           nmax, partial);
        const int st = check_hip(hipGetLastError(), "level_histogram_kernel");
        if (st != ATH_OK) return st;
        hipLaunchKernelGGL(entropy_finish_kernel, dim3(nimg), dim3(kThreads),
                           0, s, partial, nblk, out);
        return check_hip(hipGetLastError(), "entropy_finish_kernel");
    }
    if (metric == 2) {
        hipLaunchKernelGGL(autocorr_kernel<T>, dim3(nimg), dim3(kThreads), 0,
                           s, images, stride, nrows, ncols, out);
        return check_hip(hipGetLastError(), "autocorr_kernel");
    }
    const int nchunk = (int)((n + kChunk - 1) / kChunk);
    auto* partial = static_cast<float*>(workspace);
    hipLaunchKernelGGL(sharpness_chunk_kernel<T>, dim3(nchunk, nimg),
                       dim3(kThreads), 0, s, images, stride, nrows, ncols,
                       partial);
    const int st = check_hip(hipGetLastError(), "sharpness_chunk_kernel");
    if (st != ATH_OK) return st;
    hipLaunchKernelGGL(sharpness_finish_kernel, dim3(nimg), dim3(kThreads), 0,
                       s, partial, nchunk, n, out);
    return check_hip(hipGetLastError(), "sharpness_finish_kernel");
}

int slice_metrics(const void* images, int dtype, int nimg, int nrows,
                  int ncols, long long stride, double nmin, double nmax,
                  int metric, double* out, void* workspace, hipStream_t s) {
    if (dtype == ATH_F64)
        return metrics_typed(static_cast<const double*>(images), nimg, nrows,
                             ncols, stride, nmin, nmax, metric, out,
                             workspace, s);
    return metrics_typed(static_cast<const float*>(images), nimg, nrows, ncols,
                         stride, nmin, nmax, metric, out, workspace, s);
}

int gaussian_axis(const float* in, float* out, int nouter, int length,
                  int ninner, const double* weights, int radius,
                  hipStream_t s) {
    if (radius > kMaxRadius)
        return set_error(ATH_ERR_UNSUPPORTED, "gaussian radius above 64");
    if (ninner == 1) return gaussian_rows(in, out, nouter, length, weights,
                                          radius, s);
    hipLaunchKernelGGL(gaussian_inner_kernel,
                       dim3((ninner + kThreads - 1) / kThreads, nouter),
                       dim3(kThreads), 0, s, in, out, length, ninner, weights,
                       radius);
    return check_hip(hipGetLastError(), "gaussian_inner_kernel launch");
}

}  // namespace ath
