// Direct Fourier inversion (DFI) on the GPU: algotom's dfi_reconstruction
// (algotom/rec/reconstruction.py:687-907) for one slice at a time.
//
// Per slice, in reference order:
//   1. dfi_shift_kernel     the odd-width, centre-shifted sinogram
//                           (scipy.ndimage.shift order 3, mode "nearest",
//                           :749-781): one workgroup per row, the 12-sample
//                           edge-padded line and its cubic B-spline
//                           coefficients in LDS, values rounded to the
//                           input's float type;
//   2. dfi_angles_kernel    (only with angles, :731-746) 360-degree fold,
//                           the stacked [rows; flipped] sinogram shifted
//                           along the angle axis (order 3, mode "wrap");
//   3. dfi_rows_fwd_kernel  pad (numpy.pad modes), -log, ifftshift, the
//                           length-N DFT by Bluestein's chirp-z method over
//                           a power-of-two LDS FFT (two real rows per
//                           complex transform), fftshift, smoothing window,
//                           then the quintic B-spline prefilter along the
//                           row (:787-790 and the first axis of
//                           map_coordinates' prefilter);
//   4. dfi_cols_kernel      the quintic prefilter along the angle axis;
//   5. dfi_grid_kernel      generate_mapping_coordinate (:687-728) per
//                           Cartesian frequency, order-5 evaluation on the
//                           (theta, r) grid (mode "reflect"), the circle
//                           mask, written ifftshifted and transposed as the
//                           Hermitian part (the real part of the inverse
//                           depends on nothing else), half of it;
//   6. dfi_cols_inv_kernel  inverse DFT along the first axis (Bluestein) of
//                           that half, keeping the rows the crop keeps;
//      dfi_transpose_kernel;
//      dfi_rows_out_kernel  inverse DFT along the second axis of the
//                           Hermitian rows (two real rows per transform),
//                           fftshift, crop [pad, pad + ncol), ratio mask --
//                           the real part the reference takes (:800-801).
//
// Spline conventions (scipy.ndimage 1.15): prefilter = the exact inverse
// of the sampled B-spline, applied as its impulse response h (closed form
// from the filter's poles; |h[k]| < 1e-18 h[0] past the stored half
// lengths) to the symmetric extension of the line: half-sample symmetric
// for modes "nearest" (after the 12-sample edge padding) and "reflect",
// whole-sample for "wrap".  Evaluation sums the order + 1 nearest knots
// weighted by the centred B-spline; knots outside the line are clamped
// ("nearest"), folded half-sample ("reflect") or whole-sample ("wrap").
// All arithmetic after the shift is float64, as numpy's complex128 FFTs.
#include "ath_internal.h"
#include "fft_lds.h"
#include "log32.h"
#include "spline.h"

namespace ath {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxLine = 8192 + 2 * kNearestPad;  // longest shifted line
constexpr double kPi = 3.141592653589793;         // numpy.pi

// ---- plan layout ---------------------------------------------------------
struct DfiPlan {
    int N, M;
    size_t tw, chirp, bhat, window, h3, h5, total;  // byte offsets
};

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int bluestein_length(int N) {
    int M = 1;
    while (M < 2 * N - 1) M <<= 1;
    return M < 4 ? 4 : M;
}

DfiPlan plan_layout(int N) {
    DfiPlan p{};
    p.N = N;
    p.M = bluestein_length(N);
    size_t o = 0;
    p.tw = o;     o = align256(o + (size_t)2 * p.M * sizeof(double2));
    p.chirp = o;  o = align256(o + (size_t)N * sizeof(double2));
    p.bhat = o;   o = align256(o + (size_t)p.M * sizeof(double2));
    p.window = o; o = align256(o + (size_t)N * sizeof(double));
    p.h3 = o;     o = align256(o + (size_t)(kK3 + 1) * sizeof(double));
    p.h5 = o;     o = align256(o + (size_t)(kK5 + 1) * sizeof(double));
    p.total = o;
    return p;
}

template <typename T>
__host__ __device__ T* at(void* base, size_t off) {
    return reinterpret_cast<T*>(static_cast<char*>(base) + off);
}
template <typename T>
__host__ __device__ const T* at(const void* base, size_t off) {
    return reinterpret_cast<const T*>(static_cast<const char*>(base) + off);
}

// numpy.remainder(a, b) for b > 0
__device__ __forceinline__ double np_mod(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0 && m < 0.0) m += b;
    return m;
}

template <typename T>
__device__ __forceinline__ double round_to(double v) {
    return (double)(T)v;
}

// ---- plan kernels ---------------------------------------------------------
// chirp[n] = exp(-i pi n^2 / N), the exponent reduced mod 2N exactly
__global__ void chirp_kernel(double2* chirp, int N) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) {
        const long long q = ((long long)n * n) % (2LL * N);
        double s, c;
        sincospi((double)q / (double)N, &s, &c);
        chirp[n] = make_double2(c, -s);
    }
}

// bhat = FFT_M(b) / M, b[m] = b[M - m] = exp(+i pi m^2 / N) for 0 <= m < N,
// zero elsewhere (one workgroup)
template <int MAXB>
__global__ __launch_bounds__(kFftThreads) void bhat_kernel(
    const double2* __restrict__ chirp, const double2* __restrict__ tw, int N,
    int M, double2* __restrict__ bhat) {
    extern __shared__ double2 lds[];
    for (int k = threadIdx.x; k < M; k += kFftThreads) {
        double2 v = make_double2(0.0, 0.0);
        if (k < N) v = conjg(chirp[k]);
        else if (k > M - N) v = conjg(chirp[M - k]);
        lds[k] = v;
    }
    __syncthreads();
    lds_fft<-1, MAXB>(lds, M, tw);
    const double inv = 1.0 / (double)M;
    for (int k = threadIdx.x; k < M; k += kFftThreads)
        bhat[k] = cscale(lds[k], inv);
}

// window (ones without one) and the inverse-filter taps.  1 / B(q) over the
// poles z_i of the sampled B-spline (|z_i| < 1, z_i + 1/z_i = a_i):
//   cubic:   h[k] = (1 - z)^2 g_z[k]
//   quintic: h[k] = C (z2 g_z2[k] - z1 g_z1[k]) / (z1 z2 (a1 - a2)),
//            C = (1 - z1)^2 (1 - z2)^2,
// with g_z[k] = z^|k| / (1 - z^2), the response of 1/((1 - zq)(1 - z/q)).
__global__ void dfi_tables_kernel(const double* __restrict__ window, int N,
                                  double* __restrict__ win,
                                  double* __restrict__ h3,
                                  double* __restrict__ h5) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) win[i] = window ? window[i] : 1.0;
    if (i <= kK3) h3[i] = cubic_tap(i);
    if (i <= kK5) {
        const double w1 = -13.0 + sqrt(105.0), w2 = -13.0 - sqrt(105.0);
        const double z1 = 2.0 / (w1 - sqrt(w1 * w1 - 4.0));
        const double z2 = 2.0 / (w2 - sqrt(w2 * w2 - 4.0));
        const double C = (1.0 - z1) * (1.0 - z1) * (1.0 - z2) * (1.0 - z2);
        const double g1 = pow(z1, (double)i) / (1.0 - z1 * z1);
        const double g2 = pow(z2, (double)i) / (1.0 - z2 * z2);
        h5[i] = C * (z2 * g2 - z1 * g1) / (z1 * z2 * (w1 - w2));
    }
}

// ---- 1. centre shift --------------------------------------------------------
// in: (nrow, ncol) rows at row_stride; out S: (nrow, ncol1) float64 holding
// values of the input's float type Tr.  line = the odd-width row edge-padded
// by 12 on both sides (L1 = ncol1 + 24), coefficients by the half-sample
// symmetric prefilter, out[j] = sum_q beta3(pos - k_q) c[clamp(k_q)] at
// pos = (j - xshift) + 12.
template <typename Tin, typename Tr>
__global__ __launch_bounds__(kThreads) void dfi_shift_kernel(
    const Tin* __restrict__ in, long long row_stride, int nrow, int ncol,
    int ncol1, double xshift, const double* __restrict__ h3,
    double* __restrict__ S) {
    extern __shared__ double sh[];
    __shared__ double hs[kK3 + 1];
    const int L1 = ncol1 + 2 * kNearestPad;
    double* line = sh;
    double* coef = sh + L1;
    const int tid = threadIdx.x;
    if (tid <= kK3) hs[tid] = h3[tid];
    for (int row = blockIdx.x; row < nrow; row += gridDim.x) {
        const Tin* src = in + (long long)row * row_stride;
        for (int p = tid; p < L1; p += kThreads) {
            int q = min(max(p - kNearestPad, 0), ncol1 - 1);
            q = min(q, ncol - 1);  // the odd-width edge column
            line[p] = (double)src[q];
        }
        __syncthreads();
        for (int p = tid; p < L1; p += kThreads) {
            double c = hs[0] * line[p];
            for (int t = 1; t <= kK3; ++t)
                c += hs[t] * (line[fold_half(p - t, L1)] +
                              line[fold_half(p + t, L1)]);
            coef[p] = c;
        }
        __syncthreads();
        double* dst = S + (size_t)row * ncol1;
        for (int j = tid; j < ncol1; j += kThreads) {
            const double pos = ((double)j - xshift) + (double)kNearestPad;
            const int k0 = (int)floor(pos) - 1;
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = k0 + q;
                v += beta3(pos - (double)k) * coef[min(max(k, 0), L1 - 1)];
            }
            dst[j] = round_to<Tr>(v);
        }
        __syncthreads();
    }
}

// The same shift for lines longer than kMaxLine (the line and its
// coefficients no longer fit in LDS together).  Outputs go in segments of
// kShiftSeg; a segment needs the coefficients of a contiguous knot range
// [c0, c1] (clamped to the line) and those the line values within kK3 of
// it (a half-sample fold near an end lands inside the staged range, which
// then starts or ends at the line's end).  line(p) is read from global
// memory by the same index rule as the LDS form.
constexpr int kShiftSeg = 4096;

template <typename Tin, typename Tr>
__global__ __launch_bounds__(kThreads) void dfi_shift_seg_kernel(
    const Tin* __restrict__ in, long long row_stride, int nrow, int ncol,
    int ncol1, double xshift, const double* __restrict__ h3,
    double* __restrict__ S) {
    constexpr int kCoef = kShiftSeg + 4;
    __shared__ double line[kCoef + 2 * kK3];
    __shared__ double coef[kCoef];
    __shared__ double hs[kK3 + 1];
    const int L1 = ncol1 + 2 * kNearestPad;
    const int tid = threadIdx.x;
    if (tid <= kK3) hs[tid] = h3[tid];
    auto knot = [&](int j) {  // first knot of output j, clamped
        const double pos = ((double)j - xshift) + (double)kNearestPad;
        return min(max((int)floor(pos) - 1, 0), L1 - 1);
    };
    const int nseg = (ncol1 + kShiftSeg - 1) / kShiftSeg;
    for (int item = blockIdx.x; item < nrow * nseg; item += gridDim.x) {
        const int row = item / nseg, j0 = (item % nseg) * kShiftSeg;
        const int j1 = min(j0 + kShiftSeg, ncol1) - 1;
        const Tin* src = in + (long long)row * row_stride;
        const int c0 = knot(j0), c1 = min(knot(j1) + 3, L1 - 1);
        const int b0 = max(c0 - kK3, 0), b1 = min(c1 + kK3, L1 - 1);
        __syncthreads();  // previous item's LDS consumed (and hs staged)
        for (int p = b0 + tid; p <= b1; p += kThreads) {
            int q = min(max(p - kNearestPad, 0), ncol1 - 1);
            q = min(q, ncol - 1);  // the odd-width edge column
            line[p - b0] = (double)src[q];
        }
        __syncthreads();
        for (int p = c0 + tid; p <= c1; p += kThreads) {
            double c = hs[0] * line[p - b0];
            for (int t = 1; t <= kK3; ++t)
                c += hs[t] * (line[fold_half(p - t, L1) - b0] +
                              line[fold_half(p + t, L1) - b0]);
            coef[p - c0] = c;
        }
        __syncthreads();
        double* dst = S + (size_t)row * ncol1;
        for (int j = j0 + tid; j <= j1; j += kThreads) {
            const double pos = ((double)j - xshift) + (double)kNearestPad;
            const int k0 = (int)floor(pos) - 1;
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = k0 + q;
                v += beta3(pos - (double)k) *
                     coef[min(max(k, 0), L1 - 1) - c0];
            }
            dst[j] = round_to<Tr>(v);
        }
    }
}

// ---- 2. angle handling ----------------------------------------------------
// src(r, j) = the (folded) sinogram row r < rows; v360 = [src[:rows - 1];
// fliplr(src)] (L = 2 rows - 1 rows); out[i] = sum_q beta3(pos - k_q)
// c[fold_whole(k_q)] with pos = i - shift wrapped into [0, L - 1) when
// outside [0, L - 1] and c the whole-sample prefilter of v360 along rows.
template <typename Tr>
__device__ __forceinline__ double angle_src(const double* S, int nrow_in,
                                            int ncol1, int rows, int fold,
                                            int r, int j) {
    const double a = S[(size_t)r * ncol1 + j];
    if (!fold) return a;
    const double b = S[(size_t)(nrow_in - rows + r) * ncol1 + (ncol1 - 1 - j)];
    return round_to<Tr>(round_to<Tr>(a + b) / 2.0);
}

template <typename Tr>
__global__ __launch_bounds__(kThreads) void dfi_angles_kernel(
    const double* __restrict__ S, int nrow_in, int ncol1, int rows, int fold,
    double shift, int flip, const double* __restrict__ h3,
    double* __restrict__ S2) {
    const int L = 2 * rows - 1;
    for (int i = blockIdx.y; i < rows; i += gridDim.y) {
        double pos = (double)i - shift;
        if (L > 1 && (pos < 0.0 || pos > (double)(L - 1)))
            pos = np_mod(pos, (double)(L - 1));
        const int k0 = (int)floor(pos) - 1;
        for (int j = blockIdx.x * kThreads + threadIdx.x; j < ncol1;
             j += gridDim.x * kThreads) {
            double v = 0.0;
            for (int q = 0; q < 4; ++q) {
                const int k = fold_whole(k0 + q, L);
                double c = 0.0;
                for (int t = -kK3; t <= kK3; ++t) {
                    const int r = fold_whole(k + t, L);
                    const double x =
                        r < rows - 1
                            ? angle_src<Tr>(S, nrow_in, ncol1, rows, fold, r, j)
                            : angle_src<Tr>(S, nrow_in, ncol1, rows, fold,
                                            r - (rows - 1), ncol1 - 1 - j);
                    c += h3[t < 0 ? -t : t] * x;
                }
                v += beta3(pos - (double)(k0 + q)) * c;
            }
            v = round_to<Tr>(v);
            if (flip)
                S2[(size_t)(rows - 1 - i) * ncol1 + (ncol1 - 1 - j)] = v;
            else
                S2[(size_t)i * ncol1 + j] = v;
        }
    }
}

// ---- Bluestein DFT over the LDS FFT ----------------------------------------
// nk = (n^2 + k^2 - (k - n)^2) / 2 turns the length-N DFT into a circular
// convolution of length M >= 2N - 1:  DFT(a)[k] = chirp[k] * conv[k],
// conv = IFFT_M(FFT_M(chirp * a) * bhat).  On entry buf[n] = chirp[n] a[n]
// (n < N, zero up to M); on exit buf[k] = conv[k] (bhat carries the 1/M).
template <int MAXB, bool QT, int NT>
__device__ __forceinline__ void bluestein_core(double2* buf, int M,
                                               const double2* __restrict__ tw,
                                               const double2* tq,
                                               const double2* __restrict__ bhat) {
    // the rotated first radix-8 pass (conflict-free LDS writes; without the
    // quarter table only)
    lds_fft<-1, MAXB, QT, NT, 1, true>(buf, M, tw, tq);
    for (int k = threadIdx.x; k < M; k += NT)
        buf[k] = cmul(buf[k], bhat[k]);
    __syncthreads();
    lds_fft<+1, MAXB, QT, NT, 1, true>(buf, M, tw, tq);
}

// The quarter twiddle table sits in LDS beside the buffer for M <= 4096.
template <int MAXB, int NT>
constexpr bool quarter_table() { return MAXB * NT <= 1024; }

// ---- 3. row spectra + row prefilter ----------------------------------------
// -log of one padded sample in the input's precision; values <= 0 become 1
// and raise the flag (reconstruction.py:862-867)
template <typename Tr>
__device__ __forceinline__ double minus_log(double v, int* flag) {
    if (v <= 0.0) {
        v = 1.0;
        *flag = 1;
    }
    if constexpr (sizeof(Tr) == 4)
        return (double)neg_log_f32((float)v);  // v is a float32 value
    return round_to<Tr>(-log(v));
}

// Consecutive row-prefilter outputs per thread.  Odd, so that the lanes'
// 16-byte LDS reads, kFirRun * 16 bytes apart, fall on distinct banks.
constexpr int kFirRun = 5;

// Two real rows per Bluestein transform: z = x1 + i x2, Z = DFT(z), then
// X1[k] = (Z[k] + conj Z[N-k]) / 2 and X2[k] = -i (Z[k] - conj Z[N-k]) / 2.
// Each spectrum times the window at its fftshifted index kp = (k - m - 1)
// mod N stays in natural order (X1 in buf[0, N), X2 in buf[N, 2N); M >= 2N);
// the prefilter reads spec[kp] = buf[(kp + m + 1) mod N].
template <int MAXB, int NT, typename Tr>
__global__ __launch_bounds__(NT) void dfi_rows_fwd_kernel(
    const double* __restrict__ S, int rows, int ncol1, int pad, int pad_mode,
    int apply_log, const void* __restrict__ plan, int N, int M, size_t o_tw,
    size_t o_chirp, size_t o_bhat, size_t o_win, size_t o_h5,
    double2* __restrict__ X, int* __restrict__ flag) {
    constexpr bool QT = quarter_table<MAXB, NT>();
    extern __shared__ double2 lds[];
    double2* buf = lds;
    double2* tq = lds + M;
    const double2* tw = at<double2>(plan, o_tw);
    const double2* chirp = at<double2>(plan, o_chirp);
    const double2* bhat = at<double2>(plan, o_bhat);
    const double* win = at<double>(plan, o_win);
    const double* h5 = at<double>(plan, o_h5);
    const int tid = threadIdx.x;
    const int m = (N - 1) / 2;
    const int npairs = (rows + 1) / 2;
    if (QT) stage_quarter<NT>(tq, tw, M);
    for (int pr = blockIdx.x; pr < npairs; pr += gridDim.x) {
        const int r1 = 2 * pr;
        const bool two = r1 + 1 < rows;
        const double* s1 = S + (size_t)r1 * ncol1;
        const double* s2 = s1 + ncol1;
        for (int n = tid; n < M; n += NT) {
            double2 a = make_double2(0.0, 0.0);
            if (n < N) {
                int u = n + m;  // ifftshift
                if (u >= N) u -= N;
                const int c = pad_source(u - pad, ncol1, pad_mode);
                double v1 = c >= 0 ? s1[c] : 0.0;
                double v2 = (two && c >= 0) ? s2[c] : 0.0;
                if (apply_log) {
                    v1 = minus_log<Tr>(v1, flag);
                    if (two) v2 = minus_log<Tr>(v2, flag);
                }
                const double2 ch = chirp[n];
                a = make_double2(ch.x * v1 - ch.y * v2, ch.x * v2 + ch.y * v1);
            }
            buf[n] = a;
        }
        __syncthreads();
        bluestein_core<MAXB, QT, NT>(buf, M, tw, tq, bhat);
        for (int k = tid; k <= m; k += NT) {
            const int kn = k == 0 ? 0 : N - k;
            const double2 zk = cmul(chirp[k], buf[k]);
            const double2 zn = cmul(chirp[kn], buf[kn]);
            const double2 x1 =
                make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
            const double2 x2 =
                make_double2(0.5 * (zk.y + zn.y), 0.5 * (zn.x - zk.x));
            const int pk = k > m ? k - m - 1 : k + m;  // (k - m - 1) mod N
            const int pn = kn > m ? kn - m - 1 : kn + m;
            buf[k] = cscale(x1, win[pk]);
            buf[N + k] = cscale(x2, win[pk]);
            if (k != 0) {
                buf[kn] = cscale(conjg(x1), win[pn]);
                buf[N + kn] = cscale(conjg(x2), win[pn]);
            }
        }
        __syncthreads();
        // quintic prefilter along each row (half-sample symmetric in kp),
        // kFirRun outputs per thread sharing their loads
        double2* d1 = X + (size_t)r1 * N;
        double2* d2 = d1 + N;
        const int ngroups = (N + kFirRun - 1) / kFirRun;
        for (int g = tid; g < ngroups; g += NT) {
            const int k0 = g * kFirRun;
            double2 a1[kFirRun], a2[kFirRun];
#pragma unroll
            for (int r = 0; r < kFirRun; ++r) {
                a1[r] = make_double2(0.0, 0.0);
                a2[r] = make_double2(0.0, 0.0);
            }
            for (int i = 0; i < kFirRun + 2 * kK5; ++i) {
                int src = fold_half(k0 - kK5 + i, N) + m + 1;
                if (src >= N) src -= N;
                const double2 v1 = buf[src];
                const double2 v2 = buf[N + src];
#pragma unroll
                for (int r = 0; r < kFirRun; ++r) {
                    int d = i - kK5 - r;
                    d = d < 0 ? -d : d;
                    if (d <= kK5) {
                        const double h = h5[d];
                        a1[r].x += h * v1.x;
                        a1[r].y += h * v1.y;
                        a2[r].x += h * v2.x;
                        a2[r].y += h * v2.y;
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < kFirRun; ++r) {
                if (k0 + r < N) {
                    d1[k0 + r] = a1[r];
                    if (two) d2[k0 + r] = a2[r];
                }
            }
        }
        __syncthreads();
    }
}

// ---- 4. column prefilter ----------------------------------------------------
// out[r][c] = sum_t h5[|t|] X[fold_half(r + t)][c]: one column per lane,
// kColRun consecutive rows per thread (each loaded row feeds every output
// within kK5 of it).
constexpr int kColRun = 16;

__global__ __launch_bounds__(kThreads) void dfi_cols_kernel(
    const double2* __restrict__ X, int rows, int N,
    const double* __restrict__ h5, double2* __restrict__ out) {
    __shared__ double hs[kK5 + 1];
    if (threadIdx.x <= kK5) hs[threadIdx.x] = h5[threadIdx.x];
    __syncthreads();
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * kColRun;
    if (c >= N || r0 >= rows) return;
    double2 acc[kColRun];
#pragma unroll
    for (int q = 0; q < kColRun; ++q) acc[q] = make_double2(0.0, 0.0);
    for (int i = 0; i < kColRun + 2 * kK5; ++i) {
        const int r = fold_half(r0 - kK5 + i, rows);
        const double2 v = X[(size_t)r * N + c];
#pragma unroll
        for (int q = 0; q < kColRun; ++q) {
            int d = i - kK5 - q;
            d = d < 0 ? -d : d;
            if (d <= kK5) {
                const double h = hs[d];
                acc[q].x += h * v.x;
                acc[q].y += h * v.y;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kColRun; ++q)
        if (r0 + q < rows) out[(size_t)(r0 + q) * N + c] = acc[q];
}

// ---- 5. gridding --------------------------------------------------------------
// The float32 (theta, r) of generate_mapping_coordinate for X = m - x,
// Y = y - m (float64 sqrt / atan2 on the integer grid, rounded as the
// reference rounds them), as the (row, column) position in coef.
__device__ __forceinline__ void grid_coords(int rows, int N, int m, int X,
                                            int Y, double& tt, double& rr) {
    const double r_max = (double)m;
    float r = (float)fmin(fmax(sqrt((double)(X * X + Y * Y)), 0.0), r_max);
    double theta = kPi + atan2((double)Y, (double)X);
    if (theta > kPi) {
        r = -r;
        theta -= kPi;
    }
    rr = (double)(float)fmin(fmax((double)r + r_max, 0.0), (double)(N - 1));
    tt = (double)(float)(theta * ((double)rows - 1.0) / kPi);
}

// sum_ij beta5(tt - k_i) beta5(rr - l_j) coef[k_i][l_j] over the 6 x 6
// nearest knots, folded half-sample.
__device__ __forceinline__ double2 grid_eval(const double2* __restrict__ coef,
                                             int rows, int N, double tt,
                                             double rr) {
    const int s0 = (int)floor(tt) - 2, s1 = (int)floor(rr) - 2;
    double wx[6];
    int cx[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        wx[j] = beta5(rr - (double)(s1 + j));
        cx[j] = fold_half(s1 + j, N);
    }
    double2 val = make_double2(0.0, 0.0);
#pragma unroll 1
    for (int i = 0; i < 6; ++i) {
        const double wy = beta5(tt - (double)(s0 + i));
        const double2* crow = coef + (size_t)fold_half(s0 + i, rows) * N;
        double2 rs = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const double2 c = crow[cx[j]];
            rs.x += wx[j] * c.x;
            rs.y += wx[j] * c.y;
        }
        val.x += wy * rs.x;
        val.y += wy * rs.y;
    }
    return val;
}

// T[v][u] = H[u][v] for the columns v <= m, H the Hermitian part
// (Gs[u][v] + conj Gs[-u][-v]) / 2 of Gs = ifftshift(G) (Gs[u][v] =
// G[(u + m) % N][(v + m) % N]), transposed; G[y][x] = mask ? sum beta5
// beta5 coef : 0 at (theta, r) of (X, Y) = (m - x, y - m).  The reference
// keeps Re(ifft2(Gs)) = ifft2(H); the columns v > m of H are the conjugates
// of these and are never formed.  (G itself is Hermitian only to the
// float32 rounding of its coordinates, r' = m -+ r, hence the explicit
// average.)  The mirrored frequency (-X, -Y) has the same theta and r' near
// 2m - r; every coef row is conjugate-symmetric about column m (spectra of
// real rows, symmetric window, symmetric prefilter and folds), so when its
// rounded theta is the same and its knot window the mirror image, conj of
// its value reuses the first point's 36 coefficients in reversed column
// order; otherwise it is gathered on its own.
__global__ __launch_bounds__(kThreads) void dfi_grid_kernel(
    const double2* __restrict__ coef, int rows, int N,
    double2* __restrict__ T) {
    const int m = (N - 1) / 2;
    for (int v = blockIdx.y; v <= m; v += gridDim.y) {
        int x = v + m;
        if (x >= N) x -= N;
        const int X = m - x;
        double2* dst = T + (size_t)v * N;
        for (int u = blockIdx.x * kThreads + threadIdx.x; u < N;
             u += gridDim.x * kThreads) {
            int y = u + m;
            if (y >= N) y -= N;
            const int Y = y - m;
            double2 h = make_double2(0.0, 0.0);
            if (X * X + Y * Y <= m * m) {
                double tta, rra, ttb, rrb;
                grid_coords(rows, N, m, X, Y, tta, rra);
                grid_coords(rows, N, m, -X, -Y, ttb, rrb);
                const int s1 = (int)floor(rra) - 2;
                const int s1b = (int)floor(rrb) - 2;
                if (tta == ttb && s1b == 2 * m - s1 - 5) {
                    const int s0 = (int)floor(tta) - 2;
                    double wa[6], wb[6];
                    int cx[6];
#pragma unroll
                    for (int j = 0; j < 6; ++j) {
                        wa[j] = beta5(rra - (double)(s1 + j));
                        wb[j] = beta5(rrb - (double)(s1b + 5 - j));
                        cx[j] = fold_half(s1 + j, N);
                    }
                    double2 sa = make_double2(0.0, 0.0);
                    double2 sb = make_double2(0.0, 0.0);
#pragma unroll 1
                    for (int i = 0; i < 6; ++i) {
                        const double wy = beta5(tta - (double)(s0 + i));
                        const double2* crow =
                            coef + (size_t)fold_half(s0 + i, rows) * N;
                        double2 ra = make_double2(0.0, 0.0);
                        double2 rb = make_double2(0.0, 0.0);
#pragma unroll
                        for (int j = 0; j < 6; ++j) {
                            const double2 c = crow[cx[j]];
                            ra.x += wa[j] * c.x;
                            ra.y += wa[j] * c.y;
                            rb.x += wb[j] * c.x;
                            rb.y += wb[j] * c.y;
                        }
                        sa.x += wy * ra.x;
                        sa.y += wy * ra.y;
                        sb.x += wy * rb.x;
                        sb.y += wy * rb.y;
                    }
                    // sb = conj of the mirrored frequency's value
                    h = make_double2(0.5 * (sa.x + sb.x), 0.5 * (sa.y + sb.y));
                } else {
                    const double2 a = grid_eval(coef, rows, N, tta, rra);
                    const double2 b = grid_eval(coef, rows, N, ttb, rrb);
                    h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
                }
            }
            dst[u] = h;
        }
    }
}

// ---- 6. inverse DFTs ----------------------------------------------------------
// Output index kc (crop) <- DFT index (kc + pad + m + 1) mod N (fftshift).
__device__ __forceinline__ int crop_index(int kc, int pad, int m, int N) {
    int k = kc + pad + m + 1;  // < 2N
    return k >= N ? k - N : k;
}

// Stage A: inverse DFT (transform / N) along the first axis for the columns
// v <= m: rows of `in` (nrows x N), by Bluestein on the conjugate; ncol
// outputs kept per row -> out (nrows x ncol).
template <int MAXB, int NT>
__global__ __launch_bounds__(NT) void dfi_cols_inv_kernel(
    const double2* __restrict__ in, int nrows, int ncol, int pad,
    const void* __restrict__ plan, int N, int M, size_t o_tw, size_t o_chirp,
    size_t o_bhat, double2* __restrict__ out) {
    constexpr bool QT = quarter_table<MAXB, NT>();
    extern __shared__ double2 lds[];
    double2* buf = lds;
    double2* tq = lds + M;
    const double2* tw = at<double2>(plan, o_tw);
    const double2* chirp = at<double2>(plan, o_chirp);
    const double2* bhat = at<double2>(plan, o_bhat);
    const int tid = threadIdx.x;
    const int m = (N - 1) / 2;
    const double invN = 1.0 / (double)N;
    if (QT) stage_quarter<NT>(tq, tw, M);
    for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
        const double2* src = in + (size_t)row * N;
        for (int n = tid; n < M; n += NT)
            buf[n] = n < N ? cmul(chirp[n], conjg(src[n]))
                           : make_double2(0.0, 0.0);
        __syncthreads();
        bluestein_core<MAXB, QT, NT>(buf, M, tw, tq, bhat);
        for (int kc = tid; kc < ncol; kc += NT) {
            const int k = crop_index(kc, pad, m, N);
            out[(size_t)row * ncol + kc] =
                cscale(conjg(cmul(chirp[k], buf[k])), invN);
        }
        __syncthreads();
    }
}

// Stage B: each output row Y has the first-pass values a[v] for v <= m
// (row Y of `in`, m + 1 values) and a[N - v] = conj(a[v]) (Hermitian), so
// its inverse DFT is real.  Two rows per transform: c = a1 + i a2 ->
// IDFT(c) = y1 + i y2.  Real parts cropped, the ratio mask (mask_r2 < 0:
// none; zero times the value keeps its sign, as the reference's float
// mask product), stored as To at out[Y * out_row_stride + kc].
template <int MAXB, int NT, typename To>
__global__ __launch_bounds__(NT) void dfi_rows_out_kernel(
    const double2* __restrict__ in, int ncol, int pad,
    const void* __restrict__ plan, int N, int M, size_t o_tw, size_t o_chirp,
    size_t o_bhat, To* __restrict__ out, long long out_row_stride,
    double mask_r2) {
    constexpr bool QT = quarter_table<MAXB, NT>();
    extern __shared__ double2 lds[];
    double2* buf = lds;
    double2* tq = lds + M;
    const double2* tw = at<double2>(plan, o_tw);
    const double2* chirp = at<double2>(plan, o_chirp);
    const double2* bhat = at<double2>(plan, o_bhat);
    const int tid = threadIdx.x;
    const int m = (N - 1) / 2;
    const int h = m + 1;
    const double invN = 1.0 / (double)N;
    const int mc = ncol / 2;
    const int npairs = (ncol + 1) / 2;
    if (QT) stage_quarter<NT>(tq, tw, M);
    for (int pr = blockIdx.x; pr < npairs; pr += gridDim.x) {
        const int y1 = 2 * pr;
        const bool two = y1 + 1 < ncol;
        const double2* a1 = in + (size_t)y1 * h;
        const double2* a2 = a1 + h;
        for (int n = tid; n < M; n += NT) {
            double2 c = make_double2(0.0, 0.0);
            if (n < N) {
                const double2 p = n <= m ? a1[n] : conjg(a1[N - n]);
                double2 q = make_double2(0.0, 0.0);
                if (two) q = n <= m ? a2[n] : conjg(a2[N - n]);
                // conj(p + i q), times the chirp
                c = cmul(chirp[n], make_double2(p.x - q.y, -(p.y + q.x)));
            }
            buf[n] = c;
        }
        __syncthreads();
        bluestein_core<MAXB, QT, NT>(buf, M, tw, tq, bhat);
        for (int kc = tid; kc < ncol; kc += NT) {
            const int k = crop_index(kc, pad, m, N);
            const double2 y = conjg(cmul(chirp[k], buf[k]));
            double v1 = y.x * invN, v2 = y.y * invN;
            if (mask_r2 >= 0.0) {
                const int dx = kc - mc, d1 = y1 - mc, d2 = y1 + 1 - mc;
                if ((double)(d1 * d1 + dx * dx) > mask_r2) v1 = v1 * 0.0;
                if ((double)(d2 * d2 + dx * dx) > mask_r2) v2 = v2 * 0.0;
            }
            out[(long long)y1 * out_row_stride + kc] = (To)v1;
            if (two) out[(long long)(y1 + 1) * out_row_stride + kc] = (To)v2;
        }
        __syncthreads();
    }
}

// A (rows x cols) -> A^T (cols x rows), 32 x 32 LDS tiles
__global__ __launch_bounds__(kThreads) void dfi_transpose_kernel(
    const double2* __restrict__ A, int rows, int cols,
    double2* __restrict__ At) {
    __shared__ double2 tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int tr = blockIdx.y; tr * 32 < rows; tr += gridDim.y)
        for (int tc = blockIdx.x; tc * 32 < cols; tc += gridDim.x) {
            for (int k = ty; k < 32; k += 8) {
                const int r = tr * 32 + k, c = tc * 32 + tx;
                if (r < rows && c < cols)
                    tile[k][tx] = A[(size_t)r * cols + c];
            }
            __syncthreads();
            for (int k = ty; k < 32; k += 8) {
                const int c = tc * 32 + k, r = tr * 32 + tx;
                if (r < rows && c < cols)
                    At[(size_t)c * rows + r] = tile[tx][k];
            }
            __syncthreads();
        }
}

// ---- 7. Bluestein lengths 16384 and 32768 (padded widths 4097 .. 16383) ----
// The convolution of length M = LV * H (LV = 2 or 4) runs as H-point LDS
// transforms with a per-workgroup global scratch of M values (conv_big and
// conv_big4 in fft_lds.h); bhat carries 1/M.
// bhat = FFT_M(b) / M for M = 2H (one workgroup; b as in bhat_kernel)
__global__ __launch_bounds__(kBigNT) void bhat_big_kernel(
    const double2* __restrict__ chirp, const double2* __restrict__ tw, int N,
    double2* __restrict__ bhat) {
    extern __shared__ double2 lds[];
    constexpr int H = kBigH, NT = kBigNT, M = 2 * kBigH;
    const int tid = threadIdx.x;
    auto b = [&](int i) -> double2 {
        if (i < N) return conjg(chirp[i]);
        if (i > M - N) return conjg(chirp[M - i]);
        return make_double2(0.0, 0.0);
    };
    for (int j = tid; j < H; j += NT) lds[j] = b(2 * j);
    __syncthreads();
    lds_fft<-1, kBigB, false, NT>(lds, H, tw);
    for (int k = tid; k < H; k += NT) bhat[k] = lds[k];
    __syncthreads();
    for (int j = tid; j < H; j += NT) lds[j] = b(2 * j + 1);
    __syncthreads();
    lds_fft<-1, kBigB, false, NT>(lds, H, tw);
    const double inv = 1.0 / (double)M;
    for (int k = tid; k < H; k += NT) {
        const double2 e = bhat[k];
        const double2 o = cmul(lds[k], tw[k]);
        bhat[k] = cscale(cadd(e, o), inv);
        bhat[k + H] = cscale(csub(e, o), inv);
    }
}

// bhat = FFT_M(b) / M for M = 4H: the quarters' transforms through bhat
// itself (F_0..F_2 in its first 3H values), then the radix-4 combination,
// each thread owning the four outputs of its k.
__global__ __launch_bounds__(kBigNT) void bhat_big4_kernel(
    const double2* __restrict__ chirp, const double2* __restrict__ tw, int N,
    double2* __restrict__ bhat) {
    extern __shared__ double2 lds[];
    constexpr int H = kBigH, NT = kBigNT, M = kBig4M;
    auto b = [&](int i) -> double2 {
        if (i < N) return conjg(chirp[i]);
        if (i > M - N) return conjg(chirp[M - i]);
        return make_double2(0.0, 0.0);
    };
    fft_big4_quarters(lds, bhat, M, tw, b);
    const double inv = 1.0 / (double)M;
    for (int k = threadIdx.x; k < H; k += NT) {
        double2 x[4];
        big4_combine(bhat[k], bhat[H + k], bhat[2 * H + k], lds[k], k, tw, x);
#pragma unroll
        for (int q = 0; q < 4; ++q) bhat[k + q * H] = cscale(x[q], inv);
    }
}

// Row pairs: pad, -log, the two rows' spectra split and windowed into rows
// r1 and r1 + 1 of Xn in natural order (the row prefilter follows in
// dfi_rows_fir_kernel).
template <int LV, typename Tr>
__global__ __launch_bounds__(kBigNT) void dfi_rows_fwd_big_kernel(
    const double* __restrict__ S, int rows, int ncol1, int pad, int pad_mode,
    int apply_log, const void* __restrict__ plan, int N, size_t o_tw,
    size_t o_chirp, size_t o_bhat, size_t o_win, double2* __restrict__ Xn,
    double2* __restrict__ scratch_all, int* __restrict__ flag) {
    extern __shared__ double2 buf[];
    double2* scratch = scratch_all + (size_t)blockIdx.x * LV * kBigH;
    const double2* tw = at<double2>(plan, o_tw);
    const double2* chirp = at<double2>(plan, o_chirp);
    const double2* bhat = at<double2>(plan, o_bhat);
    const double* win = at<double>(plan, o_win);
    const int tid = threadIdx.x;
    const int m = (N - 1) / 2;
    const int npairs = (rows + 1) / 2;
    for (int pr = blockIdx.x; pr < npairs; pr += gridDim.x) {
        const int r1 = 2 * pr;
        const bool two = r1 + 1 < rows;
        const double* s1 = S + (size_t)r1 * ncol1;
        const double* s2 = s1 + ncol1;
        auto load = [&](int n) -> double2 {
            int u = n + m;  // ifftshift
            if (u >= N) u -= N;
            const int c = pad_source(u - pad, ncol1, pad_mode);
            double v1 = c >= 0 ? s1[c] : 0.0;
            double v2 = (two && c >= 0) ? s2[c] : 0.0;
            if (apply_log) {
                v1 = minus_log<Tr>(v1, flag);
                if (two) v2 = minus_log<Tr>(v2, flag);
            }
            const double2 ch = chirp[n];
            return make_double2(ch.x * v1 - ch.y * v2, ch.x * v2 + ch.y * v1);
        };
        conv_lv<LV>(buf, scratch, N, tw, bhat, load);
        double2* d1 = Xn + (size_t)r1 * N;
        double2* d2 = d1 + N;
        for (int k = tid; k <= m; k += kBigNT) {
            const int kn = k == 0 ? 0 : N - k;
            const double2 zk =
                cmul(chirp[k], conv_lv_out<LV>(buf, scratch, k));
            const double2 zn =
                cmul(chirp[kn], conv_lv_out<LV>(buf, scratch, kn));
            const double2 x1 =
                make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
            const double2 x2 =
                make_double2(0.5 * (zk.y + zn.y), 0.5 * (zn.x - zk.x));
            const int pk = k > m ? k - m - 1 : k + m;  // (k - m - 1) mod N
            const int pn = kn > m ? kn - m - 1 : kn + m;
            d1[k] = cscale(x1, win[pk]);
            if (two) d2[k] = cscale(x2, win[pk]);
            if (k != 0) {
                d1[kn] = cscale(conjg(x1), win[pn]);
                if (two) d2[kn] = cscale(conjg(x2), win[pn]);
            }
        }
        __syncthreads();
    }
}

// The quintic prefilter along rows of Xn (natural order; spec[kp] =
// Xn[(kp + m + 1) mod N]) into X (fftshifted order).  A workgroup stages
// the spec segment its kFirSpan outputs need (plus kK5 either side) in LDS,
// then each thread forms kFirRun consecutive outputs from it.
constexpr int kFirSpan = kThreads * kFirRun;

__global__ __launch_bounds__(kThreads) void dfi_rows_fir_kernel(
    const double2* __restrict__ Xn, int rows, int N,
    const double* __restrict__ h5, double2* __restrict__ X) {
    __shared__ double2 seg[kFirSpan + 2 * kK5];
    const int m = (N - 1) / 2;
    const int tid = threadIdx.x;
    for (int row = blockIdx.y; row < rows; row += gridDim.y) {
        const double2* src = Xn + (size_t)row * N;
        double2* dst = X + (size_t)row * N;
        for (int c0 = blockIdx.x * kFirSpan; c0 < N;
             c0 += gridDim.x * kFirSpan) {
            for (int i = tid; i < kFirSpan + 2 * kK5; i += kThreads) {
                int q = fold_half(c0 - kK5 + i, N) + m + 1;
                if (q >= N) q -= N;
                seg[i] = src[q];
            }
            __syncthreads();
            const int l0 = tid * kFirRun;  // local index of the first output
            double2 a[kFirRun];
#pragma unroll
            for (int r = 0; r < kFirRun; ++r) a[r] = make_double2(0.0, 0.0);
            for (int i = 0; i < kFirRun + 2 * kK5; ++i) {
                const double2 v = seg[l0 + i];
#pragma unroll
                for (int r = 0; r < kFirRun; ++r) {
                    int d = i - kK5 - r;
                    d = d < 0 ? -d : d;
                    if (d <= kK5) {
                        const double h = h5[d];
                        a[r].x += h * v.x;
                        a[r].y += h * v.y;
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < kFirRun; ++r)
                if (c0 + l0 + r < N) dst[c0 + l0 + r] = a[r];
            __syncthreads();
        }
    }
}

// Stage A at M = LV * H (as dfi_cols_inv_kernel).
template <int LV>
__global__ __launch_bounds__(kBigNT) void dfi_cols_inv_big_kernel(
    const double2* __restrict__ in, int nrows, int ncol, int pad,
    const void* __restrict__ plan, int N, size_t o_tw, size_t o_chirp,
    size_t o_bhat, double2* __restrict__ out,
    double2* __restrict__ scratch_all) {
    extern __shared__ double2 buf[];
    double2* scratch = scratch_all + (size_t)blockIdx.x * LV * kBigH;
    const double2* tw = at<double2>(plan, o_tw);
    const double2* chirp = at<double2>(plan, o_chirp);
    const double2* bhat = at<double2>(plan, o_bhat);
    const int m = (N - 1) / 2;
    const double invN = 1.0 / (double)N;
    for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
        const double2* src = in + (size_t)row * N;
        auto load = [&](int n) -> double2 {
            return cmul(chirp[n], conjg(src[n]));
        };
        conv_lv<LV>(buf, scratch, N, tw, bhat, load);
        for (int kc = threadIdx.x; kc < ncol; kc += kBigNT) {
            const int k = crop_index(kc, pad, m, N);
            out[(size_t)row * ncol + kc] = cscale(
                conjg(cmul(chirp[k], conv_lv_out<LV>(buf, scratch, k))),
                invN);
        }
        __syncthreads();
    }
}

// Stage B at M = LV * H (as dfi_rows_out_kernel).
template <int LV, typename To>
__global__ __launch_bounds__(kBigNT) void dfi_rows_out_big_kernel(
    const double2* __restrict__ in, int ncol, int pad,
    const void* __restrict__ plan, int N, size_t o_tw, size_t o_chirp,
    size_t o_bhat, To* __restrict__ out, long long out_row_stride,
    double mask_r2, double2* __restrict__ scratch_all) {
    extern __shared__ double2 buf[];
    double2* scratch = scratch_all + (size_t)blockIdx.x * LV * kBigH;
    const double2* tw = at<double2>(plan, o_tw);
    const double2* chirp = at<double2>(plan, o_chirp);
    const double2* bhat = at<double2>(plan, o_bhat);
    const int m = (N - 1) / 2;
    const int h = m + 1;
    const double invN = 1.0 / (double)N;
    const int mc = ncol / 2;
    const int npairs = (ncol + 1) / 2;
    for (int pr = blockIdx.x; pr < npairs; pr += gridDim.x) {
        const int y1 = 2 * pr;
        const bool two = y1 + 1 < ncol;
        const double2* a1 = in + (size_t)y1 * h;
        const double2* a2 = a1 + h;
        auto load = [&](int n) -> double2 {
            const double2 p = n <= m ? a1[n] : conjg(a1[N - n]);
            double2 q = make_double2(0.0, 0.0);
            if (two) q = n <= m ? a2[n] : conjg(a2[N - n]);
            return cmul(chirp[n], make_double2(p.x - q.y, -(p.y + q.x)));
        };
        conv_lv<LV>(buf, scratch, N, tw, bhat, load);
        for (int kc = threadIdx.x; kc < ncol; kc += kBigNT) {
            const int k = crop_index(kc, pad, m, N);
            const double2 y =
                conjg(cmul(chirp[k], conv_lv_out<LV>(buf, scratch, k)));
            double v1 = y.x * invN, v2 = y.y * invN;
            if (mask_r2 >= 0.0) {
                const int dx = kc - mc, d1 = y1 - mc, d2 = y1 + 1 - mc;
                if ((double)(d1 * d1 + dx * dx) > mask_r2) v1 = v1 * 0.0;
                if ((double)(d2 * d2 + dx * dx) > mask_r2) v2 = v2 * 0.0;
            }
            out[(long long)y1 * out_row_stride + kc] = (To)v1;
            if (two) out[(long long)(y1 + 1) * out_row_stride + kc] = (To)v2;
        }
        __syncthreads();
    }
}

// ---- dispatch ---------------------------------------------------------------------
// Transform lengths up to 4096 run with 256 threads (MAXB butterflies each,
// the quarter twiddle table in LDS); 8192 runs with 1024 threads of 2
// butterflies: four waves per SIMD for the single workgroup that its 128 KB
// of LDS leaves room for (measured 1.6x faster than 256 threads of 8).
inline size_t fft_lds_bytes(int M) {
    return (size_t)M * sizeof(double2) +
           (M <= 4096 ? (size_t)(M / 4) * sizeof(double2) : 0);
}

template <int MAXB, int NT, typename Tr>
int launch_rows_fwd(const double* S, int rows, int ncol1, int pad,
                    int pad_mode, int apply_log, const void* plan,
                    const DfiPlan& p, double2* X, int* flag, hipStream_t s) {
    auto k = dfi_rows_fwd_kernel<MAXB, NT, Tr>;
    const size_t lds = fft_lds_bytes(p.M);
    prepare_lds(k, lds);
    const int grid = std::min((rows + 1) / 2, resident_blocks(k, lds, NT));
    hipLaunchKernelGGL(k, dim3(grid), dim3(NT), lds, s, S, rows, ncol1, pad,
                       pad_mode, apply_log, plan, p.N, p.M, p.tw, p.chirp,
                       p.bhat, p.window, p.h5, X, flag);
    return check_hip(hipGetLastError(), "dfi_rows_fwd_kernel launch");
}

template <int MAXB, int NT>
int launch_cols_inv(const double2* in, int nrows, int ncol, int pad,
                    const void* plan, const DfiPlan& p, double2* out,
                    hipStream_t s) {
    auto k = dfi_cols_inv_kernel<MAXB, NT>;
    const size_t lds = fft_lds_bytes(p.M);
    prepare_lds(k, lds);
    const int grid = std::min(nrows, resident_blocks(k, lds, NT));
    hipLaunchKernelGGL(k, dim3(grid), dim3(NT), lds, s, in, nrows, ncol, pad,
                       plan, p.N, p.M, p.tw, p.chirp, p.bhat, out);
    return check_hip(hipGetLastError(), "dfi_cols_inv_kernel launch");
}

template <int MAXB, int NT, typename To>
int launch_rows_out(const double2* in, int ncol, int pad, const void* plan,
                    const DfiPlan& p, To* out, long long out_row_stride,
                    double mask_r2, hipStream_t s) {
    auto k = dfi_rows_out_kernel<MAXB, NT, To>;
    const size_t lds = fft_lds_bytes(p.M);
    prepare_lds(k, lds);
    const int grid = std::min((ncol + 1) / 2, resident_blocks(k, lds, NT));
    hipLaunchKernelGGL(k, dim3(grid), dim3(NT), lds, s, in, ncol, pad, plan,
                       p.N, p.M, p.tw, p.chirp, p.bhat, out, out_row_stride,
                       mask_r2);
    return check_hip(hipGetLastError(), "dfi_rows_out_kernel launch");
}

template <typename Tr>
int rows_fwd(int maxb, const double* S, int rows, int ncol1, int pad,
             int pad_mode, int apply_log, const void* plan, const DfiPlan& p,
             double2* X, int* flag, hipStream_t s) {
    switch (maxb) {
    case 1: return launch_rows_fwd<1, 256, Tr>(S, rows, ncol1, pad, pad_mode, apply_log, plan, p, X, flag, s);
    case 2: return launch_rows_fwd<2, 256, Tr>(S, rows, ncol1, pad, pad_mode, apply_log, plan, p, X, flag, s);
    case 4: return launch_rows_fwd<4, 256, Tr>(S, rows, ncol1, pad, pad_mode, apply_log, plan, p, X, flag, s);
    default: return launch_rows_fwd<2, 1024, Tr>(S, rows, ncol1, pad, pad_mode, apply_log, plan, p, X, flag, s);
    }
}

int cols_inv(int maxb, const double2* in, int nrows, int ncol, int pad,
             const void* plan, const DfiPlan& p, double2* out, hipStream_t s) {
    switch (maxb) {
    case 1: return launch_cols_inv<1, 256>(in, nrows, ncol, pad, plan, p, out, s);
    case 2: return launch_cols_inv<2, 256>(in, nrows, ncol, pad, plan, p, out, s);
    case 4: return launch_cols_inv<4, 256>(in, nrows, ncol, pad, plan, p, out, s);
    default: return launch_cols_inv<2, 1024>(in, nrows, ncol, pad, plan, p, out, s);
    }
}

template <typename To>
int rows_out(int maxb, const double2* in, int ncol, int pad, const void* plan,
             const DfiPlan& p, To* out, long long out_row_stride,
             double mask_r2, hipStream_t s) {
    switch (maxb) {
    case 1: return launch_rows_out<1, 256, To>(in, ncol, pad, plan, p, out, out_row_stride, mask_r2, s);
    case 2: return launch_rows_out<2, 256, To>(in, ncol, pad, plan, p, out, out_row_stride, mask_r2, s);
    case 4: return launch_rows_out<4, 256, To>(in, ncol, pad, plan, p, out, out_row_stride, mask_r2, s);
    default: return launch_rows_out<2, 1024, To>(in, ncol, pad, plan, p, out, out_row_stride, mask_r2, s);
    }
}

// The M = LV * H stages: one workgroup of kBigNT threads per CU, each with
// its scratch slot.
template <int LV, typename Tr>
int launch_rows_fwd_big(const double* S, int nrow, int ncol1, int pad,
                        int pad_mode, int apply_log, const void* plan,
                        const DfiPlan& p, double2* X, double2* scr, int nscr,
                        int* flag, hipStream_t s) {
    const size_t lds = (size_t)kBigH * sizeof(double2);
    auto k = dfi_rows_fwd_big_kernel<LV, Tr>;
    prepare_lds(k, lds);
    const int grid = std::min((nrow + 1) / 2, nscr);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kBigNT), lds, s, S, nrow, ncol1,
                       pad, pad_mode, apply_log, plan, p.N, p.tw, p.chirp,
                       p.bhat, p.window, X, scr, flag);
    return check_hip(hipGetLastError(), "dfi_rows_fwd_big_kernel launch");
}

template <int LV>
int launch_cols_inv_big(const double2* in, int h, int ncol, int pad,
                        const void* plan, const DfiPlan& p, double2* out,
                        double2* scr, int nscr, hipStream_t s) {
    const size_t lds = (size_t)kBigH * sizeof(double2);
    auto k = dfi_cols_inv_big_kernel<LV>;
    prepare_lds(k, lds);
    hipLaunchKernelGGL(k, dim3(std::min(h, nscr)), dim3(kBigNT), lds, s, in,
                       h, ncol, pad, plan, p.N, p.tw, p.chirp, p.bhat, out,
                       scr);
    return check_hip(hipGetLastError(), "dfi_cols_inv_big_kernel launch");
}

template <int LV, typename To>
int launch_rows_out_big(const double2* in, int ncol, int pad,
                        const void* plan, const DfiPlan& p, To* out,
                        long long out_row_stride, double mask_r2,
                        double2* scr, int nscr, hipStream_t s) {
    const size_t lds = (size_t)kBigH * sizeof(double2);
    auto k = dfi_rows_out_big_kernel<LV, To>;
    prepare_lds(k, lds);
    const int grid = std::min((ncol + 1) / 2, nscr);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kBigNT), lds, s, in, ncol, pad,
                       plan, p.N, p.tw, p.chirp, p.bhat, out, out_row_stride,
                       mask_r2, scr);
    return check_hip(hipGetLastError(), "dfi_rows_out_big_kernel launch");
}

// Workspace: S1 (shifted sinogram), S2 (angle-handled), and two ping-pong
// regions: R1 holds the row coefficients (rows x N), then the half grid
// ((m + 1) x N), then the transposed first pass (ncol x (m + 1)); R2 holds
// the column coefficients (rows x N), then the first pass ((m + 1) x ncol).
struct DfiWork {
    size_t s1, s2, r1, r2, scr, total;
    int nscr;  // scratch slots (workgroups) of M values for M > 8192
};

// One M-value scratch per compute unit (the big kernels' 128 KB of LDS
// allow one workgroup per CU).
inline int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess || cus < 1) {
        (void)hipGetLastError();
        return 256;
    }
    return cus;
}

DfiWork work_layout(int nrow_in, int ncol, int nrow, int N) {
    const int ncol1 = ncol + (ncol % 2 == 0 ? 1 : 0);
    const size_t h = (size_t)(N - 1) / 2 + 1;
    DfiWork w{};
    size_t o = 0;
    w.s1 = o; o = align256(o + (size_t)nrow_in * ncol1 * sizeof(double));
    w.s2 = o; o = align256(o + (size_t)nrow * ncol1 * sizeof(double));
    const size_t r1 = std::max({(size_t)nrow * N, h * N, (size_t)ncol * h});
    const size_t r2 = std::max((size_t)nrow * N, h * ncol);
    w.r1 = o; o = align256(o + r1 * sizeof(double2));
    w.r2 = o; o = align256(o + r2 * sizeof(double2));
    const int M = bluestein_length(N);
    w.nscr = M > kBigH ? device_cus() : 0;
    w.scr = o; o = align256(o + (size_t)w.nscr * M * sizeof(double2));
    w.total = o;
    return w;
}

template <typename Tin, typename Tr>
int shift_rows(const Tin* in, long long row_stride, int nrow, int ncol,
               int ncol1, double xshift, const double* h3, double* S,
               hipStream_t s) {
    if (ncol1 + 2 * kNearestPad > kMaxLine) {
        const int nseg = (ncol1 + kShiftSeg - 1) / kShiftSeg;
        const int grid = (int)std::min<long long>((long long)nrow * nseg,
                                                  8192);
        hipLaunchKernelGGL((dfi_shift_seg_kernel<Tin, Tr>), dim3(grid),
                           dim3(kThreads), 0, s, in, row_stride, nrow, ncol,
                           ncol1, xshift, h3, S);
        return check_hip(hipGetLastError(), "dfi_shift_seg_kernel launch");
    }
    auto k = dfi_shift_kernel<Tin, Tr>;
    const size_t lds = 2 * (size_t)(ncol1 + 2 * kNearestPad) * sizeof(double);
    prepare_lds(k, 2 * (size_t)kMaxLine * sizeof(double));
    const int grid = std::min(nrow, 4096);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), lds, s, in, row_stride,
                       nrow, ncol, ncol1, xshift, h3, S);
    return check_hip(hipGetLastError(), "dfi_shift_kernel launch");
}

}  // namespace

size_t dfi_plan_bytes(int width) {
    if (width < 3 || width % 2 == 0 || bluestein_length(width) > kBig4M)
        return 0;
    return plan_layout(width).total;
}

int dfi_plan_build(const double* window, int width, void* plan,
                   hipStream_t s) {
    const DfiPlan p = plan_layout(width);
    const bool big = p.M > kBigH;
    // the FFT's table: length-M transforms take exp(-2 pi i n / 2M); the
    // length-2H path runs H-point transforms, exp(-2 pi i n / 2H)
    const int L = big ? p.M : 2 * p.M;
    hipLaunchKernelGGL(twiddle_kernel, dim3((L + 255) / 256), dim3(256), 0, s,
                       at<double2>(plan, p.tw), L);
    hipLaunchKernelGGL(chirp_kernel, dim3((p.N + 255) / 256), dim3(256), 0, s,
                       at<double2>(plan, p.chirp), p.N);
    const int tab = std::max(p.N, kK5 + 1);
    hipLaunchKernelGGL(dfi_tables_kernel, dim3((tab + 255) / 256), dim3(256),
                       0, s, window, p.N, at<double>(plan, p.window),
                       at<double>(plan, p.h3), at<double>(plan, p.h5));
    if (big) {
        const size_t lds_big = (size_t)kBigH * sizeof(double2);
        auto kern = p.M == kBig4M ? bhat_big4_kernel : bhat_big_kernel;
        prepare_lds(kern, lds_big);
        hipLaunchKernelGGL(kern, dim3(1), dim3(kBigNT), lds_big, s,
                           at<double2>(plan, p.chirp),
                           at<double2>(plan, p.tw), p.N,
                           at<double2>(plan, p.bhat));
        return check_hip(hipGetLastError(), "dfi plan kernels");
    }
    const size_t lds = (size_t)p.M * sizeof(double2);
#define ATH_BHAT(B)                                                           \
    {                                                                         \
        prepare_lds(bhat_kernel<B>, lds);                                     \
        hipLaunchKernelGGL(bhat_kernel<B>, dim3(1), dim3(kFftThreads), lds,   \
                           s, at<double2>(plan, p.chirp),                     \
                           at<double2>(plan, p.tw), p.N, p.M,                 \
                           at<double2>(plan, p.bhat));                        \
    }
    switch (maxb_for(p.M)) {
    case 1: ATH_BHAT(1) break;
    case 2: ATH_BHAT(2) break;
    case 4: ATH_BHAT(4) break;
    default: ATH_BHAT(8) break;
    }
#undef ATH_BHAT
    return check_hip(hipGetLastError(), "dfi plan kernels");
}

size_t dfi_workspace_bytes(int nrow_in, int ncol, int nrow, int width) {
    return work_layout(nrow_in, ncol, nrow, width).total;
}

int dfi_reconstruction(const void* sino, int dtype, int nrow_in, int nslice,
                       int ncol, long long row_stride, long long slice_stride,
                       double xshift, int pad, int pad_mode, int apply_log,
                       int angle_mode, double row_shift, int flip, int nrow,
                       const void* plan, int width, void* workspace,
                       void* out, int out_dtype, long long out_row_stride,
                       long long out_slice_stride, double mask_r2, int* flag,
                       hipStream_t s) {
    const DfiPlan p = plan_layout(width);
    const DfiWork w = work_layout(nrow_in, ncol, nrow, width);
    const int ncol1 = ncol + (ncol % 2 == 0 ? 1 : 0);
    const int maxb = maxb_for(p.M);
    const bool f64 = dtype == ATH_F64;
    double* S1 = at<double>(workspace, w.s1);
    double* S2 = at<double>(workspace, w.s2);
    double2* R1 = at<double2>(workspace, w.r1);
    double2* R2 = at<double2>(workspace, w.r2);
    double2* scr = at<double2>(workspace, w.scr);
    const bool big = p.M > kBigH;
    const int lv = p.M / kBigH;  // 2 or 4 when big
    const double* h3 = at<double>(plan, p.h3);
    const double* h5 = at<double>(plan, p.h5);
    const size_t esz = f64 ? sizeof(double) : sizeof(float);
    for (int sl = 0; sl < nslice; ++sl) {
        const void* in = static_cast<const char*>(sino) +
                         (size_t)sl * (size_t)slice_stride * esz;
        int st = f64 ? shift_rows<double, double>(
                           static_cast<const double*>(in), row_stride,
                           nrow_in, ncol, ncol1, xshift, h3, S1, s)
                     : shift_rows<float, float>(
                           static_cast<const float*>(in), row_stride, nrow_in,
                           ncol, ncol1, xshift, h3, S1, s);
        if (st != ATH_OK) return st;
        const double* S = S1;
        if (angle_mode != 0) {
            const dim3 grid = row_grid(nrow, ncol1, kThreads);
            const int fold = angle_mode == 2 ? 1 : 0;
            if (f64)
                hipLaunchKernelGGL(dfi_angles_kernel<double>, grid,
                                   dim3(kThreads), 0, s, S1, nrow_in, ncol1,
                                   nrow, fold, row_shift, flip, h3, S2);
            else
                hipLaunchKernelGGL(dfi_angles_kernel<float>, grid,
                                   dim3(kThreads), 0, s, S1, nrow_in, ncol1,
                                   nrow, fold, row_shift, flip, h3, S2);
            st = check_hip(hipGetLastError(), "dfi_angles_kernel launch");
            if (st != ATH_OK) return st;
            S = S2;
        }
        if (big) {
            // spectra into R2 (natural order), then the row prefilter into R1
            if (f64)
                st = lv == 4 ? launch_rows_fwd_big<4, double>(
                                   S, nrow, ncol1, pad, pad_mode, apply_log,
                                   plan, p, R2, scr, w.nscr, flag, s)
                             : launch_rows_fwd_big<2, double>(
                                   S, nrow, ncol1, pad, pad_mode, apply_log,
                                   plan, p, R2, scr, w.nscr, flag, s);
            else
                st = lv == 4 ? launch_rows_fwd_big<4, float>(
                                   S, nrow, ncol1, pad, pad_mode, apply_log,
                                   plan, p, R2, scr, w.nscr, flag, s)
                             : launch_rows_fwd_big<2, float>(
                                   S, nrow, ncol1, pad, pad_mode, apply_log,
                                   plan, p, R2, scr, w.nscr, flag, s);
            if (st != ATH_OK) return st;
            const dim3 fgrid((p.N + kFirSpan - 1) / kFirSpan,
                             std::min(nrow, 65535));
            hipLaunchKernelGGL(dfi_rows_fir_kernel, fgrid, dim3(kThreads), 0,
                               s, R2, nrow, p.N, h5, R1);
            st = check_hip(hipGetLastError(), "dfi_rows_fir_kernel launch");
        } else {
            st = f64 ? rows_fwd<double>(maxb, S, nrow, ncol1, pad, pad_mode,
                                        apply_log, plan, p, R1, flag, s)
                     : rows_fwd<float>(maxb, S, nrow, ncol1, pad, pad_mode,
                                       apply_log, plan, p, R1, flag, s);
        }
        if (st != ATH_OK) return st;
        {
            const dim3 grid((p.N + 63) / 64,
                            (nrow + 4 * kColRun - 1) / (4 * kColRun));
            hipLaunchKernelGGL(dfi_cols_kernel, grid, dim3(kThreads), 0, s, R1,
                               nrow, p.N, h5, R2);
            st = check_hip(hipGetLastError(), "dfi_cols_kernel launch");
            if (st != ATH_OK) return st;
        }
        const int h = (p.N - 1) / 2 + 1;  // columns v <= m
        hipLaunchKernelGGL(dfi_grid_kernel, row_grid(h, p.N, kThreads),
                           dim3(kThreads), 0, s, R2, nrow, p.N, R1);
        st = check_hip(hipGetLastError(), "dfi_grid_kernel launch");
        if (st != ATH_OK) return st;
        if (big) {
            st = lv == 4 ? launch_cols_inv_big<4>(R1, h, ncol, pad, plan, p,
                                                  R2, scr, w.nscr, s)
                         : launch_cols_inv_big<2>(R1, h, ncol, pad, plan, p,
                                                  R2, scr, w.nscr, s);
        } else {
            st = cols_inv(maxb, R1, h, ncol, pad, plan, p, R2, s);
        }
        if (st != ATH_OK) return st;
        {
            const dim3 grid(std::min((ncol + 31) / 32, 65535),
                            std::min((h + 31) / 32, 65535));
            hipLaunchKernelGGL(dfi_transpose_kernel, grid, dim3(kThreads), 0,
                               s, R2, h, ncol, R1);
            st = check_hip(hipGetLastError(), "dfi_transpose_kernel launch");
            if (st != ATH_OK) return st;
        }
        char* o = static_cast<char*>(out) +
                  (size_t)sl * (size_t)out_slice_stride *
                      (out_dtype == ATH_F64 ? sizeof(double) : sizeof(float));
        if (big) {
            if (out_dtype == ATH_F64) {
                double* od = reinterpret_cast<double*>(o);
                st = lv == 4 ? launch_rows_out_big<4, double>(
                                   R1, ncol, pad, plan, p, od, out_row_stride,
                                   mask_r2, scr, w.nscr, s)
                             : launch_rows_out_big<2, double>(
                                   R1, ncol, pad, plan, p, od, out_row_stride,
                                   mask_r2, scr, w.nscr, s);
            } else {
                float* of = reinterpret_cast<float*>(o);
                st = lv == 4 ? launch_rows_out_big<4, float>(
                                   R1, ncol, pad, plan, p, of, out_row_stride,
                                   mask_r2, scr, w.nscr, s)
                             : launch_rows_out_big<2, float>(
                                   R1, ncol, pad, plan, p, of, out_row_stride,
                                   mask_r2, scr, w.nscr, s);
            }
        } else {
            st = out_dtype == ATH_F64
                     ? rows_out<double>(maxb, R1, ncol, pad, plan, p,
                                        reinterpret_cast<double*>(o),
                                        out_row_stride, mask_r2, s)
                     : rows_out<float>(maxb, R1, ncol, pad, plan, p,
                                       reinterpret_cast<float*>(o),
                                       out_row_stride, mask_r2, s);
        }
        if (st != ATH_OK) return st;
    }
    return ATH_OK;
}

}  // namespace ath
