// 2D circular ramp filter of BPF (back-projection filtering) on gfx950.
//
// algotom's apply_circular_ramp_filter (reconstruction.py:524-560) pads the
// back-projected image by `pad` on every side (numpy.pad, one mode for both
// axes) to W x W, multiplies its shifted 2D spectrum by the circular ramp
// window R2 and keeps the real part of the inverse.  That is a W x W circular
// convolution with the real kernel g2 = Re(ifft2(ifftshift(R2))).  As in the
// 1D filter, it is evaluated exactly as a LINEAR convolution with the kernel
// tiled 2 x 2 (K[i][j] = g2[i mod W][j mod W], i, j < 2W) on a power-of-two
// L x L grid, L >= 2W - pad, whose outputs (W+pad+y, W+pad+x), y, x < ncol,
// never wrap.  All arithmetic is float64, in three passes over one image:
//
//   rows_fwd_kernel   one workgroup per padded row u < W: real FFT of length
//                     L (L/2 packed complex samples in LDS + split), the
//                     L/2+1 bins stored row-major T[u][k];
//   cols_kernel       one workgroup per bin column k <= L/2: the W values
//                     T[.][k] zero-extended to L, complex FFT of length L,
//                     times the kernel spectrum column Khat[k][.], inverse
//                     FFT, the ncol needed rows stored U[y][k];
//   rows_inv_kernel   one workgroup per output row y: Hermitian merge, inverse
//                     real FFT, crop, optional circle mask and pi/(N-1) scale
//                     in the reference's order, float32 or float64 store.
//
// The kernel spectrum Khat (L/2+1 columns of L complex values, 1/(M L)
// folded in) is built once per (W, window) by the same two forward passes
// over the 2W rows of K.  The column transform holds L complex doubles in
// LDS, so L <= 8192 (ncol up to about 3800 with the default pad); L = 16384
// (ncol up to about 7960) runs the column pass as the two-level transform
// of fft_lds.h (cols_big_kernel) with a per-workgroup global scratch.
// Kernel samples at index >= L are dropped rather than wrapped: a kept
// output (W+pad+y, W+pad+x) only meets K at offsets pad+1 .. 2W-pad-1 < L
// in each axis, so no needed term is lost and none aliases.
#include "fft_lds.h"

namespace ath {
namespace {

// Kernels by size class CLS = maxb_for(length) (butterflies per thread of a
// 256-thread workgroup): the thread count and butterflies actually used.
// Lengths 4096 and 8192 run 512 and 1024 threads of two butterflies, for
// more waves per SIMD behind the LDS latency.
template <int CLS> struct F2Cfg { static constexpr int maxb = CLS, nt = 256; };
template <> struct F2Cfg<4> { static constexpr int maxb = 2, nt = 512; };
template <> struct F2Cfg<8> { static constexpr int maxb = 2, nt = 1024; };

constexpr int kSrcImage = 0;   // rows of the padded image
constexpr int kSrcKernel = 1;  // rows of the tiled kernel K

struct Plan2d {
    const double2* tw;    // L entries, exp(-2 pi i n / L)
    const double2* tw2;   // 2L entries, exp(-2 pi i n / 2L)
    const double2* khat;  // (M + 1) x L, column k contiguous
};

__host__ __device__ inline Plan2d plan_view(const void* plan, int L) {
    const double2* base = static_cast<const double2*>(plan);
    return Plan2d{base, base + L, base + 3 * (size_t)L};
}

// Forward real FFT of one row: image rows with numpy.pad on both axes, or
// rows of the tiled kernel.  T row r of image k: T + (k * nrows + r) * (M+1).
template <int MAXB, int SRC>
__global__ __launch_bounds__(F2Cfg<MAXB>::nt) void rows_fwd_kernel(
    const float* __restrict__ images, long long row_stride,
    long long img_stride, const double* __restrict__ g2, int ncol, int pad,
    int pad_mode, int width, const double2* __restrict__ tw,
    double2* __restrict__ T, int M) {
    constexpr int NT_ = F2Cfg<MAXB>::nt;
    extern __shared__ double2 buf[];
    const int tid = threadIdx.x;
    const int u = blockIdx.x;
    const int nrows = gridDim.x;
    const int img = blockIdx.y;
    const float* src_row = nullptr;
    const double* g_row = nullptr;
    if (SRC == kSrcImage) {
        const int su = pad_source(u - pad, ncol, pad_mode);
        if (su >= 0)
            src_row = images + img * img_stride + (long long)su * row_stride;
    } else {
        g_row = g2 + (size_t)(u % width) * width;
    }
    for (int n = tid; n < M; n += NT_) {
        double v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int col = 2 * n + h;
            v[h] = 0.0;
            if (SRC == kSrcImage) {
                if (src_row && col < width) {
                    const int sv = pad_source(col - pad, ncol, pad_mode);
                    if (sv >= 0) v[h] = (double)src_row[sv];
                }
            } else if (col < 2 * width) {
                v[h] = g_row[col < width ? col : col - width];
            }
        }
        buf[n] = make_double2(v[0], v[1]);
    }
    __syncthreads();
    lds_fft<-1, F2Cfg<MAXB>::maxb, false, NT_>(buf, M, tw);
    double2* out = T + ((size_t)img * nrows + u) * (M + 1);
    for (int k = tid; k <= M / 2; k += NT_) {
        const int km = (M - k) & (M - 1);
        double2 xk, xm;
        split_real(buf[k], buf[km], tw[k], xk, xm);
        out[k] = xk;
        out[M - k] = xm;
    }
}

// Column pass.  SPECTRUM: store the forward transform of column k, scaled,
// as Khat[k][.].  Otherwise multiply by Khat[k][.], transform back and keep
// rows r0 .. r0+nout-1 into U[y][k].
template <int MAXB, bool SPECTRUM>
__global__ __launch_bounds__(F2Cfg<MAXB>::nt) void cols_kernel(
    const double2* __restrict__ T, int nrows_in, const double2* __restrict__ tw2,
    double2* __restrict__ khat_out, const double2* __restrict__ khat,
    double2* __restrict__ U, int r0, int nout, int M, double scale) {
    constexpr int NT_ = F2Cfg<MAXB>::nt;
    extern __shared__ double2 buf[];
    const int tid = threadIdx.x;
    const int k = blockIdx.x;
    const int img = blockIdx.y;
    const int L = 2 * M;
    const double2* col = T + (size_t)img * nrows_in * (M + 1) + k;
    for (int i = tid; i < L; i += NT_)
        buf[i] = i < nrows_in ? col[(size_t)i * (M + 1)] : make_double2(0.0, 0.0);
    __syncthreads();
    lds_fft<-1, F2Cfg<MAXB>::maxb, false, NT_>(buf, L, tw2);
    if (SPECTRUM) {
        double2* dst = khat_out + (size_t)k * L;
        for (int i = tid; i < L; i += NT_) dst[i] = cscale(buf[i], scale);
        return;
    }
    const double2* kc = khat + (size_t)k * L;
    for (int i = tid; i < L; i += NT_) buf[i] = cmul(buf[i], kc[i]);
    __syncthreads();
    lds_fft<1, F2Cfg<MAXB>::maxb, false, NT_>(buf, L, tw2);
    double2* dst = U + (size_t)img * nout * (M + 1) + k;
    for (int y = tid; y < nout; y += NT_)
        dst[(size_t)y * (M + 1)] = buf[r0 + y];
}

// Column pass at L = LV * H (16384: conv_big / fft_big_forward; 32768:
// conv_big4 / fft_big4_forward), persistent over the (image, bin column)
// pairs; workgroup b uses scratch slot b (L values).  tw: the plan's table
// exp(-2 pi i n / L), which is the table both forms need.
template <bool SPECTRUM, int LV>
__global__ __launch_bounds__(kBigNT) void cols_big_kernel(
    const double2* __restrict__ T, int nrows_in, const double2* __restrict__ tw,
    double2* __restrict__ khat_out, const double2* __restrict__ khat,
    double2* __restrict__ U, int r0, int nout, int M, double scale,
    int nimg, double2* __restrict__ scratch_all) {
    extern __shared__ double2 buf[];
    double2* scratch = scratch_all + (size_t)blockIdx.x * LV * kBigH;
    const int L = 2 * M;
    const long long ncols = (long long)(M + 1) * nimg;
    for (long long c = blockIdx.x; c < ncols; c += gridDim.x) {
        const int img = (int)(c / (M + 1));  // block-uniform
        const int k = (int)(c - (long long)img * (M + 1));
        const double2* col = T + (size_t)img * nrows_in * (M + 1) + k;
        auto load = [&](int n) -> double2 { return col[(size_t)n * (M + 1)]; };
        if (SPECTRUM) {
            double2* dst = khat_out + (size_t)k * L;
            fft_lv_forward<LV>(
                buf, scratch, nrows_in, tw, load,
                [&](int i, double2 v) { dst[i] = cscale(v, scale); });
        } else {
            conv_lv<LV>(buf, scratch, nrows_in, tw, khat + (size_t)k * L,
                        load);
            double2* dst = U + (size_t)img * nout * (M + 1) + k;
            for (int y = threadIdx.x; y < nout; y += kBigNT)
                dst[(size_t)y * (M + 1)] =
                    conv_lv_out<LV>(buf, scratch, r0 + y);
            __syncthreads();
        }
    }
}

// Scratch slots of the L = 16384 column pass: one per compute unit (its
// 128 KB of LDS leaves room for one workgroup per CU).
inline int big_slots() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess || cus < 1) {
        (void)hipGetLastError();
        return 256;
    }
    return cus;
}

inline size_t big_scratch_bytes(int L) {
    return L == 2 * kBigH || L == 4 * kBigH
               ? (size_t)big_slots() * L * sizeof(double2)
               : 0;
}

struct Finish2d {
    int ncol, c0;             // output columns, first kept sample
    long long out_row_stride, out_img_stride;
    int use_mask, mask_c;
    double mask_r2;
    int scale;                // multiply by pi / (nrow - 1)
    double pi64, denom64;     // float64 output
    float pi32, denom32;      // float32 output
};

// Row inverse: Hermitian merge of U[y][0..M], inverse FFT, crop, finish.
template <int MAXB, typename Tout>
__global__ __launch_bounds__(F2Cfg<MAXB>::nt) void rows_inv_kernel(
    const double2* __restrict__ U, const double2* __restrict__ tw,
    Tout* __restrict__ out, Finish2d f, int M) {
    constexpr int NT_ = F2Cfg<MAXB>::nt;
    extern __shared__ double2 buf[];
    const int tid = threadIdx.x;
    const int y = blockIdx.x;
    const int img = blockIdx.y;
    const double2* Y = U + ((size_t)img * gridDim.x + y) * (M + 1);
    for (int k = tid; k <= M / 2; k += NT_) {
        const int km = (M - k) & (M - 1);
        const double2 wk = tw[k];
        const double2 yk = Y[k];
        const double2 ym = Y[M - k];
        // Zy[k] = (Yk + conj Ym)/2 + i conj(wk) (Yk - conj Ym)/2
        const double2 ek = cscale(cadd(yk, conjg(ym)), 0.5);
        const double2 ok = rot<1>(cmul(conjg(wk),
                                       cscale(csub(yk, conjg(ym)), 0.5)));
        buf[k] = cadd(ek, ok);
        if (k != 0 && k != km) {
            // Zy[M-k] = (Ym + conj Yk)/2 - i wk (Ym - conj Yk)/2
            const double2 em = cscale(cadd(ym, conjg(yk)), 0.5);
            const double2 om = rot<-1>(cmul(wk,
                                            cscale(csub(ym, conjg(yk)), 0.5)));
            buf[km] = cadd(em, om);
        }
    }
    __syncthreads();
    lds_fft<1, F2Cfg<MAXB>::maxb, false, NT_>(buf, M, tw);
    const double* v = reinterpret_cast<const double*>(buf) + f.c0;
    Tout* row = out + img * f.out_img_stride + (long long)y * f.out_row_stride;
    for (int x = tid; x < f.ncol; x += NT_) {
        bool keep = true;
        if (f.use_mask) {
            const long long dx = x - f.mask_c, dy = y - f.mask_c;
            keep = (double)(dx * dx + dy * dy) <= f.mask_r2;
        }
        if constexpr (sizeof(Tout) == 8) {
            // float64 result of the reference's 2D branch: recon * mask,
            // then * pi, then / (nrow - 1), all float64
            double r = v[x] * (keep ? 1.0 : 0.0);
            if (f.scale) r = r * f.pi64 / f.denom64;
            row[x] = r;
        } else {
            // 3D branch: float32 slices, then * mask, * pi, / (nrow - 1)
            float r = (float)v[x] * (keep ? 1.0f : 0.0f);
            if (f.scale) {
                r = r * f.pi32;
                r = r / f.denom32;
            }
            row[x] = r;
        }
    }
}

// mask and pi/(N-1) of output (x, y) in the reference's order and dtype
template <typename Tout>
__device__ __forceinline__ Tout finish_value(double v, int x, int y,
                                             const Finish2d& f) {
    bool keep = true;
    if (f.use_mask) {
        const long long dx = x - f.mask_c, dy = y - f.mask_c;
        keep = (double)(dx * dx + dy * dy) <= f.mask_r2;
    }
    if constexpr (sizeof(Tout) == 8) {
        double r = v * (keep ? 1.0 : 0.0);
        if (f.scale) r = r * f.pi64 / f.denom64;
        return r;
    } else {
        float r = (float)v * (keep ? 1.0f : 0.0f);
        if (f.scale) {
            r = r * f.pi32;
            r = r / f.denom32;
        }
        return r;
    }
}

// ---- L = 32768: rows as two-half transforms (fft_lds.h) -------------------
// Forward real FFT of one row as rows_fwd_kernel, the packed 16384 points
// as two 8192-point halves.  tw: exp(-2 pi i n / L).
template <int SRC>
__global__ __launch_bounds__(kBigNT) void rows_fwd_big_kernel(
    const float* __restrict__ images, long long row_stride,
    long long img_stride, const double* __restrict__ g2, int ncol, int pad,
    int pad_mode, int width, const double2* __restrict__ tw,
    double2* __restrict__ T) {
    extern __shared__ double2 buf[];
    constexpr int M = 2 * kBigH;
    const int u = blockIdx.x, nrows = gridDim.x, img = blockIdx.y;
    const float* src_row = nullptr;
    const double* g_row = nullptr;
    if (SRC == kSrcImage) {
        const int su = pad_source(u - pad, ncol, pad_mode);
        if (su >= 0)
            src_row = images + img * img_stride + (long long)su * row_stride;
    } else {
        g_row = g2 + (size_t)(u % width) * width;
    }
    auto val = [&](int col) -> double {
        if (SRC == kSrcImage) {
            if (src_row && col < width) {
                const int sv = pad_source(col - pad, ncol, pad_mode);
                if (sv >= 0) return (double)src_row[sv];
            }
            return 0.0;
        }
        return col < 2 * width ? g_row[col < width ? col : col - width] : 0.0;
    };
    BigRegs r;
    big_forward_halves(buf, tw, val, r);
    double2* out = T + ((size_t)img * nrows + u) * (M + 1);
    big_pairs(buf, tw, r, [&](int k, double2 zk, double2 zm) {
        double2 xk, xm;
        split_real(zk, zm, tw[k], xk, xm);
        out[k] = xk;
        out[M - k] = xm;
    });
}

// Row inverse as rows_inv_kernel: the Hermitian merge of U[y][0..M] pair by
// pair, the two-half inverse, crop, finish.
template <typename Tout>
__global__ __launch_bounds__(kBigNT) void rows_inv_big_kernel(
    const double2* __restrict__ U, const double2* __restrict__ tw,
    Tout* __restrict__ out, Finish2d f) {
    extern __shared__ double2 buf[];
    constexpr int H = kBigH, M = 2 * H, NT = kBigNT;
    const int tid = threadIdx.x;
    const int y = blockIdx.x;
    const int img = blockIdx.y;
    const double2* Y = U + ((size_t)img * gridDim.x + y) * (M + 1);
    BigRegs r;
#pragma unroll
    for (int b = 0; b < kBigKB; ++b) {
        const int k = tid + NT * b;
        double2 z0, z1, z2, z3;  // the inverse's inputs at k, H+k, H-k, 2H-k
        if (k == 0) {
            // packed bin 0 takes the DC and Nyquist values; bin H pairs
            // with itself
            merge_pair(Y[0], Y[M], 0, M, tw, z0, z0);
            merge_pair(Y[H], Y[H], H, M, tw, z1, z1);
            z2 = z0;
            z3 = z1;
        } else {
            merge_pair(Y[k], Y[M - k], k, M, tw, z0, z3);
            merge_pair(Y[H - k], Y[H + k], H - k, M, tw, z2, z1);
        }
        big_store_uv(buf, tw, k, z0, z1, z2, z3, r.v[b]);
    }
    if (tid == 0) {
        double2 z0, z1;
        merge_pair(Y[H / 2], Y[M - H / 2], H / 2, M, tw, z0, z1);
        big_store_mid(buf, tw, z0, z1, r.mid);
    }
    __syncthreads();
    const double* v = reinterpret_cast<const double*>(buf);
    Tout* row = out + img * f.out_img_stride + (long long)y * f.out_row_stride;
    big_inverse_phases(buf, tw, r, [&](int h) {
        for (int x = tid; x < f.ncol; x += NT) {
            const int q = f.c0 + x;
            if (((q >> 1) & 1) == h)
                row[x] = finish_value<Tout>(v[2 * (q >> 2) + (q & 1)], x, y,
                                            f);
        }
    });
}

template <int MAXB_R, int MAXB_C>
int build_plan(const double* g2, int width, int L, void* plan, void* work,
               hipStream_t s) {
    const int M = L / 2;
    double2* tw = static_cast<double2*>(plan);
    double2* tw2 = tw + L;
    double2* khat = tw + 3 * (size_t)L;
    hipLaunchKernelGGL(twiddle_kernel, dim3((L + 255) / 256), dim3(256), 0, s,
                       tw, L);
    hipLaunchKernelGGL(twiddle_kernel, dim3((2 * L + 255) / 256), dim3(256),
                       0, s, tw2, 2 * L);
    int st = check_hip(hipGetLastError(), "twiddle_kernel launch");
    if (st != ATH_OK) return st;
    auto* T = static_cast<double2*>(work);
    // rows of K at or beyond L never reach a kept output (see the header)
    const int krows = std::min(2 * width, L);
    auto rk = rows_fwd_kernel<MAXB_R, kSrcKernel>;
    const size_t lds_r = (size_t)M * sizeof(double2);
    prepare_lds(rk, 1024 * (size_t)MAXB_R * sizeof(double2));
    hipLaunchKernelGGL(rk, dim3(krows, 1), dim3(F2Cfg<MAXB_R>::nt), lds_r, s,
                       nullptr, 0LL, 0LL, g2, width, 0, ATH_PAD_EDGE, width,
                       tw, T, M);
    st = check_hip(hipGetLastError(), "rows_fwd_kernel (kernel) launch");
    if (st != ATH_OK) return st;
    auto ck = cols_kernel<MAXB_C, true>;
    const size_t lds_c = (size_t)L * sizeof(double2);
    prepare_lds(ck, 1024 * (size_t)MAXB_C * sizeof(double2));
    hipLaunchKernelGGL(ck, dim3(M + 1, 1), dim3(F2Cfg<MAXB_C>::nt), lds_c, s, T,
                       krows, tw2, khat, nullptr, nullptr, 0, 0, M,
                       1.0 / ((double)M * (double)L));
    return check_hip(hipGetLastError(), "cols_kernel (spectrum) launch");
}

// L = 16384: the row passes as for L <= 8192 (class 8: 1024 threads), the
// column pass two-level.  L = 32768: the row passes as two-half
// transforms, the column pass radix-4 two-level.  work: the row-spectrum
// rows, then the scratch.
template <int LV>
int build_plan_big(const double* g2, int width, int L, void* plan,
                   void* work, hipStream_t s) {
    const int M = L / 2;
    double2* tw = static_cast<double2*>(plan);
    double2* tw2 = tw + L;
    double2* khat = tw + 3 * (size_t)L;
    hipLaunchKernelGGL(twiddle_kernel, dim3((L + 255) / 256), dim3(256), 0, s,
                       tw, L);
    hipLaunchKernelGGL(twiddle_kernel, dim3((2 * L + 255) / 256), dim3(256),
                       0, s, tw2, 2 * L);
    int st = check_hip(hipGetLastError(), "twiddle_kernel launch");
    if (st != ATH_OK) return st;
    auto* T = static_cast<double2*>(work);
    const int krows = std::min(2 * width, L);
    auto* scr = T + (size_t)krows * (M + 1);
    const size_t lds_r = (size_t)(LV == 4 ? kBigH : M) * sizeof(double2);
    if constexpr (LV == 4) {
        auto rk = rows_fwd_big_kernel<kSrcKernel>;
        prepare_lds(rk, lds_r);
        hipLaunchKernelGGL(rk, dim3(krows, 1), dim3(kBigNT), lds_r, s,
                           nullptr, 0LL, 0LL, g2, width, 0, ATH_PAD_EDGE,
                           width, tw, T);
    } else {
        auto rk = rows_fwd_kernel<8, kSrcKernel>;
        prepare_lds(rk, 1024 * (size_t)8 * sizeof(double2));
        hipLaunchKernelGGL(rk, dim3(krows, 1), dim3(F2Cfg<8>::nt), lds_r, s,
                           nullptr, 0LL, 0LL, g2, width, 0, ATH_PAD_EDGE,
                           width, tw, T, M);
    }
    st = check_hip(hipGetLastError(), "rows_fwd_kernel (kernel) launch");
    if (st != ATH_OK) return st;
    auto ck = cols_big_kernel<true, LV>;
    const size_t lds_c = (size_t)kBigH * sizeof(double2);
    prepare_lds(ck, lds_c);
    const int grid = std::min(M + 1, big_slots());
    hipLaunchKernelGGL(ck, dim3(grid), dim3(kBigNT), lds_c, s, T, krows, tw,
                       khat, nullptr, nullptr, 0, 0, M,
                       1.0 / ((double)M * (double)L), 1, scr);
    return check_hip(hipGetLastError(), "cols_big_kernel (spectrum) launch");
}

template <int MAXB_R, int MAXB_C, typename Tout>
int run_filter(const float* images, int nimg, int ncol, long long row_stride,
               long long img_stride, int pad, int pad_mode, const void* plan,
               int L, void* work, Tout* out, const Finish2d& f,
               hipStream_t s) {
    const int M = L / 2;
    const int width = ncol + 2 * pad;
    const Plan2d p = plan_view(plan, L);
    auto* T = static_cast<double2*>(work);
    auto* U = T + (size_t)nimg * width * (M + 1);
    auto rk = rows_fwd_kernel<MAXB_R, kSrcImage>;
    const size_t lds_r = (size_t)M * sizeof(double2);
    prepare_lds(rk, 1024 * (size_t)MAXB_R * sizeof(double2));
    hipLaunchKernelGGL(rk, dim3(width, nimg), dim3(F2Cfg<MAXB_R>::nt), lds_r, s,
                       images,
                       row_stride, img_stride, nullptr, ncol, pad, pad_mode,
                       width, p.tw, T, M);
    int st = check_hip(hipGetLastError(), "rows_fwd_kernel launch");
    if (st != ATH_OK) return st;
    auto ck = cols_kernel<MAXB_C, false>;
    const size_t lds_c = (size_t)L * sizeof(double2);
    prepare_lds(ck, 1024 * (size_t)MAXB_C * sizeof(double2));
    hipLaunchKernelGGL(ck, dim3(M + 1, nimg), dim3(F2Cfg<MAXB_C>::nt), lds_c, s,
                       T,
                       width, p.tw2, nullptr, p.khat, U, width + pad, ncol, M,
                       1.0);
    st = check_hip(hipGetLastError(), "cols_kernel launch");
    if (st != ATH_OK) return st;
    auto ik = rows_inv_kernel<MAXB_R, Tout>;
    prepare_lds(ik, 1024 * (size_t)MAXB_R * sizeof(double2));
    hipLaunchKernelGGL(ik, dim3(ncol, nimg), dim3(F2Cfg<MAXB_R>::nt), lds_r, s,
                       U,
                       p.tw, out, f, M);
    return check_hip(hipGetLastError(), "rows_inv_kernel launch");
}

template <int LV, typename Tout>
int run_filter_big(const float* images, int nimg, int ncol,
                   long long row_stride, long long img_stride, int pad,
                   int pad_mode, const void* plan, int L, void* work,
                   Tout* out, const Finish2d& f, hipStream_t s) {
    const int M = L / 2;
    const int width = ncol + 2 * pad;
    const Plan2d p = plan_view(plan, L);
    auto* T = static_cast<double2*>(work);
    auto* U = T + (size_t)nimg * width * (M + 1);
    auto* scr = U + (size_t)nimg * ncol * (M + 1);
    const size_t lds_r = (size_t)(LV == 4 ? kBigH : M) * sizeof(double2);
    if constexpr (LV == 4) {
        auto rk = rows_fwd_big_kernel<kSrcImage>;
        prepare_lds(rk, lds_r);
        hipLaunchKernelGGL(rk, dim3(width, nimg), dim3(kBigNT), lds_r, s,
                           images, row_stride, img_stride, nullptr, ncol, pad,
                           pad_mode, width, p.tw, T);
    } else {
        auto rk = rows_fwd_kernel<8, kSrcImage>;
        prepare_lds(rk, 1024 * (size_t)8 * sizeof(double2));
        hipLaunchKernelGGL(rk, dim3(width, nimg), dim3(F2Cfg<8>::nt), lds_r,
                           s, images, row_stride, img_stride, nullptr, ncol,
                           pad, pad_mode, width, p.tw, T, M);
    }
    int st = check_hip(hipGetLastError(), "rows_fwd_kernel launch");
    if (st != ATH_OK) return st;
    auto ck = cols_big_kernel<false, LV>;
    const size_t lds_c = (size_t)kBigH * sizeof(double2);
    prepare_lds(ck, lds_c);
    const long long ncols = (long long)(M + 1) * nimg;
    const int grid = (int)std::min<long long>(ncols, big_slots());
    hipLaunchKernelGGL(ck, dim3(grid), dim3(kBigNT), lds_c, s, T, width, p.tw,
                       nullptr, p.khat, U, width + pad, ncol, M, 1.0, nimg,
                       scr);
    st = check_hip(hipGetLastError(), "cols_big_kernel launch");
    if (st != ATH_OK) return st;
    if constexpr (LV == 4) {
        auto ik = rows_inv_big_kernel<Tout>;
        prepare_lds(ik, lds_r);
        hipLaunchKernelGGL(ik, dim3(ncol, nimg), dim3(kBigNT), lds_r, s, U,
                           p.tw, out, f);
    } else {
        auto ik = rows_inv_kernel<8, Tout>;
        prepare_lds(ik, 1024 * (size_t)8 * sizeof(double2));
        hipLaunchKernelGGL(ik, dim3(ncol, nimg), dim3(F2Cfg<8>::nt), lds_r, s,
                           U, p.tw, out, f, M);
    }
    return check_hip(hipGetLastError(), "rows_inv_kernel launch");
}

}  // namespace

size_t circular_plan_bytes(int L) {
    const size_t M = (size_t)L / 2;
    return (3 * (size_t)L + (M + 1) * (size_t)L) * sizeof(double2);
}

size_t circular_plan_work_bytes(int width, int L) {
    return 2 * (size_t)width * ((size_t)L / 2 + 1) * sizeof(double2) +
           big_scratch_bytes(L);
}

size_t circular_filter_work_bytes(int ncol, int pad, int L, int nimg) {
    const size_t width = (size_t)ncol + 2 * (size_t)pad;
    return (size_t)nimg * (width + (size_t)ncol) * ((size_t)L / 2 + 1) *
               sizeof(double2) +
           big_scratch_bytes(L);
}

int circular_plan_build(const double* g2, int width, int L, void* plan,
                        void* work, hipStream_t s) {
    switch (L) {
    case 64: case 128: case 256: case 512: case 1024:
        return build_plan<1, 1>(g2, width, L, plan, work, s);
    case 2048: return build_plan<1, 2>(g2, width, L, plan, work, s);
    case 4096: return build_plan<2, 4>(g2, width, L, plan, work, s);
    case 8192: return build_plan<4, 8>(g2, width, L, plan, work, s);
    case 16384: return build_plan_big<2>(g2, width, L, plan, work, s);
    case 32768: return build_plan_big<4>(g2, width, L, plan, work, s);
    }
    return set_error(ATH_ERR_UNSUPPORTED,
                     "circular ramp filter: FFT length above 32768");
}

template <typename Tout>
static int dispatch_filter(const float* images, int nimg, int ncol,
                           long long row_stride, long long img_stride,
                           int pad, int pad_mode, const void* plan, int L,
                           void* work, Tout* out, const Finish2d& f,
                           hipStream_t s) {
#define ATH_F2D(R_, C_)                                                      \
    return run_filter<R_, C_, Tout>(images, nimg, ncol, row_stride,          \
                                    img_stride, pad, pad_mode, plan, L, work, \
                                    out, f, s);
    switch (L) {
    case 64: case 128: case 256: case 512: case 1024: ATH_F2D(1, 1)
    case 2048: ATH_F2D(1, 2)
    case 4096: ATH_F2D(2, 4)
    case 8192: ATH_F2D(4, 8)
    case 16384:
        return run_filter_big<2, Tout>(images, nimg, ncol, row_stride,
                                       img_stride, pad, pad_mode, plan, L,
                                       work, out, f, s);
    case 32768:
        return run_filter_big<4, Tout>(images, nimg, ncol, row_stride,
                                       img_stride, pad, pad_mode, plan, L,
                                       work, out, f, s);
    }
#undef ATH_F2D
    return set_error(ATH_ERR_UNSUPPORTED,
                     "circular ramp filter: FFT length above 32768");
}

int circular_ramp_filter(const float* images, int nimg, int ncol,
                         long long row_stride, long long img_stride, int pad,
                         int pad_mode, const void* plan, int L, void* work,
                         void* out, int out_dtype, long long out_row_stride,
                         long long out_img_stride, double mask_radius,
                         int scale_nrow, hipStream_t s) {
    Finish2d f;
    f.ncol = ncol;
    f.c0 = ncol + 3 * pad;  // W + pad
    f.out_row_stride = out_row_stride;
    f.out_img_stride = out_img_stride;
    f.use_mask = mask_radius >= 0.0;
    f.mask_c = ncol / 2;
    f.mask_r2 = mask_radius * mask_radius;
    f.scale = scale_nrow >= 2;
    f.pi64 = 3.141592653589793;
    f.denom64 = (double)(scale_nrow - 1);
    f.pi32 = (float)3.141592653589793;
    f.denom32 = (float)(scale_nrow - 1);
    if (out_dtype == ATH_F64)
        return dispatch_filter<double>(images, nimg, ncol, row_stride,
                                       img_stride, pad, pad_mode, plan, L,
                                       work, static_cast<double*>(out), f, s);
    return dispatch_filter<float>(images, nimg, ncol, row_stride, img_stride,
                                  pad, pad_mode, plan, L, work,
                                  static_cast<float*>(out), f, s);
}

}  // namespace ath
