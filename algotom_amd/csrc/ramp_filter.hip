// Ramp filter stage of FBP on gfx950.
//
// algotom's apply_ramp_filter (reconstruction.py:118-153) pads every row to
// W = ncol + 2*pad columns, multiplies its shifted spectrum by the ramp
// window and keeps the real part of the shifted inverse.  That is exactly a
// length-W circular convolution of the padded row with a real kernel g
// (g = Re(ifftshift(ifft(ifftshift(R)))); W is mostly 4 x prime here, so no
// fast length-W transform exists).  This kernel evaluates it exactly as a
// LINEAR convolution with the doubled kernel [g, g] on a power-of-two
// length L >= 2W - pad, where the outputs W+pad .. W+pad+ncol-1 are free of
// wrap-around.  Everything is float64: one workgroup owns one row, keeps the
// row as L/2 packed complex samples (x[2n] + i x[2n+1]) in LDS, runs a
// radix-4 Stockham FFT, multiplies by the kernel's half spectrum (the real
// FFT split and the inverse merge fused in one pass) and runs the inverse.
// The pass is HBM-bound: one read and one write of every sample.
#include "fft_lds.h"
#include "log32.h"

namespace ath {
namespace {

constexpr int kThreads = 256;

// Row kernels by size class CLS = maxb_for(M) (the butterflies a 256-thread
// workgroup would give each thread): the workgroup's thread count and the
// butterflies per thread actually used.  Larger transforms get more threads
// (more waves per SIMD to hide LDS latency) at two butterflies each:
// M = 4096 as 512 threads (two workgroups of 80 KB per CU), M = 8192 as 1024
// threads (one workgroup of 128 KB per CU).
template <int CLS> struct RowCfg { static constexpr int maxb = CLS, nt = 256; };
template <> struct RowCfg<4> { static constexpr int maxb = 2, nt = 512; };
template <> struct RowCfg<8> { static constexpr int maxb = 2, nt = 1024; };

template <typename Tin>
__device__ __forceinline__ double load_sample(const Tin* row, int src,
                                              int apply_log);
template <>
__device__ __forceinline__ double load_sample<float>(const float* row, int src,
                                                     int apply_log) {
    const float x = row[src];
    // -np.log of a float32 array is float32 (reconstruction.py:436)
    return apply_log ? (double)neg_log_f32(x) : (double)x;
}
template <>
__device__ __forceinline__ double load_sample<double>(const double* row,
                                                      int src, int apply_log) {
    const double x = row[src];
    return apply_log ? -log(x) : x;
}

// ---- numpy.pad's value modes (ATH_PAD_LINEAR_RAMP .. ATH_PAD_MINIMUM) ----
// With numpy's defaults (end_values 0, stat_length None) every row needs
// two numbers, computed by the row's workgroup before it fills the padded
// row: the pad value of the statistic modes (the same on both sides), or
// the left / right steps edge / pad of linear_ramp (float64).
struct PadVals {
    double lo, hi;
};

template <typename T> struct KeyOf;
template <> struct KeyOf<float> { using K = unsigned int; };
template <> struct KeyOf<double> { using K = unsigned long long; };

// order-preserving map between floating-point values and unsigned keys
template <typename T, typename K>
__device__ __forceinline__ K to_key(T v) {
    constexpr K sign = (K)1 << (sizeof(K) * 8 - 1);
    const K u = __builtin_bit_cast(K, v);
    return (u & sign) ? (K)~u : (K)(u | sign);
}
template <typename T, typename K>
__device__ __forceinline__ T from_key(K k) {
    constexpr K sign = (K)1 << (sizeof(K) * 8 - 1);
    return __builtin_bit_cast(T, (k & sign) ? (K)(k ^ sign) : (K)~k);
}

// The k-th smallest (0-based) of n keys in LDS: radix select, 4 bits per
// pass from the top, `hist` 16 ints of LDS.  All NT threads call it; it
// ends synchronised.  Invariant: more than k keys match the prefix.
template <int NT, typename K>
__device__ K lds_select(const K* keys, int n, int k, int* hist) {
    const int tid = threadIdx.x;
    K prefix = 0, mask = 0;
    for (int shift = (int)sizeof(K) * 8 - 4; shift >= 0; shift -= 4) {
        if (tid < 16) hist[tid] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += NT) {
            const K key = keys[i];
            if ((key & mask) == prefix)
                atomicAdd(&hist[(int)((key >> shift) & 15)], 1);
        }
        __syncthreads();
        int d = 0, below = 0;
        while (below + hist[d] <= k) below += hist[d++];
        k -= below;
        prefix |= (K)d << shift;
        mask |= (K)15 << shift;
        __syncthreads();  // every thread has read hist before the reset
    }
    return prefix;
}

template <int NT>
__device__ double block_reduce(double v, int mode, double* red) {
    auto comb = [mode](double a, double b) {
        return mode == ATH_PAD_MAXIMUM ? fmax(a, b)
               : mode == ATH_PAD_MINIMUM ? fmin(a, b) : a + b;
    };
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = comb(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < NT / 64; ++w) r = comb(r, red[w]);
    __syncthreads();  // red free for the caller
    return r;
}

// PadVals of one row (sample c = the row's value after the optional log).
// buf: the row's LDS (M >= ncol + 1 double2, M >= 32) used as scratch; all
// NT threads; ends synchronised.
template <int NT, typename Tin>
__device__ PadVals row_pad_values(double2* buf, const Tin* __restrict__ row,
                                  int ncol, int pad, int mode,
                                  int apply_log) {
    const int tid = threadIdx.x;
    PadVals pv{0.0, 0.0};
    if (mode == ATH_PAD_LINEAR_RAMP) {
        // numpy.linspace(0, edge, pad, endpoint=False, dtype=row type):
        // value j = j * (edge / pad) evaluated in float64, then rounded to
        // the row's type
        pv.lo = load_sample<Tin>(row, 0, apply_log) / (double)pad;
        pv.hi = load_sample<Tin>(row, ncol - 1, apply_log) / (double)pad;
        return pv;
    }
    if (mode == ATH_PAD_MEDIAN) {
        using K = typename KeyOf<Tin>::K;
        K* keys = reinterpret_cast<K*>(buf);
        int* hist = reinterpret_cast<int*>(keys + ncol + (ncol & 1));
        for (int c = tid; c < ncol; c += NT)
            keys[c] = to_key<Tin, K>((Tin)load_sample<Tin>(row, c, apply_log));
        __syncthreads();
        // numpy.median: the middle value, or the mean of the two middle
        // values in the row's type
        const int k2 = ncol / 2;
        const Tin hi = from_key<Tin, K>(lds_select<NT, K>(keys, ncol, k2,
                                                          hist));
        Tin med = hi;
        if ((ncol & 1) == 0) {
            const Tin lo = from_key<Tin, K>(lds_select<NT, K>(keys, ncol,
                                                              k2 - 1, hist));
            med = (Tin)(lo + hi) / (Tin)2;
        }
        pv.lo = pv.hi = (double)med;
        return pv;
    }
    double acc = mode == ATH_PAD_MAXIMUM   ? -HUGE_VAL
                 : mode == ATH_PAD_MINIMUM ? HUGE_VAL : 0.0;
    for (int c = tid; c < ncol; c += NT) {
        const double v = load_sample<Tin>(row, c, apply_log);
        acc = mode == ATH_PAD_MAXIMUM   ? fmax(acc, v)
              : mode == ATH_PAD_MINIMUM ? fmin(acc, v) : acc + v;
    }
    double r = block_reduce<NT>(acc, mode, reinterpret_cast<double*>(buf));
    if (mode == ATH_PAD_MEAN) r = (double)((Tin)r / (Tin)ncol);
    pv.lo = pv.hi = r;
    return pv;
}

// value of padded sample k (outside [pad, pad + ncol)) in a value mode
template <typename Tin>
__device__ __forceinline__ double pad_value(const PadVals& pv, int mode,
                                            int k, int pad, int ncol) {
    if (mode != ATH_PAD_LINEAR_RAMP) return pv.lo;
    // left ramp j = k; right ramp reversed: j = pad - 1 - (k - pad - ncol)
    const bool left = k < pad;
    const int j = left ? k : 2 * pad + ncol - 1 - k;
    return (double)(Tin)((double)j * (left ? pv.lo : pv.hi));
}

// Sample k of the zero-extended padded row (0 beyond the padded width W):
// the pad index map, the optional log, and numpy.pad's value modes.  ADJ
// embeds the row with zeros instead (the transpose of the padding).
template <typename Tin, bool ADJ>
__device__ __forceinline__ double padded_sample(const Tin* __restrict__ row,
                                                int k, int width, int ncol,
                                                int pad, int pad_mode,
                                                int apply_log, bool value_mode,
                                                const PadVals& pv) {
    if (k >= width) return 0.0;
    const int s = ADJ ? pad_source(k - pad, ncol, ATH_PAD_CONSTANT)
                      : pad_source(k - pad, ncol, pad_mode);
    if (s >= 0) return load_sample<Tin>(row, s, apply_log);
    if (value_mode) return pad_value<Tin>(pv, pad_mode, k, pad, ncol);
    return 0.0;
}

// The kernel product in the packed domain for the index pair (k, km = (M -
// k) mod M), 0 <= k <= M/2, of Z = FFT_M(p[2n] + i p[2n+1]): the real-FFT
// split, the products with the kernel's spectrum G (M + 1 values) and the
// merge for the inverse.  zk, zm: Z[k], Z[km] in, the inverse's input at k
// and km out (zm untouched when k == km).  tw: the length-2M table.
__device__ __forceinline__ void spectral_pair(double2& zk, double2& zm, int k,
                                              int M,
                                              const double2* __restrict__ tw,
                                              const double2* __restrict__ G) {
    const int km = (M - k) & (M - 1);
    double2 xk, xm;
    split_real(zk, k == km ? zk : zm, tw[k], xk, xm);
    merge_pair(cmul(xk, G[k]), cmul(xm, G[M - k]), k, M, tw, zk, zm);
}

// Filters one row in LDS buffer `buf` (M double2): pad, rfft, multiply by
// the kernel half spectrum G, irfft, crop into dst.  All 256 threads.
// ADJ: the transpose of the filter -- the row is embedded with zeros
// (pad_mode is then the FORWARD pass's mode, used only to fold), the kernel
// spectrum is that of the reversed taps, and every one of the W outputs of
// the circular correlation is folded back through the transpose of the pad
// index map instead of cropped.
template <int MAXB, typename Tin, typename Tout, bool ADJ = false>
__device__ __forceinline__ void filter_row(
    double2* buf, const double2* tq, const Tin* __restrict__ src_row,
    Tout* __restrict__ dst, int ncol, int pad, int pad_mode, int apply_log,
    const double2* __restrict__ tw, const double2* __restrict__ G, int M) {
    constexpr bool QT = MAXB <= 4;  // quarter table in LDS (M <= 4096)
    constexpr int B = RowCfg<MAXB>::maxb, NT = RowCfg<MAXB>::nt;
    const int tid = threadIdx.x;
    const int width = ncol + 2 * pad;
    PadVals pv{0.0, 0.0};
    bool value_mode = false;
    if constexpr (!ADJ) {
        value_mode = pad_mode >= ATH_PAD_LINEAR_RAMP && pad > 0;
        if (value_mode)
            pv = row_pad_values<NT, Tin>(buf, src_row, ncol, pad, pad_mode,
                                         apply_log);
    }
    // packed, padded, zero-extended row: buf[n] = (p[2n], p[2n+1])
    for (int n = tid; n < M; n += NT)
        buf[n] = make_double2(
            padded_sample<Tin, ADJ>(src_row, 2 * n, width, ncol, pad,
                                    pad_mode, apply_log, value_mode, pv),
            padded_sample<Tin, ADJ>(src_row, 2 * n + 1, width, ncol, pad,
                                    pad_mode, apply_log, value_mode, pv));
    __syncthreads();
    lds_fft<-1, B, QT, NT, 1, true>(buf, M, tw, tq);

    // rfft split, multiply by the kernel spectrum, merge for the inverse
    for (int k = tid; k <= M / 2; k += NT) {
        const int km = (M - k) & (M - 1);
        double2 zk = buf[k], zm = buf[km];
        spectral_pair(zk, zm, k, M, tw, G);
        buf[k] = zk;
        if (km != k) buf[km] = zm;
    }
    __syncthreads();
    lds_fft<1, B, QT, NT, 1, true>(buf, M, tw, tq);

    const double* y = reinterpret_cast<const double*>(buf);
    if (!ADJ) {
        const int off = width + pad;
        for (int c = tid; c < ncol; c += NT)
            store_elem(dst + c, y[off + c]);
    } else {
        // z[n] = y[W + n], n in [0, W); out[c] = sum of z[n] over the n whose
        // pad source is c (the interior n = pad + c, plus folded pad samples)
        const double* z = y + width;
        for (int c = tid; c < ncol; c += NT) {
            double v = z[pad + c];
            if (pad_mode == ATH_PAD_EDGE) {
                if (c == 0)
                    for (int n = 0; n < pad; ++n) v += z[n];
                if (c == ncol - 1)
                    for (int n = pad + ncol; n < width; ++n) v += z[n];
            } else if (pad_mode != ATH_PAD_CONSTANT) {
                for (int n = 0; n < pad; ++n)
                    if (pad_source(n - pad, ncol, pad_mode) == c) v += z[n];
                for (int n = pad + ncol; n < width; ++n)
                    if (pad_source(n - pad, ncol, pad_mode) == c) v += z[n];
            }
            store_elem(dst + c, v);
        }
    }
    __syncthreads();  // buf free for the caller's next row
}

// LDS of the row kernels: the row (M complex) then, for M <= 4096, the
// quarter twiddle table (M/4 complex).
template <int MAXB>
constexpr size_t row_lds_bytes(int M) {
    return (size_t)(M + (MAXB <= 4 ? M / 4 : 0)) * sizeof(double2);
}

// Persistent: each workgroup stages the twiddles once, then filters rows
// blockIdx.x, blockIdx.x + gridDim.x, ... of the (nrow, nslice, ncol) stack.
template <int MAXB, typename Tin, typename Tout>
__global__ __launch_bounds__(RowCfg<MAXB>::nt) void ramp_filter_kernel(
    const Tin* __restrict__ sino, Tout* __restrict__ out, size_t rows,
    int nslice, int ncol, int pad, int pad_mode, int apply_log,
    const double2* __restrict__ tw, const double2* __restrict__ spec,
    int nwin, int M) {
    extern __shared__ double2 buf[];
    double2* tq = buf + M;
    if (MAXB <= 4)  // read after the row's barrier
        stage_quarter<RowCfg<MAXB>::nt>(tq, tw, M);
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const double2* G =
            spec + (nwin > 1 ? (row / nslice) : 0) * (size_t)(M + 1);
        filter_row<MAXB, Tin, Tout>(buf, tq, sino + row * (size_t)ncol,
                                    out + row * (size_t)ncol, ncol, pad,
                                    pad_mode, apply_log, tw, G, M);
    }
}

// Transpose of ramp_filter_kernel (float32 in and out), persistent.
template <int MAXB>
__global__ __launch_bounds__(RowCfg<MAXB>::nt) void ramp_filter_adjoint_kernel(
    const float* __restrict__ in, float* __restrict__ out, size_t rows,
    int nslice, int ncol, int pad, int pad_mode,
    const double2* __restrict__ tw, const double2* __restrict__ spec,
    int nwin, int M) {
    extern __shared__ double2 buf[];
    double2* tq = buf + M;
    if (MAXB <= 4) stage_quarter<RowCfg<MAXB>::nt>(tq, tw, M);
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const double2* G =
            spec + (nwin > 1 ? (row / nslice) : 0) * (size_t)(M + 1);
        filter_row<MAXB, float, float, true>(buf, tq, in + row * (size_t)ncol,
                                             out + row * (size_t)ncol, ncol,
                                             pad, pad_mode, 0, tw, G, M);
    }
}

// Rows of many stacks sharing the FFT length L = 2M (hence the twiddles),
// persistent; a row's stack is found by binary search over row_begin.
template <int MAXB, typename Tin, typename Tout>
__global__ __launch_bounds__(RowCfg<MAXB>::nt) void ramp_filter_batched_kernel(
    const ath_filter_job* __restrict__ jobs, int njobs, long long total_rows,
    int pad_mode, int apply_log, int M) {
    extern __shared__ double2 buf[];
    double2* tq = buf + M;
    if (MAXB <= 4)
        stage_quarter<RowCfg<MAXB>::nt>(
            tq, static_cast<const double2*>(jobs[0].plan), M);
    for (long long grow = blockIdx.x; grow < total_rows; grow += gridDim.x) {
        int lo = 0, hi = njobs - 1;  // last job with row_begin <= grow
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].row_begin <= grow) lo = mid;
            else hi = mid - 1;
        }
        const ath_filter_job jb = jobs[lo];
        const size_t row = (size_t)(grow - jb.row_begin);
        const double2* tw = static_cast<const double2*>(jb.plan);
        const double2* G =
            tw + 2 * M +
            (jb.nwin > 1 ? (row / jb.nslice) : 0) * (size_t)(M + 1);
        filter_row<MAXB, Tin, Tout>(
            buf, tq, static_cast<const Tin*>(jb.sino) + row * (size_t)jb.ncol,
            static_cast<Tout*>(jb.out) + row * (size_t)jb.ncol, jb.ncol,
            jb.pad, pad_mode, apply_log, tw, G, M);
    }
}

// spec[r][k] = rfft_L([g_r, g_r, 0...])[k] / M, k = 0..M, one block per row.
template <int MAXB>
__global__ __launch_bounds__(kThreads) void kernel_spectrum_kernel(
    const double* __restrict__ taps, int width, const double2* __restrict__ tw,
    double2* __restrict__ spec, int M) {
    extern __shared__ double2 buf[];
    const int tid = threadIdx.x;
    const double* g = taps + (size_t)blockIdx.x * width;
    double2* G = spec + (size_t)blockIdx.x * (M + 1);
    for (int n = tid; n < M; n += kThreads) {
        double v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = 2 * n + h;
            v[h] = k < 2 * width ? g[k < width ? k : k - width] : 0.0;
        }
        buf[n] = make_double2(v[0], v[1]);
    }
    __syncthreads();
    lds_fft<-1, MAXB>(buf, M, tw);
    const double inv = 1.0 / (double)M;
    for (int k = tid; k <= M / 2; k += kThreads) {
        const int km = (M - k) & (M - 1);
        double2 xk, xm;
        split_real(buf[k], buf[km], tw[k], xk, xm);
        G[k] = cscale(xk, inv);
        G[M - k] = cscale(xm, inv);
    }
}

// ---- FFT length L = 32768: M = 16384 = 2H packed points --------------------
// One workgroup of kBigNT threads per row through the two-half transforms of
// fft_lds.h (BigRegs, big_forward_halves, big_z, big_store_uv,
// big_inverse_phases): H = 8192 complex float64 in LDS and no global
// scratch.
// the row's LDS plus room for the median's histogram past ncol keys
constexpr size_t kBigRowLds = (size_t)kBigH * sizeof(double2) + 256;
constexpr int kBigRowNT = kBigNT;

// ADJ (the filter's transpose, as in filter_row): the W outputs
// z[n] = y[W + n] fold back through the transpose of the pad index map.
// Outputs c that receive folded pad samples (c <= pad or c >= ncol - pad - 1
// for the index-map modes) need values of both inverse phases, so the odd
// phase's values they use (their own z[pad + c] and the pad samples) wait
// in `stash` (kAdjStash doubles of LDS past the row); every other output is
// stored by the phase that computes it.
constexpr int kAdjMaxPad = 700;
constexpr int kAdjStash = 4 * kAdjMaxPad + 2;

template <typename Tin, typename Tout, bool ADJ = false>
__device__ __forceinline__ void filter_row_big(
    double2* buf, const Tin* __restrict__ src_row,
                               Tout* __restrict__ dst, int ncol, int pad,
                               int pad_mode, int apply_log,
                               const double2* __restrict__ tw,
                               const double2* __restrict__ G,
                               double* stash = nullptr) {
    constexpr int H = kBigH, M = 2 * H, NT = kBigRowNT;
    const int tid = threadIdx.x;
    const int width = ncol + 2 * pad;
    PadVals pv{0.0, 0.0};
    const bool value_mode =
        !ADJ && pad_mode >= ATH_PAD_LINEAR_RAMP && pad > 0;
    if (value_mode)
        pv = row_pad_values<NT, Tin>(buf, src_row, ncol, pad, pad_mode,
                                     apply_log);
    auto val = [&](int k) {
        return padded_sample<Tin, ADJ>(src_row, k, width, ncol, pad,
                                       pad_mode, apply_log, value_mode, pv);
    };
    BigRegs r;
    big_forward_halves(buf, tw, val, r);
    // kernel product: U into r, V into buf (every index has one owner)
#pragma unroll
    for (int b = 0; b < kBigKB; ++b) {
        const int k = tid + NT * b;
        double2 z0, z1;  // Z[k], Z[H+k]
        big_z(buf, tw, k, r.v[b][0], z0, z1);
        double2 z2 = z0, z3 = z1;  // Z[H-k], Z[2H-k]
        if (k == 0) {
            // the self pairs Z[0] and Z[H]
            spectral_pair(z0, z0, 0, M, tw, G);
            spectral_pair(z1, z1, H, M, tw, G);
        } else {
            big_z(buf, tw, H - k, r.v[b][1], z2, z3);
            spectral_pair(z0, z3, k, M, tw, G);      // (k, M-k)
            spectral_pair(z2, z1, H - k, M, tw, G);  // (H-k, H+k)
        }
        big_store_uv(buf, tw, k, z0, z1, z2, z3, r.v[b]);
    }
    if (tid == 0) {  // the one pair (H/2, 3H/2)
        double2 z0, z1;
        big_z(buf, tw, H / 2, r.mid, z0, z1);
        spectral_pair(z0, z1, H / 2, M, tw, G);
        big_store_mid(buf, tw, z0, z1, r.mid);
    }
    __syncthreads();
    // IFFT_H(V) gives the odd outputs y[2j+1], then IFFT_H(U) the even
    // ones.  Real sample q is component q & 1 of y[q >> 1], at buf[q >> 2]
    // of the transform of parity (q >> 1) & 1.
    const double* y = reinterpret_cast<const double*>(buf);
    auto phase_of = [](int q) { return (q >> 1) & 1; };
    auto at = [&](int q) { return y[2 * (q >> 2) + (q & 1)]; };
    // ADJ: output c gets folds when c <= pad or c >= ncol - pad - 1; its
    // interior value waits at slot c or pad + 1 + c - (ncol - pad - 1), a
    // pad sample n at 2 pad + 2 + (n < pad ? n : n - ncol)
    const int lo_end = pad, hi_beg = ncol - pad - 1;
    auto folded = [&](int c) {
        return pad_mode != ATH_PAD_CONSTANT && (c <= lo_end || c >= hi_beg);
    };
    auto int_slot = [&](int c) {
        return c <= lo_end ? c : pad + 1 + (c - hi_beg);
    };
    auto pad_slot = [&](int n) {
        return 2 * pad + 2 + (n < pad ? n : n - ncol);
    };
    big_inverse_phases(buf, tw, r, [&](int h) {
        if constexpr (!ADJ) {
            // output c is padded sample q = W + pad + c
            for (int c = tid; c < ncol; c += NT) {
                const int q = width + pad + c;
                if (phase_of(q) == h) store_elem(dst + c, at(q));
            }
        } else {
            // z[n] = y[W + n]; this phase's values of the unfolded outputs
            // are stored, the odd phase's values the folds need wait
            for (int n = tid; n < width; n += NT) {
                const int q = width + n;
                if (phase_of(q) != h) continue;
                const bool interior = n >= pad && n < pad + ncol;
                const int c = n - pad;
                if (interior && !folded(c)) {
                    store_elem(dst + c, at(q));
                } else if (h == 1) {
                    stash[interior ? int_slot(c) : pad_slot(n)] = at(q);
                }
            }
            if (h == 0) {
                // the folded outputs, in filter_row's order: the interior
                // value, then the pad samples whose source is c
                auto z = [&](int n) {
                    const int q = width + n;
                    if (phase_of(q) == 0) return at(q);
                    const bool interior = n >= pad && n < pad + ncol;
                    return stash[interior ? int_slot(n - pad) : pad_slot(n)];
                };
                for (int c = tid; c < ncol; c += NT) {
                    if (!folded(c)) continue;
                    double v = z(pad + c);
                    if (pad_mode == ATH_PAD_EDGE) {
                        if (c == 0)
                            for (int n = 0; n < pad; ++n) v += z(n);
                        if (c == ncol - 1)
                            for (int n = pad + ncol; n < width; ++n)
                                v += z(n);
                    } else {
                        for (int n = 0; n < pad; ++n)
                            if (pad_source(n - pad, ncol, pad_mode) == c)
                                v += z(n);
                        for (int n = pad + ncol; n < width; ++n)
                            if (pad_source(n - pad, ncol, pad_mode) == c)
                                v += z(n);
                    }
                    store_elem(dst + c, v);
                }
            }
        }
    });
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kBigRowNT) void ramp_filter_big_kernel(
    const Tin* __restrict__ sino, Tout* __restrict__ out, size_t rows,
    int nslice, int ncol, int pad, int pad_mode, int apply_log,
    const double2* __restrict__ tw, const double2* __restrict__ spec,
    int nwin) {
    extern __shared__ double2 buf[];
    constexpr int M = 2 * kBigH;
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const double2* G =
            spec + (nwin > 1 ? (row / nslice) : 0) * (size_t)(M + 1);
        filter_row_big<Tin, Tout>(buf, sino + row * (size_t)ncol,
                                  out + row * (size_t)ncol, ncol, pad,
                                  pad_mode, apply_log, tw, G);
    }
}

// Transpose of ramp_filter_big_kernel (float32 in and out), persistent;
// the stash follows the row's LDS.
constexpr size_t kBigAdjLds = kBigRowLds + (size_t)kAdjStash * sizeof(double);

__global__ __launch_bounds__(kBigRowNT) void ramp_filter_adjoint_big_kernel(
    const float* __restrict__ in, float* __restrict__ out, size_t rows,
    int nslice, int ncol, int pad, int pad_mode,
    const double2* __restrict__ tw, const double2* __restrict__ spec,
    int nwin) {
    extern __shared__ double2 buf[];
    double* stash = reinterpret_cast<double*>(
        reinterpret_cast<char*>(buf) + kBigRowLds);
    constexpr int M = 2 * kBigH;
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const double2* G =
            spec + (nwin > 1 ? (row / nslice) : 0) * (size_t)(M + 1);
        filter_row_big<float, float, true>(buf, in + row * (size_t)ncol,
                                           out + row * (size_t)ncol, ncol,
                                           pad, pad_mode, 0, tw, G, stash);
    }
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kBigRowNT) void ramp_filter_batched_big_kernel(
    const ath_filter_job* __restrict__ jobs, int njobs, long long total_rows,
    int pad_mode, int apply_log) {
    extern __shared__ double2 buf[];
    constexpr int M = 2 * kBigH;
    for (long long grow = blockIdx.x; grow < total_rows; grow += gridDim.x) {
        int lo = 0, hi = njobs - 1;  // last job with row_begin <= grow
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].row_begin <= grow) lo = mid;
            else hi = mid - 1;
        }
        const ath_filter_job jb = jobs[lo];
        const size_t row = (size_t)(grow - jb.row_begin);
        const double2* tw = static_cast<const double2*>(jb.plan);
        const double2* G =
            tw + 2 * M +
            (jb.nwin > 1 ? (row / jb.nslice) : 0) * (size_t)(M + 1);
        filter_row_big<Tin, Tout>(
            buf, static_cast<const Tin*>(jb.sino) + row * (size_t)jb.ncol,
            static_cast<Tout*>(jb.out) + row * (size_t)jb.ncol, jb.ncol,
            jb.pad, pad_mode, apply_log, tw, G);
    }
}

// spec[r][k] = rfft_L([g_r, g_r, 0...])[k] / M at L = 32768, one block per
// window row: the forward halves as in filter_row_big, then the real split
// of the pairs (k, M-k) and (H-k, H+k).
__global__ __launch_bounds__(kBigRowNT) void kernel_spectrum_big_kernel(
    const double* __restrict__ taps, int width,
    const double2* __restrict__ tw, double2* __restrict__ spec) {
    extern __shared__ double2 buf[];
    constexpr int M = 2 * kBigH;
    const double* g = taps + (size_t)blockIdx.x * width;
    double2* G = spec + (size_t)blockIdx.x * (M + 1);
    auto val = [&](int k) {
        return k < 2 * width ? g[k < width ? k : k - width] : 0.0;
    };
    BigRegs r;
    big_forward_halves(buf, tw, val, r);
    const double inv = 1.0 / (double)M;
    big_pairs(buf, tw, r, [&](int k, double2 zk, double2 zm) {
        double2 xk, xm;
        split_real(zk, zm, tw[k], xk, xm);
        G[k] = cscale(xk, inv);
        G[M - k] = cscale(xm, inv);
    });
}

// Size class of the row kernels for a packed length M: maxb_for(M) up to
// M = 8192, kBigClass for the two-half path at M = 16384, 0 above.
constexpr int kBigClass = 16;
inline int filter_class(int M) {
    if (M == 2 * kBigH) return kBigClass;
    return maxb_for(M);
}

template <typename Tin, typename Tout>
int launch_filter_big(const void* sino, void* out, int nrow, int nslice,
                      int ncol, int pad, int pad_mode, int apply_log,
                      const void* plan, int nwin, hipStream_t stream) {
    const double2* tw = static_cast<const double2*>(plan);
    const double2* spec = tw + 4 * kBigH;
    auto kern = ramp_filter_big_kernel<Tin, Tout>;
    prepare_lds(kern, kBigRowLds);
    const size_t rows = (size_t)nrow * nslice;
    const size_t grid = std::min(
        rows, (size_t)resident_blocks(kern, kBigRowLds, kBigRowNT));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kBigRowNT),
                       kBigRowLds, stream, static_cast<const Tin*>(sino),
                       static_cast<Tout*>(out), rows, nslice, ncol, pad,
                       pad_mode, apply_log, tw, spec, nwin);
    return check_hip(hipGetLastError(), "ramp_filter_big_kernel launch");
}

template <int MAXB, typename Tin, typename Tout>
int launch_filter(const void* sino, void* out, int nrow, int nslice, int ncol,
                  int pad, int pad_mode, int apply_log, const void* plan,
                  int fft_len, int nwin, hipStream_t stream) {
  if constexpr (MAXB == kBigClass) {
    return launch_filter_big<Tin, Tout>(sino, out, nrow, nslice, ncol, pad,
                                        pad_mode, apply_log, plan, nwin,
                                        stream);
  } else {
    const int M = fft_len / 2;
    const double2* tw = static_cast<const double2*>(plan);
    const double2* spec = tw + fft_len;
    const size_t lds = row_lds_bytes<MAXB>(M);
    auto kern = ramp_filter_kernel<MAXB, Tin, Tout>;
    prepare_lds(kern, row_lds_bytes<MAXB>(1024 * MAXB));
    const size_t rows = (size_t)nrow * nslice;
    constexpr int NT = RowCfg<MAXB>::nt;
    const size_t grid = std::min(
        rows, (size_t)resident_blocks(kern, row_lds_bytes<MAXB>(1024 * MAXB),
                                      NT));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), lds, stream,
                       static_cast<const Tin*>(sino), static_cast<Tout*>(out),
                       rows, nslice, ncol, pad, pad_mode, apply_log, tw, spec,
                       nwin, M);
    return check_hip(hipGetLastError(), "ramp_filter_kernel launch");
  }
}

template <int MAXB, typename Tin>
int dispatch_out(int out_dtype, const void* sino, void* out, int nrow,
                 int nslice, int ncol, int pad, int pad_mode, int apply_log,
                 const void* plan, int fft_len, int nwin, hipStream_t s) {
    switch (out_dtype) {
    case ATH_F32:
        return launch_filter<MAXB, Tin, float>(sino, out, nrow, nslice, ncol,
                                               pad, pad_mode, apply_log, plan,
                                               fft_len, nwin, s);
    case ATH_F64:
        return launch_filter<MAXB, Tin, double>(sino, out, nrow, nslice, ncol,
                                                pad, pad_mode, apply_log, plan,
                                                fft_len, nwin, s);
    case ATH_F16:
        return launch_filter<MAXB, Tin, __half>(sino, out, nrow, nslice, ncol,
                                                pad, pad_mode, apply_log, plan,
                                                fft_len, nwin, s);
    case ATH_BF16:
        return launch_filter<MAXB, Tin, __hip_bfloat16>(
            sino, out, nrow, nslice, ncol, pad, pad_mode, apply_log, plan,
            fft_len, nwin, s);
    }
    return set_error(ATH_ERR_INVALID_ARG, "ramp filter: bad output dtype");
}

template <int MAXB, typename Tin, typename Tout>
int launch_filter_batched(const ath_filter_job* jobs, int njobs,
                          long long total_rows, int fft_len, int pad_mode,
                          int apply_log, hipStream_t stream) {
  if constexpr (MAXB == kBigClass) {
    auto kern = ramp_filter_batched_big_kernel<Tin, Tout>;
    prepare_lds(kern, kBigRowLds);
    const long long grid = std::min<long long>(
        total_rows, resident_blocks(kern, kBigRowLds, kBigRowNT));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kBigRowNT),
                       kBigRowLds, stream, jobs, njobs, total_rows, pad_mode,
                       apply_log);
    return check_hip(hipGetLastError(),
                     "ramp_filter_batched_big_kernel launch");
  } else {
    const int M = fft_len / 2;
    const size_t lds = row_lds_bytes<MAXB>(M);
    auto kern = ramp_filter_batched_kernel<MAXB, Tin, Tout>;
    prepare_lds(kern, row_lds_bytes<MAXB>(1024 * MAXB));
    constexpr int NT = RowCfg<MAXB>::nt;
    const long long grid = std::min<long long>(
        total_rows,
        resident_blocks(kern, row_lds_bytes<MAXB>(1024 * MAXB), NT));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), lds, stream,
                       jobs, njobs, total_rows, pad_mode, apply_log, M);
    return check_hip(hipGetLastError(), "ramp_filter_batched_kernel launch");
  }
}

template <int MAXB, typename Tin>
int batched_out(int out_dtype, const ath_filter_job* jobs, int njobs,
                long long total_rows, int fft_len, int pad_mode,
                int apply_log, hipStream_t s) {
    switch (out_dtype) {
    case ATH_F32:
        return launch_filter_batched<MAXB, Tin, float>(
            jobs, njobs, total_rows, fft_len, pad_mode, apply_log, s);
    case ATH_F64:
        return launch_filter_batched<MAXB, Tin, double>(
            jobs, njobs, total_rows, fft_len, pad_mode, apply_log, s);
    case ATH_F16:
        return launch_filter_batched<MAXB, Tin, __half>(
            jobs, njobs, total_rows, fft_len, pad_mode, apply_log, s);
    case ATH_BF16:
        return launch_filter_batched<MAXB, Tin, __hip_bfloat16>(
            jobs, njobs, total_rows, fft_len, pad_mode, apply_log, s);
    }
    return set_error(ATH_ERR_INVALID_ARG, "ramp filter: bad output dtype");
}

template <int MAXB>
int batched_in(int in_dtype, int out_dtype, const ath_filter_job* jobs,
               int njobs, long long total_rows, int fft_len, int pad_mode,
               int apply_log, hipStream_t s) {
    if (in_dtype == ATH_F32)
        return batched_out<MAXB, float>(out_dtype, jobs, njobs, total_rows,
                                        fft_len, pad_mode, apply_log, s);
    if (in_dtype == ATH_F64)
        return batched_out<MAXB, double>(out_dtype, jobs, njobs, total_rows,
                                         fft_len, pad_mode, apply_log, s);
    return set_error(ATH_ERR_INVALID_ARG,
                     "ramp filter: input dtype must be float32 or float64");
}

template <int MAXB>
int dispatch_in(int in_dtype, int out_dtype, const void* sino, void* out,
                int nrow, int nslice, int ncol, int pad, int pad_mode,
                int apply_log, const void* plan, int fft_len, int nwin,
                hipStream_t s) {
    if (in_dtype == ATH_F32)
        return dispatch_out<MAXB, float>(out_dtype, sino, out, nrow, nslice,
                                         ncol, pad, pad_mode, apply_log, plan,
                                         fft_len, nwin, s);
    if (in_dtype == ATH_F64)
        return dispatch_out<MAXB, double>(out_dtype, sino, out, nrow, nslice,
                                          ncol, pad, pad_mode, apply_log, plan,
                                          fft_len, nwin, s);
    return set_error(ATH_ERR_INVALID_ARG,
                     "ramp filter: input dtype must be float32 or float64");
}

template <int MAXB>
int launch_spectrum(const double* taps, int nwin, int width,
                    const double2* tw, double2* spec, int M,
                    hipStream_t stream) {
    const size_t lds = (size_t)M * sizeof(double2);
    auto kern = kernel_spectrum_kernel<MAXB>;
    prepare_lds(kern, (size_t)1024 * MAXB * sizeof(double2));
    hipLaunchKernelGGL(kern, dim3((unsigned)nwin), dim3(kThreads), lds,
                       stream, taps, width, tw, spec, M);
    return check_hip(hipGetLastError(), "kernel_spectrum_kernel launch");
}

}  // namespace

int fft_length(int ncol, int pad) {
    const long long width = (long long)ncol + 2LL * pad;
    const long long need = 2 * width - pad;
    long long L = 64;
    while (L < need) L <<= 1;
    return L > (1 << 30) ? -1 : (int)L;
}

size_t plan_bytes(int fft_len, int nwin) {
    const size_t M = (size_t)fft_len / 2;
    return sizeof(double2) * ((size_t)fft_len + (size_t)nwin * (M + 1));
}

int plan_build(const double* taps, int nwin, int width, int fft_len,
               void* plan, hipStream_t stream) {
    const int M = fft_len / 2;
    double2* tw = static_cast<double2*>(plan);
    double2* spec = tw + fft_len;
    hipLaunchKernelGGL(twiddle_kernel, dim3((fft_len + 255) / 256), dim3(256),
                       0, stream, tw, fft_len);
    int st = check_hip(hipGetLastError(), "twiddle_kernel launch");
    if (st != ATH_OK) return st;
    switch (filter_class(M)) {
    case 1: return launch_spectrum<1>(taps, nwin, width, tw, spec, M, stream);
    case 2: return launch_spectrum<2>(taps, nwin, width, tw, spec, M, stream);
    case 4: return launch_spectrum<4>(taps, nwin, width, tw, spec, M, stream);
    case 8: return launch_spectrum<8>(taps, nwin, width, tw, spec, M, stream);
    case kBigClass: {
        auto kern = kernel_spectrum_big_kernel;
        prepare_lds(kern, kBigRowLds);
        hipLaunchKernelGGL(kern, dim3((unsigned)nwin), dim3(kBigRowNT),
                           kBigRowLds, stream, taps, width, tw, spec);
        return check_hip(hipGetLastError(),
                         "kernel_spectrum_big_kernel launch");
    }
    }
    return set_error(ATH_ERR_UNSUPPORTED,
                     "ramp filter: FFT length above 32768 is not supported");
}

int ramp_filter_batched(const ath_filter_job* jobs, int njobs,
                        long long total_rows, int in_dtype, int out_dtype,
                        int fft_len, int pad_mode, int apply_log,
                        hipStream_t s) {
    switch (filter_class(fft_len / 2)) {
    case 1: return batched_in<1>(in_dtype, out_dtype, jobs, njobs, total_rows,
                                 fft_len, pad_mode, apply_log, s);
    case 2: return batched_in<2>(in_dtype, out_dtype, jobs, njobs, total_rows,
                                 fft_len, pad_mode, apply_log, s);
    case 4: return batched_in<4>(in_dtype, out_dtype, jobs, njobs, total_rows,
                                 fft_len, pad_mode, apply_log, s);
    case 8: return batched_in<8>(in_dtype, out_dtype, jobs, njobs, total_rows,
                                 fft_len, pad_mode, apply_log, s);
    case kBigClass:
        return batched_in<kBigClass>(in_dtype, out_dtype, jobs, njobs,
                                     total_rows, fft_len, pad_mode, apply_log,
                                     s);
    }
    return set_error(ATH_ERR_UNSUPPORTED,
                     "ramp filter: FFT length above 32768 is not supported");
}

namespace {
template <int MAXB>
int launch_adjoint(const float* in, int nrow, int nslice, int ncol, int pad,
                   int pad_mode, const void* plan, int fft_len, int nwin,
                   float* out, hipStream_t stream) {
    const int M = fft_len / 2;
    const double2* tw = static_cast<const double2*>(plan);
    const double2* spec = tw + fft_len;
    const size_t lds = row_lds_bytes<MAXB>(M);
    auto kern = ramp_filter_adjoint_kernel<MAXB>;
    prepare_lds(kern, row_lds_bytes<MAXB>(1024 * MAXB));
    const size_t rows = (size_t)nrow * nslice;
    constexpr int NT = RowCfg<MAXB>::nt;
    const size_t grid = std::min(
        rows, (size_t)resident_blocks(kern, row_lds_bytes<MAXB>(1024 * MAXB),
                                      NT));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), lds, stream, in,
                       out, rows, nslice, ncol, pad, pad_mode, tw, spec, nwin,
                       M);
    return check_hip(hipGetLastError(), "ramp_filter_adjoint_kernel launch");
}
}  // namespace

int ramp_filter_adjoint(const float* in, int nrow, int nslice, int ncol,
                        int pad, int pad_mode, const void* plan, int fft_len,
                        int nwin, float* out, hipStream_t s) {
    if (fft_len == 4 * kBigH) {
        if (pad > kAdjMaxPad || pad >= ncol)
            return set_error(ATH_ERR_UNSUPPORTED,
                             "ramp filter adjoint at FFT length 32768: pad "
                             "must be below ncol and at most 700");
        const double2* tw = static_cast<const double2*>(plan);
        const double2* spec = tw + fft_len;
        auto kern = ramp_filter_adjoint_big_kernel;
        prepare_lds(kern, kBigAdjLds);
        const size_t rows = (size_t)nrow * nslice;
        const size_t grid = std::min(
            rows, (size_t)resident_blocks(kern, kBigAdjLds, kBigRowNT));
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kBigRowNT),
                           kBigAdjLds, s, in, out, rows, nslice, ncol, pad,
                           pad_mode, tw, spec, nwin);
        return check_hip(hipGetLastError(),
                         "ramp_filter_adjoint_big_kernel launch");
    }
    switch (maxb_for(fft_len / 2)) {
    case 1: return launch_adjoint<1>(in, nrow, nslice, ncol, pad, pad_mode,
                                     plan, fft_len, nwin, out, s);
    case 2: return launch_adjoint<2>(in, nrow, nslice, ncol, pad, pad_mode,
                                     plan, fft_len, nwin, out, s);
    case 4: return launch_adjoint<4>(in, nrow, nslice, ncol, pad, pad_mode,
                                     plan, fft_len, nwin, out, s);
    case 8: return launch_adjoint<8>(in, nrow, nslice, ncol, pad, pad_mode,
                                     plan, fft_len, nwin, out, s);
    }
    return set_error(ATH_ERR_UNSUPPORTED,
                     "ramp filter adjoint: FFT length above 32768 is not "
                     "supported");
}

int ramp_filter(const void* sino, int in_dtype, int nrow, int nslice,
                int ncol, int pad, int pad_mode, int apply_log,
                const void* plan, int fft_len, int nwin, void* out,
                int out_dtype, hipStream_t s) {
    const int M = fft_len / 2;
    switch (filter_class(M)) {
    case kBigClass:
        return dispatch_in<kBigClass>(in_dtype, out_dtype, sino, out, nrow,
                                      nslice, ncol, pad, pad_mode, apply_log,
                                      plan, fft_len, nwin, s);
    case 1: return dispatch_in<1>(in_dtype, out_dtype, sino, out, nrow, nslice,
                                  ncol, pad, pad_mode, apply_log, plan,
                                  fft_len, nwin, s);
    case 2: return dispatch_in<2>(in_dtype, out_dtype, sino, out, nrow, nslice,
                                  ncol, pad, pad_mode, apply_log, plan,
                                  fft_len, nwin, s);
    case 4: return dispatch_in<4>(in_dtype, out_dtype, sino, out, nrow, nslice,
                                  ncol, pad, pad_mode, apply_log, plan,
                                  fft_len, nwin, s);
    case 8: return dispatch_in<8>(in_dtype, out_dtype, sino, out, nrow, nslice,
                                  ncol, pad, pad_mode, apply_log, plan,
                                  fft_len, nwin, s);
    }
    return set_error(ATH_ERR_UNSUPPORTED,
                     "ramp filter: FFT length above 32768 is not supported");
}

}  // namespace ath
