// LDS FFT building blocks shared by the 1D ramp filter (ramp_filter.hip)
// and the 2D circular ramp filter (filter2d.hip): float64 complex helpers, a
// radix-4 Stockham FFT over a power-of-two length held in LDS, the real-FFT
// split, the twiddle table and the large-LDS launch attribute.  Everything
// sits in an unnamed namespace: each translation unit gets its own copy.
#pragma once

#include "ath_internal.h"

namespace ath {
namespace {

constexpr int kFftThreads = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 conjg(double2 a) {
    return make_double2(a.x, -a.y);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) {
    return make_double2(a.x * s, a.y * s);
}
// multiply by -i (forward) or +i (inverse)
template <int DIR>
__device__ __forceinline__ double2 rot(double2 a) {
    return DIR < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// Quarter twiddle table layout: entry m sits at swz_tw(m), which moves only
// the low four bits of m by a function of the higher bits (a permutation of
// every aligned block of 16 entries).  The pass at span ns reads entries
// k * M / (4 ns), 64 .. 4096 bytes apart for the spans 4 .. 256, which in
// natural order fall on the same banks (up to 16-way conflicts).
__device__ __forceinline__ int swz_tw(int t) {
    return t ^ (((t >> 4) ^ (t >> 8)) & 15);
}

// Quarter twiddle table for a length-M transform: tq[m] = exp(-2 pi i m/M)
// for m < M/4, taken from the plan's length-2M table (tw[2m]), stored at
// swz_tw(m).  All threads.
template <int NT = kFftThreads>
__device__ __forceinline__ void stage_quarter(double2* tq,
                                              const double2* __restrict__ tw,
                                              int M) {
    for (int m = threadIdx.x; m < (M >> 2); m += NT) tq[swz_tw(m)] = tw[2 * m];
}

// exp(-2 pi i j/M) for j < M/2 from the quarter table: the second quarter
// is the first turned by -i (exact).
__device__ __forceinline__ double2 twiddle_half(const double2* tq, int j,
                                                int M) {
    const int q = M >> 2;
    if (j < q) return tq[swz_tw(j)];
    const double2 t = tq[swz_tw(j - q)];
    return make_double2(t.y, -t.x);
}

// 8-point DFT in place, natural order in and out (two 4-point DFTs of the
// even and odd inputs, then W8^k and the radix-2 combination).
template <int DIR>
__device__ __forceinline__ void dft4(double2& x0, double2& x1, double2& x2,
                                     double2& x3) {
    const double2 a0 = cadd(x0, x2), a1 = csub(x0, x2);
    const double2 a2 = cadd(x1, x3), a3 = rot<DIR>(csub(x1, x3));
    x0 = cadd(a0, a2);
    x1 = cadd(a1, a3);
    x2 = csub(a0, a2);
    x3 = csub(a1, a3);
}
template <int DIR>
__device__ __forceinline__ void dft8(double2 (&v)[8]) {
    constexpr double h = 0.70710678118654752440;  // sqrt(2)/2
    double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    double2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    // o_k *= W8^k (forward W8 = exp(-i pi/4); the inverse conjugates)
    if (DIR < 0) {
        o1 = make_double2((o1.x + o1.y) * h, (o1.y - o1.x) * h);
        o3 = make_double2((o3.y - o3.x) * h, -(o3.x + o3.y) * h);
    } else {
        o1 = make_double2((o1.x - o1.y) * h, (o1.x + o1.y) * h);
        o3 = make_double2(-(o3.x + o3.y) * h, (o3.x - o3.y) * h);
    }
    o2 = rot<DIR>(o2);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

// In-LDS complex FFT of length M = 2^m, natural order in and out (Stockham
// autosort: each pass reads all operands into registers, barriers, then
// writes them to their sorted place in the same buffer).  Radix-8 passes
// (4 for M = 4096 instead of 6 radix-4 ones: fewer barriers and LDS round
// trips) and at most one radix-4 and one radix-2 pass.  DIR = -1 is the
// forward transform, +1 the unscaled inverse.  tw[n] = exp(-2 pi i n / L),
// L = 2M.  Each of the NT threads handles up to MAXB butterflies a pass.
// A radix-4 butterfly fetches one twiddle w and forms w^2, w^3 by complex
// products (one table access instead of three).  With QT the twiddles come
// from the LDS quarter table tq (stage_quarter) instead of global memory.
// TWS: the table's index stride (tw[n * TWS] = exp(-2 pi i n / L)), so a
// transform can read the table of a multiple of its length.
template <int DIR, int MAXB, bool QT = false, int NT = kFftThreads,
          int TWS = 1, bool ROT = false>
__device__ void lds_fft(double2* buf, int M, const double2* __restrict__ tw,
                        const double2* tq = nullptr) {
    static_assert(TWS == 1 || !QT, "the quarter table has stride 1");
    const int tid = threadIdx.x;
    const int L = 2 * M;
    int ns = 1;
    // radix-8 passes first (M/8 butterflies: B8 per thread), then at most one
    // radix-4 and one radix-2 pass for the remaining factor
    constexpr int B8 = (MAXB + 1) / 2;
    // The first radix-8 pass writes 8 consecutive elements per lane (8-way
    // bank conflicts); with ROT each lane pair's outputs are rotated, which
    // removes them at the cost of 14 complex products and 10-90 more VGPRs.
    // Taken by the DFI Bluestein transforms (1.75 -> 1.68 ms per 1801 x 2560
    // slice), the two-level transforms below (DFI 16384-point columns 1.10
    // -> 1.00 ms, BPF 16384-point columns 3.31 -> 3.11 ms) and the 1D
    // filter's row kernels without the quarter table (7192-wide volume rows
    // 4.54 -> 4.22 ms).  Not by the 2D filter's row and column kernels:
    // there the registers cost occupancy (rows 62 -> 148 VGPRs, 8 -> 3
    // waves per SIMD; 0.43 -> 0.66 ms per 8-image row pass).  Never with the
    // LDS quarter table.
    constexpr bool kRotate = ROT && !QT;
    while (ns * 8 <= M) {
        const int nb = M >> 3;
        const int tstep = L / (ns * 8);
        double2 v[B8][8];
#pragma unroll
        for (int b = 0; b < B8; ++b) {
            const int j = tid + b * NT;
            if (j < nb) {
#pragma unroll
                for (int r = 0; r < 8; ++r) v[b][r] = buf[j + r * nb];
                const int k = j & (ns - 1);
                if (kRotate && ns == 1) {
                    // first pass: rotate the outputs by rr = (j >> 1) & 7
                    // (inputs times W8^(n rr), so output t is X[(t+rr)&7])
                    // so that the 16 lanes of a group, which write 8
                    // consecutive elements each, hit 16 distinct banks
                    const int rr = (j >> 1) & 7;
                    double2 wr = tw[rr * (L >> 3) * TWS];  // W8^rr
                    if (DIR > 0) wr = conjg(wr);
                    double2 w = wr;
#pragma unroll
                    for (int r = 1; r < 8; ++r) {
                        v[b][r] = cmul(v[b][r], w);
                        w = cmul(w, wr);
                    }
                } else if (k != 0) {
                    double2 w1 = QT ? tq[swz_tw(k * (tstep >> 1))]
                                    : tw[k * tstep * TWS];
                    if (DIR > 0) w1 = conjg(w1);
                    const double2 w2 = cmul(w1, w1);
                    const double2 w3 = cmul(w2, w1);
                    const double2 w4 = cmul(w2, w2);
                    v[b][1] = cmul(v[b][1], w1);
                    v[b][2] = cmul(v[b][2], w2);
                    v[b][3] = cmul(v[b][3], w3);
                    v[b][4] = cmul(v[b][4], w4);
                    v[b][5] = cmul(v[b][5], cmul(w4, w1));
                    v[b][6] = cmul(v[b][6], cmul(w3, w3));
                    v[b][7] = cmul(v[b][7], cmul(w4, w3));
                }
                dft8<DIR>(v[b]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < B8; ++b) {
            const int j = tid + b * NT;
            if (j < nb) {
                const int k = j & (ns - 1);
                const int base = (j - k) * 8 + k;
                if (kRotate && ns == 1) {
                    const int rr = (j >> 1) & 7;
#pragma unroll
                    for (int r = 0; r < 8; ++r)
                        buf[base + ((r + rr) & 7)] = v[b][r];
                } else {
#pragma unroll
                    for (int r = 0; r < 8; ++r) buf[base + r * ns] = v[b][r];
                }
            }
        }
        __syncthreads();
        ns *= 8;
    }
    while (ns * 4 <= M) {
        const int nb = M >> 2;
        const int tstep = L / (ns * 4);
        double2 v[MAXB][4];
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            const int j = tid + b * NT;
            if (j < nb) {
#pragma unroll
                for (int r = 0; r < 4; ++r) v[b][r] = buf[j + r * nb];
                const int k = j & (ns - 1);
                if (k != 0) {
                    double2 w1 = QT ? tq[swz_tw(k * (tstep >> 1))]
                                    : tw[k * tstep * TWS];
                    if (DIR > 0) w1 = conjg(w1);
                    const double2 w2 = cmul(w1, w1);
                    const double2 w3 = cmul(w1, w2);
                    v[b][1] = cmul(v[b][1], w1);
                    v[b][2] = cmul(v[b][2], w2);
                    v[b][3] = cmul(v[b][3], w3);
                }
                const double2 a0 = cadd(v[b][0], v[b][2]);
                const double2 a1 = csub(v[b][0], v[b][2]);
                const double2 a2 = cadd(v[b][1], v[b][3]);
                const double2 a3 = rot<DIR>(csub(v[b][1], v[b][3]));
                v[b][0] = cadd(a0, a2);
                v[b][1] = cadd(a1, a3);
                v[b][2] = csub(a0, a2);
                v[b][3] = csub(a1, a3);
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            const int j = tid + b * NT;
            if (j < nb) {
                const int k = j & (ns - 1);
                const int base = (j - k) * 4 + k;
#pragma unroll
                for (int r = 0; r < 4; ++r) buf[base + r * ns] = v[b][r];
            }
        }
        __syncthreads();
        ns *= 4;
    }
    if (ns < M) {  // one radix-2 pass, ns = M / 2
        const int nb = M >> 1;
        double2 v[2 * MAXB][2];
#pragma unroll
        for (int b = 0; b < 2 * MAXB; ++b) {
            const int j = tid + b * NT;
            if (j < nb) {
                v[b][0] = buf[j];
                double2 x1 = buf[j + nb];
                if (j != 0) {
                    double2 w = QT ? twiddle_half(tq, j, M)
                                   : tw[2 * j * TWS];
                    if (DIR > 0) w = conjg(w);
                    x1 = cmul(x1, w);
                }
                v[b][1] = csub(v[b][0], x1);
                v[b][0] = cadd(v[b][0], x1);
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < 2 * MAXB; ++b) {
            const int j = tid + b * NT;
            if (j < nb) {
                buf[j] = v[b][0];
                buf[j + ns] = v[b][1];
            }
        }
        __syncthreads();
    }
}

// Transforms of length 2H = 16384, the next size past what one workgroup
// holds in LDS in float64.  conv_big forms the circular convolution
// IFFT_2H(FFT_2H(x) S) (inverse unscaled; S carries the scale) from H-point
// LDS transforms: the even and odd samples forward (decimation in time,
// X[k] = E[k] + W^k O[k], X[k + H] = E[k] - W^k O[k], W = exp(-2 pi i/2H)),
// the product Y = X S, then the inverse transforms of U = Y[k] + Y[k + H]
// (even outputs) and V = (Y[k] - Y[k + H]) W^-k (odd outputs) -- decimation
// in frequency.  E, U and the odd outputs pass through a per-workgroup
// global scratch of 2H values.  512 threads of 4 butterflies.
constexpr int kBigH = 8192;
constexpr int kBigNT = 512;
constexpr int kBigB = 4;  // butterflies per thread: kBigH / 4 / kBigNT

// LOAD(n): x[n] for n < N (zero beyond).  tw: the H-point FFT's table,
// tw[k] = exp(-2 pi i k / 2H) = W^k for k < 2H.  On exit output n is
// big_conv(buf, scratch, n).  All kBigNT threads.
template <typename Load>
__device__ void conv_big(double2* buf, double2* scratch, int N,
                         const double2* __restrict__ tw,
                         const double2* __restrict__ S, Load load) {
    constexpr int H = kBigH, NT = kBigNT;
    const int tid = threadIdx.x;
    for (int j = tid; j < H; j += NT)
        buf[j] = 2 * j < N ? load(2 * j) : make_double2(0.0, 0.0);
    __syncthreads();
    lds_fft<-1, kBigB, false, NT, 1, true>(buf, H, tw);
    for (int k = tid; k < H; k += NT) scratch[k] = buf[k];
    __syncthreads();
    for (int j = tid; j < H; j += NT)
        buf[j] = 2 * j + 1 < N ? load(2 * j + 1) : make_double2(0.0, 0.0);
    __syncthreads();
    lds_fft<-1, kBigB, false, NT, 1, true>(buf, H, tw);
    for (int k = tid; k < H; k += NT) {
        const double2 w = tw[k];
        const double2 e = scratch[k];
        const double2 o = cmul(buf[k], w);
        const double2 ylo = cmul(cadd(e, o), S[k]);
        const double2 yhi = cmul(csub(e, o), S[k + H]);
        scratch[k] = cadd(ylo, yhi);
        buf[k] = cmul(csub(ylo, yhi), conjg(w));
    }
    __syncthreads();
    lds_fft<+1, kBigB, false, NT, 1, true>(buf, H, tw);
    for (int q = tid; q < H; q += NT) scratch[H + q] = buf[q];
    __syncthreads();
    for (int k = tid; k < H; k += NT) buf[k] = scratch[k];
    __syncthreads();
    lds_fft<+1, kBigB, false, NT, 1, true>(buf, H, tw);
}

__device__ __forceinline__ double2 big_conv(const double2* buf,
                                            const double2* scratch, int n) {
    return (n & 1) ? scratch[kBigH + (n >> 1)] : buf[n >> 1];
}

// Transforms of length 4H = 32768: the radix-4 form of conv_big.  Forward,
// decimation in time: F_r = FFT_H(x[4j + r]) for r = 0..3, then
// X[k + qH] = sum_r (-i)^(rq) W^(rk) F_r[k], W = exp(-2 pi i / 4H).  The
// product Y = X S.  Inverse, decimation in frequency:
// V_s[k] = W^(-sk) sum_q i^(sq) Y[k + qH] and y[4j + s] = IFFT_H(V_s)[j].
// F_0..F_2, then V_0..V_2 and the outputs of phases 1..3 pass through a
// per-workgroup global scratch of 4H values; phase 0 ends in buf.  tw is the
// 4H-point table (the H-point transforms read it at stride 2).
// kBigNT threads.
constexpr int kBig4M = 4 * kBigH;

// F_r = FFT_H(x[4j + r]) (LOAD(n) for n < N, zero beyond): F_0..F_2 into
// scratch[rH ..], F_3 left in buf.
template <typename Load>
__device__ void fft_big4_quarters(double2* buf, double2* scratch, int N,
                                  const double2* __restrict__ tw, Load load) {
    constexpr int H = kBigH, NT = kBigNT;
    const int tid = threadIdx.x;
    for (int r = 0; r < 4; ++r) {
        for (int j = tid; j < H; j += NT) {
            const int n = 4 * j + r;
            buf[j] = n < N ? load(n) : make_double2(0.0, 0.0);
        }
        __syncthreads();
        lds_fft<-1, kBigB, false, NT, 2, true>(buf, H, tw);
        if (r < 3) {
            for (int k = tid; k < H; k += NT) scratch[r * H + k] = buf[k];
            __syncthreads();
        }
    }
}

// the four outputs X[k + qH] from the quarters' values at k
__device__ __forceinline__ void big4_combine(double2 f0, double2 f1,
                                             double2 f2, double2 f3, int k,
                                             const double2* __restrict__ tw,
                                             double2 (&x)[4]) {
    const double2 t0 = f0;
    const double2 t1 = cmul(f1, tw[k]);
    const double2 t2 = cmul(f2, tw[2 * k]);
    const double2 t3 = cmul(f3, tw[3 * k]);
    const double2 a = cadd(t0, t2), b = csub(t0, t2);
    const double2 c = cadd(t1, t3), d = rot<-1>(csub(t1, t3));  // -i(t1-t3)
    x[0] = cadd(a, c);
    x[1] = cadd(b, d);
    x[2] = csub(a, c);
    x[3] = csub(b, d);
}

// Forward FFT of length 4H of x (LOAD as in conv_big4): STORE(k, X[k]) for
// every k < 4H.  All kBigNT threads; ends synchronised.
template <typename Load, typename Store>
__device__ void fft_big4_forward(double2* buf, double2* scratch, int N,
                                 const double2* __restrict__ tw, Load load,
                                 Store store) {
    fft_big4_quarters(buf, scratch, N, tw, load);
    for (int k = threadIdx.x; k < kBigH; k += kBigNT) {
        double2 x[4];
        big4_combine(scratch[k], scratch[kBigH + k], scratch[2 * kBigH + k],
                     buf[k], k, tw, x);
#pragma unroll
        for (int q = 0; q < 4; ++q) store(k + q * kBigH, x[q]);
    }
    __syncthreads();
}

template <typename Load>
__device__ void conv_big4(double2* buf, double2* scratch, int N,
                          const double2* __restrict__ tw,
                          const double2* __restrict__ S, Load load) {
    constexpr int H = kBigH, NT = kBigNT;
    const int tid = threadIdx.x;
    fft_big4_quarters(buf, scratch, N, tw, load);
    for (int k = tid; k < H; k += NT) {
        double2 y[4];
        big4_combine(scratch[k], scratch[H + k], scratch[2 * H + k], buf[k], k,
                     tw, y);
#pragma unroll
        for (int q = 0; q < 4; ++q) y[q] = cmul(y[q], S[k + q * H]);
        // sum_q i^(sq) Y_q for s = 0..3
        const double2 a = cadd(y[0], y[2]), b = csub(y[0], y[2]);
        const double2 c = cadd(y[1], y[3]), d = rot<1>(csub(y[1], y[3]));
        scratch[k] = cadd(a, c);
        scratch[H + k] = cmul(cadd(b, d), conjg(tw[k]));
        scratch[2 * H + k] = cmul(csub(a, c), conjg(tw[2 * k]));
        buf[k] = cmul(csub(b, d), conjg(tw[3 * k]));
    }
    __syncthreads();
    // phases 3, 2, 1 into scratch, then phase 0 in buf
    for (int s = 3; s >= 0; --s) {
        if (s < 3) {
            for (int k = tid; k < H; k += NT) buf[k] = scratch[s * H + k];
            __syncthreads();
        }
        lds_fft<1, kBigB, false, NT, 2, true>(buf, H, tw);
        if (s > 0) {
            for (int j = tid; j < H; j += NT) scratch[s * H + j] = buf[j];
            __syncthreads();
        }
    }
}

__device__ __forceinline__ double2 big4_conv(const double2* buf,
                                             const double2* scratch, int n) {
    const int s = n & 3, j = n >> 2;
    return s == 0 ? buf[j] : scratch[s * kBigH + j];
}

// fft_big_forward (LV = 2) or fft_big4_forward (LV = 4)
template <int LV, typename Load, typename Store>
__device__ __forceinline__ void fft_lv_forward(double2* buf, double2* scratch,
                                               int N,
                                               const double2* __restrict__ tw,
                                               Load load, Store store) {
    if constexpr (LV == 2) fft_big_forward(buf, scratch, N, tw, load, store);
    else fft_big4_forward(buf, scratch, N, tw, load, store);
}

// conv_big (LV = 2, M = 2H) or conv_big4 (LV = 4, M = 4H); scratch of
// LV * H values
template <int LV, typename Load>
__device__ __forceinline__ void conv_lv(double2* buf, double2* scratch, int N,
                                        const double2* __restrict__ tw,
                                        const double2* __restrict__ S,
                                        Load load) {
    if constexpr (LV == 2) conv_big(buf, scratch, N, tw, S, load);
    else conv_big4(buf, scratch, N, tw, S, load);
}

template <int LV>
__device__ __forceinline__ double2 conv_lv_out(const double2* buf,
                                               const double2* scratch, int n) {
    if constexpr (LV == 2) return (n & 1) ? scratch[kBigH + (n >> 1)]
                                          : buf[n >> 1];
    else return big4_conv(buf, scratch, n);
}

// Forward FFT of length 2H of x (LOAD as in conv_big): STORE(k, X[k]) for
// every k < 2H.  All kBigNT threads.
template <typename Load, typename Store>
__device__ void fft_big_forward(double2* buf, double2* scratch, int N,
                                const double2* __restrict__ tw, Load load,
                                Store store) {
    constexpr int H = kBigH, NT = kBigNT;
    const int tid = threadIdx.x;
    for (int j = tid; j < H; j += NT)
        buf[j] = 2 * j < N ? load(2 * j) : make_double2(0.0, 0.0);
    __syncthreads();
    lds_fft<-1, kBigB, false, NT, 1, true>(buf, H, tw);
    for (int k = tid; k < H; k += NT) scratch[k] = buf[k];
    __syncthreads();
    for (int j = tid; j < H; j += NT)
        buf[j] = 2 * j + 1 < N ? load(2 * j + 1) : make_double2(0.0, 0.0);
    __syncthreads();
    lds_fft<-1, kBigB, false, NT, 1, true>(buf, H, tw);
    for (int k = tid; k < H; k += NT) {
        const double2 e = scratch[k];
        const double2 o = cmul(buf[k], tw[k]);
        store(k, cadd(e, o));
        store(k + H, csub(e, o));
    }
    __syncthreads();
}

// Z = FFT_M(x[2n] + i x[2n+1]) -> X[k] = rfft_L(x)[k] for k and M-k.
__device__ __forceinline__ void split_real(double2 zk, double2 zm, double2 wk,
                                           double2& xk, double2& xm) {
    const double2 e = cscale(cadd(zk, conjg(zm)), 0.5);
    const double2 o = rot<-1>(cscale(csub(zk, conjg(zm)), 0.5));
    const double2 wo = cmul(wk, o);
    xk = cadd(e, wo);
    xm = conjg(csub(e, wo));
}

// The inverse of split_real for the pair (k, km = (M - k) mod M),
// 0 <= k <= M/2: from the real spectrum's values Y[k], Y[M-k] the packed
// input of the length-M inverse at k and km (zm untouched when k == km).
// tw: the length-2M table.
__device__ __forceinline__ void merge_pair(double2 yk, double2 ym, int k,
                                           int M,
                                           const double2* __restrict__ tw,
                                           double2& zk, double2& zm) {
    const int km = (M - k) & (M - 1);
    const double2 wk = tw[k];
    // Zy[k] = (Yk + conj Ym)/2 + i conj(wk) (Yk - conj Ym)/2
    const double2 ek = cscale(cadd(yk, conjg(ym)), 0.5);
    const double2 ok = rot<1>(cmul(conjg(wk),
                                   cscale(csub(yk, conjg(ym)), 0.5)));
    zk = cadd(ek, ok);
    if (k != km) {
        // Zy[M-k] = (Ym + conj Yk)/2 - i wk (Ym - conj Yk)/2
        const double2 em = cscale(cadd(ym, conjg(yk)), 0.5);
        const double2 om = rot<-1>(cmul(wk,
                                        cscale(csub(ym, conjg(yk)), 0.5)));
        zm = cadd(em, om);
    }
}

// ---- real transforms of length 4H = 32768 (M = 2H packed points) ---------
// One workgroup of kBigNT threads per row, H = 8192 complex float64 in LDS
// and no global scratch.  Forward, decimation in time: E = FFT_H(z[2j])
// moves into registers (thread t keeps E[k] and E[H-k] for its
// k = t + NT b < H/2, thread 0 also E[H/2]) while O = FFT_H(z[2j+1]) runs in
// LDS; then Z[k] = E[k] + W^k O[k] and Z[k+H] = E[k] - W^k O[k]
// (W = exp(-2 pi i/M)) at the four indices k, H-k, H+k, 2H-k, which are the
// two pairs (k, M-k) and (H-k, H+k) of the real split.  Inverse,
// decimation in frequency: U[k] = Y[k] + Y[k+H] stays in registers and
// V[k] = (Y[k] - Y[k+H]) W^-k goes to LDS; IFFT_H(V) gives the odd outputs
// y[2j+1], then IFFT_H(U) the even ones.  The H-point transforms read the
// 4H-point table at stride 2.
constexpr int kBigKB = kBigH / 2 / kBigNT;  // k slots per thread (8)

struct BigRegs {
    double2 v[kBigKB][2];
    double2 mid;
};

// E = FFT_H of the packed samples val(4j), val(4j+1) into r; then O =
// FFT_H of val(4j+2), val(4j+3) left in buf.  All threads; ends
// synchronised.
template <typename Val>
__device__ __forceinline__ void big_forward_halves(
    double2* buf, const double2* __restrict__ tw, Val val, BigRegs& r) {
    constexpr int H = kBigH, NT = kBigNT;
    const int tid = threadIdx.x;
    for (int h = 0; h < 2; ++h) {
        if (h == 1) {
#pragma unroll
            for (int b = 0; b < kBigKB; ++b) {
                const int k = tid + NT * b;
                r.v[b][0] = buf[k];
                r.v[b][1] = buf[(H - k) & (H - 1)];
            }
            if (tid == 0) r.mid = buf[H / 2];
            __syncthreads();
        }
        for (int j = tid; j < H; j += NT)
            buf[j] = make_double2(val(4 * j + 2 * h), val(4 * j + 2 * h + 1));
        __syncthreads();
        lds_fft<-1, kBigB, false, NT, 2, true>(buf, H, tw);
    }
}

// Z[k] and Z[k + H] from E[k] (register) and O[k] (buf)
__device__ __forceinline__ void big_z(const double2* buf,
                                      const double2* __restrict__ tw, int k,
                                      double2 ek, double2& zlo, double2& zhi) {
    const double2 o = cmul(buf[k], tw[2 * k]);
    zlo = cadd(ek, o);
    zhi = csub(ek, o);
}

// After big_forward_halves: put(k, Z[k], Z[(M - k) mod M]) once for every
// pair of the real split, 0 <= k <= M/2 (the self pairs 0 and H with the
// same value twice).  Each pair is one thread's.
template <typename Put>
__device__ __forceinline__ void big_pairs(const double2* buf,
                                          const double2* __restrict__ tw,
                                          const BigRegs& r, Put put) {
    constexpr int H = kBigH, NT = kBigNT;
    const int tid = threadIdx.x;
#pragma unroll
    for (int b = 0; b < kBigKB; ++b) {
        const int k = tid + NT * b;
        double2 z0, z1;
        big_z(buf, tw, k, r.v[b][0], z0, z1);
        if (k == 0) {
            put(0, z0, z0);
            put(H, z1, z1);
        } else {
            double2 z2, z3;
            big_z(buf, tw, H - k, r.v[b][1], z2, z3);
            put(k, z0, z3);
            put(H - k, z2, z1);
        }
    }
    if (tid == 0) {
        double2 z0, z1;
        big_z(buf, tw, H / 2, r.mid, z0, z1);
        put(H / 2, z0, z1);
    }
}

// From the inverse's inputs Y at k, H+k (y0, y1) and, for k > 0, at H-k,
// 2H-k (y2, y3): U into the slot's registers, V into buf.
__device__ __forceinline__ void big_store_uv(double2* buf,
                                             const double2* __restrict__ tw,
                                             int k, double2 y0, double2 y1,
                                             double2 y2, double2 y3,
                                             double2 (&u)[2]) {
    constexpr int H = kBigH;
    u[0] = cadd(y0, y1);
    buf[k] = cmul(csub(y0, y1), conjg(tw[2 * k]));
    if (k != 0) {
        u[1] = cadd(y2, y3);
        buf[H - k] = cmul(csub(y2, y3), conjg(tw[2 * (H - k)]));
    }
}

// the same for the index H/2 (y at H/2 and 3H/2), thread 0
__device__ __forceinline__ void big_store_mid(double2* buf,
                                              const double2* __restrict__ tw,
                                              double2 y0, double2 y1,
                                              double2& mid) {
    mid = cadd(y0, y1);
    buf[kBigH / 2] = cmul(csub(y0, y1), conjg(tw[kBigH]));
}

// The two inverse phases after big_store_uv / big_store_mid (callers
// synchronise first): out(1) runs with the odd outputs in buf, out(0) with
// the even ones.  Real sample q is component q & 1 of y[q >> 1], at
// buf[q >> 2] in phase (q >> 1) & 1.  Ends synchronised.
template <typename Out>
__device__ __forceinline__ void big_inverse_phases(
    double2* buf, const double2* __restrict__ tw, const BigRegs& r, Out out) {
    constexpr int H = kBigH, NT = kBigNT;
    const int tid = threadIdx.x;
    for (int h = 1; h >= 0; --h) {
        if (h == 0) {
#pragma unroll
            for (int b = 0; b < kBigKB; ++b) {
                const int k = tid + NT * b;
                buf[k] = r.v[b][0];
                if (k != 0) buf[H - k] = r.v[b][1];
            }
            if (tid == 0) buf[H / 2] = r.mid;
            __syncthreads();
        }
        lds_fft<1, kBigB, false, NT, 2, true>(buf, H, tw);
        out(h);
        __syncthreads();  // buf free for U, then for the caller
    }
}

// tw[n] = exp(-2 pi i n / L), n < L
__global__ void twiddle_kernel(double2* tw, int L) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < L) {
        double s, c;
        sincospi(2.0 * (double)n / (double)L, &s, &c);
        tw[n] = make_double2(c, -s);
    }
}

// Butterflies per thread and pass for a length-M transform (M <= 8192).
inline int maxb_for(int M) {
    if (M <= 1024) return 1;
    if (M <= 2048) return 2;
    if (M <= 4096) return 4;
    if (M <= 8192) return 8;
    return 0;
}

// Workgroups of `kernel` resident on the current device at once with `lds`
// bytes of dynamic LDS each: the grid of a persistent launch (cached per
// kernel, device, block size and LDS size; 1024 if the query fails).
template <typename K>
int resident_blocks(K kernel, size_t lds, int threads = kFftThreads) {
    int n = 0;
    if (resident_grid(reinterpret_cast<const void*>(kernel), threads, lds,
                      &n) != ATH_OK) {
        clear_error();
        return 1024;
    }
    return n;
}

// Allows dynamic LDS above 64 KB for `kernel` on the current device (a
// host-side attribute, so later launches, including graph-captured ones,
// enqueue only the kernel).  `bytes` is the largest size the kernel is
// launched with.  Set once per kernel and device, and only marked done
// when it succeeded.
template <typename K>
void prepare_lds(K kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    if (allow_dynamic_lds(reinterpret_cast<const void*>(kernel), bytes) !=
        ATH_OK)
        clear_error();
}

}  // namespace
}  // namespace ath
