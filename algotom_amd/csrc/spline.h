// Cubic and quintic B-spline pieces shared by the kernels that reproduce
// scipy.ndimage's spline interpolation (shift, map_coordinates): the
// B-spline weights, the boundary index folds and the inverse-filter taps
// that replace scipy's recursive spline prefilter by a symmetric FIR.
#pragma once

#include "ath_internal.h"

namespace ath {

constexpr int kK3 = 32;          // cubic inverse-filter half length
constexpr int kK5 = 52;          // quintic inverse-filter half length
constexpr int kNearestPad = 12;  // scipy's pre-padding for mode "nearest"

__device__ __forceinline__ double beta3(double t) {
    t = fabs(t);
    if (t < 1.0) return 2.0 / 3.0 - t * t + 0.5 * t * t * t;
    if (t < 2.0) { const double u = 2.0 - t; return u * u * u / 6.0; }
    return 0.0;
}

__device__ __forceinline__ double beta5(double t) {
    t = fabs(t);
    if (t < 1.0) {
        const double t2 = t * t;
        return 11.0 / 20.0 + t2 * (-0.5 + t2 * (0.25 - t / 12.0));
    }
    if (t < 2.0)
        return 17.0 / 40.0 +
               t * (5.0 / 8.0 +
                    t * (-7.0 / 4.0 +
                         t * (5.0 / 4.0 + t * (-3.0 / 8.0 + t / 24.0))));
    if (t < 3.0) {
        const double u = 3.0 - t, u2 = u * u;
        return u2 * u2 * u / 120.0;
    }
    return 0.0;
}

// half-sample symmetric (scipy "reflect") and whole-sample ("mirror")
// index folds onto [0, n)
__device__ __forceinline__ int fold_half(int i, int n) {
    return pad_source(i, n, ATH_PAD_SYMMETRIC);
}
__device__ __forceinline__ int fold_whole(int i, int n) {
    return pad_source(i, n, ATH_PAD_REFLECT);
}

// 1 / B(q) of the sampled cubic B-spline as a two-sided FIR:
// h[k] = (1 - z)^2 z^|k| / (1 - z^2), z = sqrt(3) - 2 (the pole inside the
// unit circle of z + 1/z = -4); |h[32]| < 1e-18 |h[0]|.
__device__ __forceinline__ double cubic_tap(int k) {
    const double w = -4.0;
    const double z = 2.0 / (w - sqrt(w * w - 4.0));
    return (1.0 - z) * (1.0 - z) * pow(z, (double)k) / (1.0 - z * z);
}

// scipy.ndimage.shift(x, sh, order=3, mode="nearest") of one line of n
// samples, by the block of NT threads.  line: L1 = n + 24 doubles holding
// the line edge-padded by kNearestPad on both sides; coef: L1 doubles of
// scratch; hs: cubic_tap(0..kK3).  Writes out(j, value) for j < n with the
// value at pos = (j - sh) + 12 of the half-sample-symmetric prefiltered
// line, sampled with clamped indices.  Ends with the LDS free for reuse.
template <int NT, typename Out>
__device__ __forceinline__ void shift_line_nearest(const double* line,
                                                   double* coef, int n,
                                                   double sh,
                                                   const double* hs,
                                                   Out out) {
    const int L1 = n + 2 * kNearestPad;
    const int tid = threadIdx.x;
    for (int p = tid; p < L1; p += NT) {
        double c = hs[0] * line[p];
        for (int t = 1; t <= kK3; ++t)
            c += hs[t] * (line[fold_half(p - t, L1)] +
                          line[fold_half(p + t, L1)]);
        coef[p] = c;
    }
    __syncthreads();
    for (int j = tid; j < n; j += NT) {
        const double pos = ((double)j - sh) + (double)kNearestPad;
        const int k0 = (int)floor(pos) - 1;
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + q;
            v += beta3(pos - (double)k) * coef[min(max(k, 0), L1 - 1)];
        }
        out(j, v);
    }
    __syncthreads();
}

}  // namespace ath
