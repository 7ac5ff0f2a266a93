// -log of a float32 sample, rounded to float32: the value of
// (float)(-log((double)x)), the arithmetic numpy applies to a float32 array
// (reconstruction.py:436, vertrec.py flat-field and log), at a fraction of
// the cost of the float64 library log.
//
// y ~ log(x) is formed in float64 with a relative error below 1e-15:
//   |x - 1| < 2^-7:  u = x - 1 is exact in float32 (Sterbenz) and
//                    log1p(u) is a degree-9 Taylor polynomial;
//   otherwise:       x = m 2^e with m in [1, 2); the top 7 bits of m pick a
//                    centre c (the cell midpoint, halved with e + 1 for
//                    m >= 1.4140625 so that c lies in [0.707, 1.414)), and
//                    log x = e ln2 + log c + log1p(t), t = (m - c) / c,
//                    |t| < 2^-8, log1p(t) a degree-6 polynomial; m - c is
//                    exact and 1/c, log c come from a 128-entry table
//                    computed at compile time (atanh series).
// Rounding y to float32 gives the same value as rounding the library's
// float64 log unless y lies within 2^-40 |y| of a float32 rounding boundary
// (probability about 3e-5 per sample); those samples, and zero, negative,
// subnormal, infinite and NaN inputs, take the library log itself.
#pragma once

#include <hip/hip_runtime.h>

namespace ath {
namespace {

struct Log32Table {
    double rc[128];  // 1 / c_j
    double lg[128];  // log(c_j)
};

constexpr double log32_series(double c) {
    // log c = 2 atanh(s), s = (c - 1)/(c + 1), |s| < 0.18: 30 terms,
    // summed from the smallest
    const double s = (c - 1.0) / (c + 1.0);
    const double s2 = s * s;
    double terms[30] = {};
    double p = s;
    for (int k = 0; k < 30; ++k) {
        terms[k] = p / (double)(2 * k + 1);
        p *= s2;
    }
    double sum = 0.0;
    for (int k = 29; k >= 0; --k) sum += terms[k];
    return 2.0 * sum;
}

constexpr Log32Table make_log32_table() {
    Log32Table t{};
    for (int j = 0; j < 128; ++j) {
        double c = 1.0 + (double)(2 * j + 1) / 256.0;
        if (j >= 53) c *= 0.5;
        t.rc[j] = 1.0 / c;
        t.lg[j] = log32_series(c);
    }
    return t;
}

__device__ constexpr Log32Table kLog32Device = make_log32_table();
constexpr Log32Table kLog32Host = make_log32_table();

__host__ __device__ inline const Log32Table& log32_table() {
#if defined(__HIP_DEVICE_COMPILE__)
    return kLog32Device;
#else
    return kLog32Host;
#endif
}

__host__ __device__ inline float log32_bits_to_float(unsigned u) {
    return __builtin_bit_cast(float, u);
}

// Copies the table into LDS (all threads of the workgroup take part; the
// caller synchronises before use).  Elementwise kernels read it from there:
// two scattered 8-byte global loads per sample would otherwise outnumber
// the kernel's own vector loads.
template <int NT>
__device__ __forceinline__ void stage_log32_table(Log32Table* s) {
    const double* src = &kLog32Device.rc[0];
    double* dst = &s->rc[0];
    for (int i = threadIdx.x; i < 256; i += NT) dst[i] = src[i];
}

// Host and device (the host build serves the exhaustive check against the
// C library's float64 log); ``tab`` is the table in any address space.
__host__ __device__ __forceinline__ float neg_log_f32(float x,
                                                      const Log32Table& tab) {
    const unsigned b = __builtin_bit_cast(unsigned, x);
    const unsigned ex = b >> 23;  // sign and exponent fields
    if (ex - 1u >= 254u) return (float)(-log((double)x));
    double y;
    const float u = x - 1.0f;
    if (fabsf(u) < 0x1p-7f) {
        const double d = (double)u;
        double p = 1.0 / 9.0;
        p = fma(p, d, -1.0 / 8.0);
        p = fma(p, d, 1.0 / 7.0);
        p = fma(p, d, -1.0 / 6.0);
        p = fma(p, d, 1.0 / 5.0);
        p = fma(p, d, -1.0 / 4.0);
        p = fma(p, d, 1.0 / 3.0);
        p = fma(p, d, -1.0 / 2.0);
        p = fma(p, d, 1.0);
        y = d * p;
    } else {
        const unsigned mb = b & 0x7fffffu;
        const int j = (int)(mb >> 16);
        const int hi = j >= 53;
        const double sc = hi ? 0.5 : 1.0;
        const double m = (1.0 + (double)mb * 0x1p-23) * sc;
        const double c = (1.0 + (double)(2 * j + 1) * 0x1p-8) * sc;
        const double t = (m - c) * tab.rc[j];
        double p = -1.0 / 6.0;
        p = fma(p, t, 1.0 / 5.0);
        p = fma(p, t, -1.0 / 4.0);
        p = fma(p, t, 1.0 / 3.0);
        p = fma(p, t, -1.0 / 2.0);
        p = fma(p, t, 1.0);
        const double e = (double)((int)ex - 127 + hi);
        y = fma(e, 0x1.62e42fefa39efp-1, tab.lg[j] + t * p);
    }
    const float f = (float)y;
    if (f == 0.0f) return -f;  // x == 1
    // the rounding boundary on y's side of f, in magnitude
    const float fa = fabsf(f);
    const unsigned fb = __builtin_bit_cast(unsigned, fa);
    const double ra = f > 0.0f ? y - (double)f : (double)f - y;
    const double nb =
        (double)log32_bits_to_float(ra >= 0.0 ? fb + 1u : fb - 1u);
    const double half = 0.5 * fabs(nb - (double)fa);
    if (fabs(fabs(ra) - half) <= 0x1p-40 * fabs(y))
        return (float)(-log((double)x));
    return -f;
}

__host__ __device__ __forceinline__ float neg_log_f32(float x) {
    return neg_log_f32(x, log32_table());
}

}  // namespace
}  // namespace ath
