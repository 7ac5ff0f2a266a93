// Adjoint of the back-projection: a forward projection (Radon transform with
// the back-projector's own linear-interpolation weights), for the backward
// pass of differentiable FBP.
//
// The back-projector adds, for pixel (x, y) and angle i with
// f = ((x - ic) c + (y - ic) s) + center in [0, ncol-1], d = floor f,
// t = f - d, the value (1 - t) s[d] + t s[d+1] (reconstruction.py:170-181).
// Its transpose therefore sends (1 - t) g(x, y) to detector bin d and
// t g(x, y) to bin d + 1.  One thread owns one (angle, bin) pair and gathers
// every pixel whose f lies in [bin - 1, bin + 1]: it walks the image along
// the axis closer to the detector normal (rows when |c| >= |s|), and per
// step tests the <= 2/max(|c|,|s|) + 3 candidate pixels with the exact
// float64 f of the back-projector.  No atomics: the sum order per bin is
// fixed, so results are deterministic, and the weights are bit-identical to
// the forward pass's.  Accumulation is float64.
#include "ath_internal.h"

namespace ath {
namespace {

constexpr int kFpThreads = 64;
constexpr int kFpMaxSlices = 16;

__device__ __forceinline__ double fp_det_pos(double xc_c, double yc, double s,
                                             double cen) {
    return fma(yc, s, xc_c) + cen;
}

template <int NS>
__global__ __launch_bounds__(kFpThreads) void forward_project_kernel(
    FpArgs a, int j0) {
    const int d = blockIdx.x * kFpThreads + threadIdx.x;
    const int i = blockIdx.y;
    const int C = a.ncol, S = a.nslice;
    if (d >= C) return;
    const int ns = min(NS, S - j0);
    const float th = -a.angles[i];
    const double c = (double)(float)cos((double)th);
    const double s = (double)(float)sin((double)th);
    const double ic = a.icenter, cen = a.center, w1 = (double)(C - 1);
    double acc[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) acc[j] = 0.0;
    const bool by_rows = fabs(c) >= fabs(s);
    const double lo_t = (double)d - 1.0, hi_t = (double)d + 1.0;
    for (int m = 0; m < C; ++m) {
        // candidate range of the other coordinate on line m
        const double mc = (double)m - ic;
        const double base = by_rows ? mc * s + cen : mc * c + cen;
        const double slope = by_rows ? c : s;
        double u0 = (lo_t - base) / slope, u1 = (hi_t - base) / slope;
        if (u0 > u1) {
            const double t = u0;
            u0 = u1;
            u1 = t;
        }
        int k0 = (int)floor(u0 + ic) - 1, k1 = (int)ceil(u1 + ic) + 1;
        k0 = max(k0, 0);
        k1 = min(k1, C - 1);
        for (int k = k0; k <= k1; ++k) {
            const int x = by_rows ? k : m;
            const int y = by_rows ? m : k;
            const double f =
                fp_det_pos(((double)x - ic) * c, (double)y - ic, s, cen);
            if (!(f >= 0.0 && f <= w1)) continue;
            const double fl = floor(f);
            const double t = f - fl;
            double w;
            if (fl == (double)d) w = 1.0 - t;
            else if (fl == (double)(d - 1)) w = t;
            else continue;
            if (a.use_mask) {
                const long long dx = x - a.mask_c, dy = y - a.mask_c;
                if (!((double)(dx * dx + dy * dy) <= a.mask_r2)) continue;
            }
            const float* px = a.image + ((size_t)y * S + j0) * C + x;
#pragma unroll
            for (int j = 0; j < NS; ++j)
                if (j < ns) acc[j] += w * (double)px[(size_t)j * C];
        }
    }
#pragma unroll
    for (int j = 0; j < NS; ++j)
        if (j < ns)
            a.sino[((size_t)i * S + j0 + j) * C + d] = (float)(acc[j] * a.scale);
}

}  // namespace

int forward_projection(const FpArgs& a, hipStream_t stream) {
    const dim3 grid((a.ncol + kFpThreads - 1) / kFpThreads, a.nrow);
    int j0 = 0;
    while (j0 < a.nslice) {
        const int rest = a.nslice - j0;
        int step;
        if (rest >= kFpMaxSlices) {
            hipLaunchKernelGGL(forward_project_kernel<kFpMaxSlices>, grid,
                               dim3(kFpThreads), 0, stream, a, j0);
            step = kFpMaxSlices;
        } else if (rest >= 4) {
            hipLaunchKernelGGL(forward_project_kernel<4>, grid,
                               dim3(kFpThreads), 0, stream, a, j0);
            step = 4;
        } else {
            hipLaunchKernelGGL(forward_project_kernel<1>, grid,
                               dim3(kFpThreads), 0, stream, a, j0);
            step = 1;
        }
        const int st = check_hip(hipGetLastError(),
                                 "forward_project_kernel launch");
        if (st != ATH_OK) return st;
        j0 += step;
    }
    return ATH_OK;
}

}  // namespace ath
