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


__device__ __forceinline__ double fp_det_pos(double xc_c, double yc, double s,
                                             double cen) {
    return fma(yc, s, xc_c) + cen;
}

// LDS-staged variant.  A workgroup = 4 waves = 4 consecutive angles x the
// same 64 detector bins x FSB slices.  The image is walked in chunks of
// kRows lines along the major axis; for each chunk the union of the four
// angles' candidate windows over the 64 (+2) bins is staged into LDS once
// (line-major, FSB slices interleaved) and every lane then gathers its
// candidates from LDS.  Each staged pixel serves up to 4 angles x 2 bins.
// Chunks whose union window is wider than kMaxMinor columns read global
// memory instead.  The walking axis keeps |slope| >= ~0.69, so 64 bins span
// <= 93 columns plus the margins and the 4 angles' spread: 128 covers every
// block of closely spaced angles, and the small tile (16 KB at 8 slices x 4
// lines) lets 6+ workgroups share a CU to hide the staging loads.
constexpr int kMaxMinor = 128;
// below this |slope| a line runs (almost) along the ray: take it whole
constexpr double kMinSlope = 1.0 / 1024.0;

template <int FSB, int kRows>
__global__ __launch_bounds__(256) void forward_project_lds_kernel(
    FpArgs a, int by_rows_block) {
    __shared__ __align__(16) float tile[kRows * kMaxMinor * FSB];
    __shared__ double s_cs[4][2];
    __shared__ double s_inv[4];   // 1 / slope, 0 for near-parallel lines
    __shared__ int s_ab[2];       // the chunk's staged column range
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int C = a.ncol, S = a.nslice, N = a.nrow;
    const int d0 = blockIdx.x * 64;
    const int i = blockIdx.y * 4 + wv;
    const int j0 = blockIdx.z * FSB;
    const int ns = min(FSB, S - j0);
    const int d = d0 + lane;
    const double ic = a.icenter, cen = a.center, w1 = (double)(C - 1);
    if (tid < 4) {
        const int ia = min(blockIdx.y * 4 + tid, N - 1);
        const float th = -a.angles[ia];
        s_cs[tid][0] = (double)(float)cos((double)th);
        s_cs[tid][1] = (double)(float)sin((double)th);
    }
    __syncthreads();
    const bool valid = (i < N) && (d < C);
    const double c = s_cs[wv][0], s = s_cs[wv][1];
    // walking direction of the whole block: rows when the first angle has
    // |cos| >= |sin| (exact windows keep any direction correct)
    const bool by_rows = by_rows_block < 0
                             ? fabs(s_cs[0][0]) >= fabs(s_cs[0][1])
                             : by_rows_block != 0;
    if (tid < 4) {
        const double sl = by_rows ? s_cs[tid][0] : s_cs[tid][1];
        s_inv[tid] = fabs(sl) >= kMinSlope ? 1.0 / sl : 0.0;
    }
    const double slope = by_rows ? c : s;
    // windows only need to be a superset (every candidate is tested with
    // the exact f), so a reciprocal replaces the per-line divisions
    const bool steep = fabs(slope) >= kMinSlope;
    const double inv_slope = steep ? 1.0 / slope : 0.0;
    double acc[FSB];
#pragma unroll
    for (int j = 0; j < FSB; ++j) acc[j] = 0.0;

    for (int m0 = 0; m0 < C; m0 += kRows) {
        const int m1 = min(m0 + kRows, C) - 1;
        // union window of the minor coordinate over 4 angles, bins
        // [d0 - 1, d0 + 64] and lines m0, m1 (windows are linear in m):
        // lanes 0-15 of wave 0 each evaluate one end, reduce min / max
        // across those 16 lanes and lane 0 publishes the column range
        if (tid < 64) {
            double lo = 1e300, hi = -1e300;
            if (tid < 16) {
                const int q = tid >> 2, e = (tid >> 1) & 1, b = tid & 1;
                const double mc = (double)(e ? m1 : m0) - ic;
                const double base =
                    (by_rows ? mc * s_cs[q][1] : mc * s_cs[q][0]) + cen;
                const double bin = b ? (double)d0 + 65.0 : (double)d0 - 2.0;
                // a near-parallel line (inverse 0) takes the whole line
                const double u = s_inv[q] != 0.0 ? (bin - base) * s_inv[q]
                                                 : (b ? 1e300 : -1e300);
                lo = u;
                hi = u;
            }
            // lanes 0-15 are closed under these xor offsets
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) {
                lo = fmin(lo, __shfl_xor(lo, off));
                hi = fmax(hi, __shfl_xor(hi, off));
            }
            // a negative inverse swaps the two bin ends: min / max over
            // all sixteen covers both orders
            if (tid == 0) {
                s_ab[0] = lo <= -1e299 ? 0
                                       : max((int)floor(lo + ic) - 1, 0);
                s_ab[1] = hi >= 1e299 ? C - 1
                                      : min((int)ceil(hi + ic) + 1, C - 1);
            }
        }
        __syncthreads();
        const int ka = s_ab[0], kb = s_ab[1];
        const int wid = kb - ka + 1;
        const bool staged = wid <= kMaxMinor;
        if (wid > 0 && staged) {
            // one pixel per thread and step: the mask test once, then its
            // FSB slices (kMaxMinor = 128 = 2^7 columns per line)
            const int nrows = m1 - m0 + 1;
            for (int e = tid; e < kRows * kMaxMinor; e += 256) {
                int r, k;
                if (by_rows) {  // lanes along x within a line
                    r = e >> 7;
                    k = e & (kMaxMinor - 1);
                } else {        // lanes along the lines (x) first
                    r = e % kRows;
                    k = e / kRows;
                }
                if (r >= nrows || k >= wid) continue;
                const int x = by_rows ? ka + k : m0 + r;
                const int y = by_rows ? m0 + r : ka + k;
                bool keep = true;
                if (a.use_mask) {
                    const long long dx = x - a.mask_c, dy = y - a.mask_c;
                    keep = (double)(dx * dx + dy * dy) <= a.mask_r2;
                }
                const float* src = a.image + ((size_t)y * S + j0) * C + x;
                float v[FSB];
#pragma unroll
                for (int js = 0; js < FSB; ++js)
                    v[js] = (keep && js < ns) ? src[(size_t)js * C] : 0.0f;
                float* dst = tile + (r * kMaxMinor + k) * FSB;
                if constexpr (FSB % 4 == 0) {
#pragma unroll
                    for (int q = 0; q < FSB / 4; ++q)
                        reinterpret_cast<float4*>(dst)[q] = make_float4(
                            v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
                } else {
#pragma unroll
                    for (int js = 0; js < FSB; ++js) dst[js] = v[js];
                }
            }
        }
        __syncthreads();
        if (valid && wid > 0) {
            const double lo_t = (double)d - 1.0, hi_t = (double)d + 1.0;
            for (int m = m0; m <= m1; ++m) {
                const double mc = (double)m - ic;
                const double base = by_rows ? mc * s + cen : mc * c + cen;
                int k0 = 0, k1 = C - 1;
                if (steep) {
                    double u0 = (lo_t - base) * inv_slope;
                    double u1 = (hi_t - base) * inv_slope;
                    if (u0 > u1) {
                        const double t = u0;
                        u0 = u1;
                        u1 = t;
                    }
                    // contributors have f in [d-1, d+1), i.e. k - ic in
                    // [u0, u1]; u carries a relative error of a few ulp
                    // (reciprocal slope), so a 1e-6 px margin keeps the
                    // candidate list a superset (each is tested exactly).
                    // Every candidate costs the whole slice loop once any
                    // lane of the wave contributes, so the list is kept
                    // to the integers inside the margin.
                    k0 = max((int)ceil(u0 + ic - 1e-6), 0);
                    k1 = min((int)floor(u1 + ic + 1e-6), C - 1);
                }
                if (staged) {
                    k0 = max(k0, ka);
                    k1 = min(k1, kb);
                }
                for (int k = k0; k <= k1; ++k) {
                    const int x = by_rows ? k : m;
                    const int y = by_rows ? m : k;
                    const double f = fp_det_pos(((double)x - ic) * c,
                                                (double)y - ic, s, cen);
                    if (!(f >= 0.0 && f <= w1)) continue;
                    const double fl = floor(f);
                    const double t = f - fl;
                    double w;
                    if (fl == (double)d) w = 1.0 - t;
                    else if (fl == (double)(d - 1)) w = t;
                    else continue;
                    if (staged) {
                        const float* q =
                            tile + ((m - m0) * kMaxMinor + (k - ka)) * FSB;
#pragma unroll
                        for (int j = 0; j < FSB; ++j)
                            acc[j] += w * (double)q[j];
                    } else {
                        if (a.use_mask) {
                            const long long dx = x - a.mask_c,
                                            dy = y - a.mask_c;
                            if (!((double)(dx * dx + dy * dy) <= a.mask_r2))
                                continue;
                        }
                        const float* px =
                            a.image + ((size_t)y * S + j0) * C + x;
#pragma unroll
                        for (int j = 0; j < FSB; ++j)
                            if (j < ns) acc[j] += w * (double)px[(size_t)j * C];
                    }
                }
            }
        }
        __syncthreads();  // tile free for the next chunk
    }
    if (valid) {
#pragma unroll
        for (int j = 0; j < FSB; ++j)
            if (j < ns)
                a.sino[((size_t)i * S + j0 + j) * C + d] =
                    (float)(acc[j] * a.scale);
    }
}

}  // namespace

// Angle blocks of 4 share one walking direction: rows when the block's
// first angle has |cos| >= |sin|.  The candidate windows are exact for any
// direction, so a block straddling 45 degrees stays correct (its windows
// are at most 2/0.70 + 3 wide).  Launched per direction-homogeneous run of
// blocks: one launch for each maximal run.
int forward_projection(const FpArgs& a, hipStream_t stream) {
    // host copy of the walking direction per block needs the angles: they
    // live on the device, so the kernel decides per block instead
    // slices per workgroup and lines per chunk measured on MI355X
    // (16 x 901 x 1024: 16/4 9.9 ms, 8/4 11.2 ms, 16/8 12.7 ms)
    const int fsb = a.nslice >= 16 ? 16
                  : a.nslice >= 8 ? 8 : (a.nslice >= 4 ? 4 : 1);
    const dim3 grid((a.ncol + 63) / 64, (a.nrow + 3) / 4,
                    (a.nslice + fsb - 1) / fsb);
    if (fsb == 16)
        hipLaunchKernelGGL((forward_project_lds_kernel<16, 4>), grid,
                           dim3(256), 0, stream, a, -1);
    else if (fsb == 8)
        hipLaunchKernelGGL((forward_project_lds_kernel<8, 4>), grid,
                           dim3(256), 0, stream, a, -1);
    else if (fsb == 4)
        hipLaunchKernelGGL((forward_project_lds_kernel<4, 8>), grid,
                           dim3(256), 0, stream, a, -1);
    else
        hipLaunchKernelGGL((forward_project_lds_kernel<1, 16>), grid,
                           dim3(256), 0, stream, a, -1);
    return check_hip(hipGetLastError(), "forward_project_lds_kernel launch");
}

}  // namespace ath
