// Linear-interpolation back-projection on gfx950.
//
// Reference arithmetic (algotom/rec/reconstruction.py:156-190, 229-264,
// 307-359), per pixel (x, y) and angle i, with numba's typing:
//   c, s  = float32 cos/sin of float32 -angle[i]
//   f     = ((x - ic) * c + (y - ic) * s) + center        (float64; ic =
//           0.5 (ncol - 1), both products exact in float64)
//   if 0 <= f <= ncol-1:  d = floor f,  val = s[d] + (s[d+1] - s[d]) * (f - d)
//                         (float32 difference, float64 product and sum)
//   accumulate val: float64 register (2D kernel) or float32 after every
//   angle (chunk and CPU kernels).
//
// Kernel design (bp_tiled_kernel): a 256-thread workgroup owns a 64 x (4*PY)
// pixel tile of SB slices.  Lane l of wave w owns pixel column x0+l and the
// PY rows y0 + w*PY + k, so one wave instruction touches 64 consecutive x
// and its LDS gathers hit consecutive columns (no bank conflicts).  Angles
// are processed in blocks of A: for each angle the workgroup computes the
// tile's detector footprint [dlo, dlo+SPAN) exactly from its four corners,
// stages (s[d], s[d+1]-s[d]) pairs of the SB slices interleaved per column
// into LDS, and every thread then gathers one 8*SB-byte record per pixel
// and angle.  f, d and the weight are computed once per pixel and angle
// and reused by all SB slices.  Angles whose footprint lies inside
// [0, ncol-1] for the whole tile (every interior tile) skip the range test;
// the rest take the exact per-pixel test.  Each pixel sums its angles in
// ascending order, as the reference does, so ATH_ACC_F64 and
// ATH_ACC_F32SEQ reproduce the reference kernels bit for bit given the same
// filtered sinogram.
#include "ath_internal.h"

namespace ath {
namespace {

constexpr int kThreads = 256;
constexpr int kTileX = 64;

constexpr int isqrt_ceil(int v) {
    int r = 0;
    while (r * r < v) ++r;
    return r;
}
// Columns staged per angle: the footprint of a TX x TY tile spans at most
// ceil(sqrt((TX-1)^2 + (TY-1)^2)) + 1 columns (|cos|, |sin| float32 with
// cos^2 + sin^2 <= 1 + 2^-23); two more for rounding slack.
constexpr int bp_span(int tx, int ty) {
    return isqrt_ceil((tx - 1) * (tx - 1) + (ty - 1) * (ty - 1)) + 3;
}

// detector position exactly as the reference evaluates it
__device__ __forceinline__ double det_pos(double xc_c, double yc, double s,
                                          double cen) {
    return fma(yc, s, xc_c) + cen;
}

template <int MODE>
struct Acc;
template <>
struct Acc<ATH_ACC_F64> {
    double v = 0.0;
    __device__ __forceinline__ void add(float2 q, double w) {
        v += (double)q.x + (double)q.y * w;
    }
    __device__ __forceinline__ void add_value(float e) { v += (double)e; }
    __device__ __forceinline__ float result() const { return (float)v; }
};
template <>
struct Acc<ATH_ACC_F32SEQ> {
    float v = 0.0f;
    __device__ __forceinline__ void add(float2 q, double w) {
        v = (float)((double)v + ((double)q.x + (double)q.y * w));
    }
    __device__ __forceinline__ void add_value(float e) {
        v = (float)((double)v + (double)e);
    }
    __device__ __forceinline__ float result() const { return v; }
};
template <>
struct Acc<ATH_ACC_F32> {
    float v = 0.0f;
    __device__ __forceinline__ void add(float2 q, float w) {
        v += fmaf(q.y, w, q.x);
    }
    __device__ __forceinline__ void add_value(float e) { v += e; }
    __device__ __forceinline__ float result() const { return v; }
};

template <int MODE>
struct Weight {
    using type = double;
    static __device__ __forceinline__ double make(double w) { return w; }
};
template <>
struct Weight<ATH_ACC_F32> {
    using type = float;
    static __device__ __forceinline__ float make(double w) { return (float)w; }
};

// mask (utility.py:160-182) and pi/(nrow-1) (reconstruction.py:493), in the
// float32 order numpy evaluates them
__device__ __forceinline__ float finalize(const BpArgs& a, float v, int x,
                                          int y) {
    if (a.use_mask) {
        const long long dx = x - a.mask_c, dy = y - a.mask_c;
        const float m = ((double)(dx * dx + dy * dy) <= a.mask_r2) ? 1.0f : 0.0f;
        v = v * m;
    }
    if (a.apply_scale) {
        v = v * a.pi32;
        v = v / a.denom;
    }
    return v;
}

template <int MODE, int SB, int PY, int A, typename Tin>
__global__ __launch_bounds__(kThreads) void bp_tiled_kernel(
    BpArgs a, int tiles_x, int tiles_y) {
    constexpr int TY = 4 * PY;
    constexpr int SPAN = bp_span(kTileX, TY);
    __shared__ float2 s_pair[A * SPAN * SB];
    __shared__ double s_cos[A], s_sin[A];
    __shared__ int s_dlo[A], s_inner[A];
    __shared__ float s_edge[A * SB * 2];

    const Tin* __restrict__ sino = static_cast<const Tin*>(a.sino);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    unsigned bid = blockIdx.x;
    const int tx = bid % tiles_x;
    bid /= tiles_x;
    const int ty = bid % tiles_y;
    const int grp = bid / tiles_y;
    const int x0 = tx * kTileX, y0 = ty * TY, j0 = grp * SB;
    const int C = a.ncol, S = a.nslice, N = a.nrow;
    const double ic = a.icenter, cen = a.center, w1 = (double)(C - 1);

    const int x = x0 + lane;
    const double xc = (double)x - ic;
    double yc[PY];
#pragma unroll
    for (int k = 0; k < PY; ++k) yc[k] = (double)(y0 + wv * PY + k) - ic;
    // corners of the whole tile (also covers lanes beyond the image edge)
    const double cx0 = (double)x0 - ic, cx1 = (double)(x0 + kTileX - 1) - ic;
    const double cy0 = (double)y0 - ic, cy1 = (double)(y0 + TY - 1) - ic;

    Acc<MODE> acc[PY][SB];

    for (int a0 = 0; a0 < N; a0 += A) {
        const int na = min(A, N - a0);
        __syncthreads();  // previous block's LDS fully consumed
        if (tid < na) {
            const float th = -a.angles[a0 + tid];
            const double c = (double)(float)cos((double)th);
            const double s = (double)(float)sin((double)th);
            const double f00 = det_pos(cx0 * c, cy0, s, cen);
            const double f01 = det_pos(cx1 * c, cy0, s, cen);
            const double f10 = det_pos(cx0 * c, cy1, s, cen);
            const double f11 = det_pos(cx1 * c, cy1, s, cen);
            const double fmn = fmin(fmin(f00, f01), fmin(f10, f11));
            const double fmx = fmax(fmax(f00, f01), fmax(f10, f11));
            s_cos[tid] = c;
            s_sin[tid] = s;
            s_dlo[tid] = (int)floor(fmn);
            s_inner[tid] = (fmn >= 0.0) && (fmx <= w1);
        }
        __syncthreads();
        const int nstage = na * SB * SPAN;
        for (int e = tid; e < nstage; e += kThreads) {
            const int k = e % SPAN;
            const int t = e / SPAN;
            const int js = t % SB;
            const int al = t / SB;
            const int j = j0 + js;
            const int col = s_dlo[al] + k;
            float yd = 0.0f, yu = 0.0f;
            if (j < S) {
                const Tin* row = sino + ((size_t)(a0 + al) * S + j) * C;
                if (col >= 0 && col < C) yd = to_f32(row[col]);
                if (col + 1 >= 0 && col + 1 < C) yu = to_f32(row[col + 1]);
            }
            s_pair[(al * SPAN + k) * SB + js] = make_float2(yd, yu - yd);
        }
        if (a.edge_pad) {
            for (int e = tid; e < na * SB; e += kThreads) {
                const int js = e % SB, al = e / SB, j = j0 + js;
                float e0 = 0.0f, e1 = 0.0f;
                if (j < S) {
                    const Tin* row = sino + ((size_t)(a0 + al) * S + j) * C;
                    e0 = to_f32(row[0]);
                    e1 = to_f32(row[C - 1]);
                }
                s_edge[2 * e] = e0;
                s_edge[2 * e + 1] = e1;
            }
        }
        __syncthreads();

        for (int al = 0; al < na; ++al) {
            const double c = s_cos[al], s = s_sin[al];
            const int dlo = s_dlo[al];
            const double xcc = xc * c;
            const float2* base = s_pair + al * SPAN * SB;
            if (s_inner[al]) {
#pragma unroll
                for (int k = 0; k < PY; ++k) {
                    const double f = det_pos(xcc, yc[k], s, cen);
                    const int d = (int)f;
                    const auto w = Weight<MODE>::make(f - (double)d);
                    const float2* q = base + (d - dlo) * SB;
#pragma unroll
                    for (int js = 0; js < SB; ++js) acc[k][js].add(q[js], w);
                }
            } else {
#pragma unroll
                for (int k = 0; k < PY; ++k) {
                    const double f = det_pos(xcc, yc[k], s, cen);
                    const bool in = (f >= 0.0) && (f <= w1);
                    const int d = (int)f;
                    const int idx = min(max(d - dlo, 0), SPAN - 1);
                    const auto w = Weight<MODE>::make(f - (double)d);
                    const float2* q = base + idx * SB;
                    if (in) {
#pragma unroll
                        for (int js = 0; js < SB; ++js)
                            acc[k][js].add(q[js], w);
                    } else if (a.edge_pad) {
                        const int side = f < 0.0 ? 0 : 1;
#pragma unroll
                        for (int js = 0; js < SB; ++js)
                            acc[k][js].add_value(
                                s_edge[2 * (al * SB + js) + side]);
                    }
                }
            }
        }
    }

    if (x < C) {
#pragma unroll
        for (int k = 0; k < PY; ++k) {
            const int y = y0 + wv * PY + k;
            if (y >= C) continue;
#pragma unroll
            for (int js = 0; js < SB; ++js) {
                const int j = j0 + js;
                if (j < S)
                    a.recon[((size_t)y * S + j) * C + x] =
                        finalize(a, acc[k][js].result(), x, y);
            }
        }
    }
}

// One thread per (x, y, slice), sinogram read straight from global memory.
template <int MODE>
__global__ void bp_naive_kernel(BpArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int j = blockIdx.z;
    const int C = a.ncol, S = a.nslice;
    if (x >= C || y >= C) return;
    const float* sino = static_cast<const float*>(a.sino);
    const double ic = a.icenter, w1 = (double)(C - 1);
    const double xc = (double)x - ic, yc = (double)y - ic;
    Acc<MODE> acc;
    for (int i = 0; i < a.nrow; ++i) {
        const float th = -a.angles[i];
        const double c = (double)(float)cos((double)th);
        const double s = (double)(float)sin((double)th);
        const double f = det_pos(xc * c, yc, s, a.center);
        const float* row = sino + ((size_t)i * S + j) * C;
        if (f >= 0.0 && f <= w1) {
            const int d = (int)f;
            const float yd = row[d];
            const float yu = d + 1 < C ? row[d + 1] : 0.0f;
            acc.add(make_float2(yd, yu - yd),
                    Weight<MODE>::make(f - (double)d));
        } else if (a.edge_pad) {
            acc.add_value(f < 0.0 ? row[0] : row[C - 1]);
        }
    }
    a.recon[((size_t)y * S + j) * C + x] = finalize(a, acc.result(), x, y);
}

template <int MODE, int SB, int PY, int A, typename Tin>
int launch_tiled(const BpArgs& a, hipStream_t stream) {
    constexpr int TY = 4 * PY;
    const int tiles_x = (a.ncol + kTileX - 1) / kTileX;
    const int tiles_y = (a.ncol + TY - 1) / TY;
    const long long groups = (a.nslice + SB - 1) / SB;
    const long long blocks = (long long)tiles_x * tiles_y * groups;
    if (blocks > 0x7fffffffLL)
        return set_error(ATH_ERR_UNSUPPORTED, "back-projection grid too large");
    hipLaunchKernelGGL((bp_tiled_kernel<MODE, SB, PY, A, Tin>),
                       dim3((unsigned)blocks), dim3(kThreads), 0, stream, a,
                       tiles_x, tiles_y);
    return check_hip(hipGetLastError(), "bp_tiled_kernel launch");
}

template <int MODE, typename Tin>
int pick_sb(const BpArgs& a, hipStream_t s) {
    constexpr int PY = 8, A = 16;
    if (a.nslice == 1) return launch_tiled<MODE, 1, PY, A, Tin>(a, s);
    if (a.nslice == 2) return launch_tiled<MODE, 2, PY, A, Tin>(a, s);
    return launch_tiled<MODE, 4, PY, A, Tin>(a, s);
}

}  // namespace

int back_projection(const BpArgs& a, int sino_dtype, int acc,
                    hipStream_t s) {
    if (sino_dtype == ATH_F32) {
        switch (acc) {
        case ATH_ACC_F64: return pick_sb<ATH_ACC_F64, float>(a, s);
        case ATH_ACC_F32SEQ: return pick_sb<ATH_ACC_F32SEQ, float>(a, s);
        case ATH_ACC_F32: return pick_sb<ATH_ACC_F32, float>(a, s);
        }
        return set_error(ATH_ERR_INVALID_ARG, "back-projection: bad acc mode");
    }
    if (acc != ATH_ACC_F32)
        return set_error(ATH_ERR_UNSUPPORTED,
                         "back-projection: float16/bfloat16 sinograms need "
                         "acc = ATH_ACC_F32");
    if (sino_dtype == ATH_F16) return pick_sb<ATH_ACC_F32, __half>(a, s);
    if (sino_dtype == ATH_BF16)
        return pick_sb<ATH_ACC_F32, __hip_bfloat16>(a, s);
    return set_error(ATH_ERR_INVALID_ARG, "back-projection: bad sinogram dtype");
}

int back_projection_naive(const BpArgs& a, int acc, hipStream_t s) {
    const dim3 block(16, 16);
    const dim3 grid((a.ncol + 15) / 16, (a.ncol + 15) / 16, a.nslice);
    switch (acc) {
    case ATH_ACC_F64:
        hipLaunchKernelGGL(bp_naive_kernel<ATH_ACC_F64>, grid, block, 0, s, a);
        break;
    case ATH_ACC_F32SEQ:
        hipLaunchKernelGGL(bp_naive_kernel<ATH_ACC_F32SEQ>, grid, block, 0, s,
                           a);
        break;
    case ATH_ACC_F32:
        hipLaunchKernelGGL(bp_naive_kernel<ATH_ACC_F32>, grid, block, 0, s, a);
        break;
    default:
        return set_error(ATH_ERR_INVALID_ARG, "back-projection: bad acc mode");
    }
    return check_hip(hipGetLastError(), "bp_naive_kernel launch");
}

}  // namespace ath
