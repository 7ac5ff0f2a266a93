// Vertical-slice reconstruction on gfx950 (algotom/rec/vertrec.py).
//
// A vertical slice is the set of points (x_j, y_j), j < width, on a line
// through the sample, reconstructed for every detector row i: recon[i][j] =
// sum over projections n of the projection row proj[n][i] interpolated at
// f = x_j cos(theta_n) + y_j sin(theta_n) + center.  The position is shared
// by all rows, so a workgroup takes 64 points (one per lane) x 64 rows and
// every lane computes f once per projection for its 16 rows; the two taps of
// a row are gathered straight from the projection row (consecutive lanes
// read consecutive columns, so the loads coalesce).
//
// The reference sums projections in chunks (vertrec.py:584-620): every
// chunk of `chunk` projections is back-projected into a fresh float32 image
// that is then added to the running float32 slice.  Two typings are
// reproduced (numba's, from the kernels' argument types):
//   kVertGpu  __vertical_back_projection_gpu(_chunk)_kernel (:176-213,
//             :256-293): float32 angles, points, centre and trig, float32
//             position r = x c + y s, f = r + center; float64 chunk sum.
//   kVertCpu  vertical_back_projection_cpu(_chunk) (:53-110, :113-173):
//             float64 angles, trig and position (points as given); the
//             chunk image is float32, stored after every projection.
// In both, val = yd + (yu - yd) (f - d) with yu - yd in float32 and the
// rest in float64; outside [0, width-1] the edge value is added only with
// edge_pad.
#include "ath_internal.h"
#include "log32.h"

namespace ath {
namespace {

constexpr int kThreads = 256;
constexpr int kPts = 64;       // points per workgroup, one per lane
// rows per lane; 4 waves -> 64 rows (1801 x 1024 x 2560, 16 lines:
// 4 / 8 / 16 rows per lane 152.2 / 144.3 / 141.7 ms)
constexpr int kRows = 16;
constexpr int kAngBlock = 64;  // projections whose trig is staged at once

struct VertArgs {
    const float* proj;    // (nproj, height, width)
    const double* angles; // nproj radians
    const double* xs;     // (nslice, width)
    const double* ys;
    const double* centers;  // per-line centres, or nullptr
    float* out;           // (nslice, height, width)
    int nproj, height, width, nslice;
    double center;
    int edge_pad, chunk, accumulate;
};

// Workgroup order (1D grid): id = (q * nslice + line) * 8 + x8 with tile
// t = 8q + x8.  Workgroups are dealt to the 8 XCDs round-robin by id, so
// every line of one tile lands on the same XCD (id % 8 = x8) and is
// dispatched within 8 * nslice ids of the others: the lines, which read
// overlapping columns of the same projection rows, share that XCD's L2
// instead of each streaming the stack from HBM.
template <int MODE>
__global__ __launch_bounds__(kThreads) void vertical_bp_kernel(VertArgs a,
                                                               int tiles_x,
                                                               int tiles) {
    __shared__ double s_c[kAngBlock], s_s[kAngBlock];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const unsigned id = blockIdx.x;
    const int x8 = id & 7;
    const unsigned rest = id >> 3;
    const int sl = (int)(rest % (unsigned)a.nslice);
    const int tile = (int)(rest / (unsigned)a.nslice) * 8 + x8;
    if (tile >= tiles) return;  // padding of the last group of 8 tiles
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int j = tx * kPts + lane;
    const int i0 = ty * (4 * kRows) + wave;
    const bool valid_j = j < a.width;
    const int width1 = a.width - 1;
    double x = 0.0, y = 0.0;
    if (valid_j) {
        x = a.xs[(size_t)sl * a.width + j];
        y = a.ys[(size_t)sl * a.width + j];
    }
    const double cen = a.centers ? a.centers[sl] : a.center;
    const float x32 = (float)x, y32 = (float)y, cen32 = (float)cen;
    double chunk64[kRows];
    float chunk32[kRows], total[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        chunk64[k] = 0.0;
        chunk32[k] = 0.0f;
        total[k] = 0.0f;
    }
    const size_t plane = (size_t)a.height * a.width;
    for (int n0 = 0; n0 < a.nproj; n0 += kAngBlock) {
        __syncthreads();
        if (tid < kAngBlock && n0 + tid < a.nproj) {
            const double th = a.angles[n0 + tid];
            if (MODE == kVertGpu) {
                const double t = (double)(float)th;
                s_c[tid] = (double)(float)cos(t);
                s_s[tid] = (double)(float)sin(t);
            } else {
                s_c[tid] = cos(th);
                s_s[tid] = sin(th);
            }
        }
        __syncthreads();
        const int nb = min(kAngBlock, a.nproj - n0);
        for (int q = 0; q < nb; ++q) {
            const int n = n0 + q;
            double fpos, frac;
            int d, u;
            bool in_range;
            if (MODE == kVertGpu) {
                const float c = (float)s_c[q], s = (float)s_s[q];
                const float p1 = x32 * c, p2 = y32 * s;
                const float f = (p1 + p2) + cen32;
                in_range = (f >= 0.0f) && ((double)f <= (double)width1);
                d = (int)floorf(f);
                u = (int)ceilf(f);
                fpos = (double)f;
                frac = (double)f - (double)d;
            } else {
                const double f = (x * s_c[q] + y * s_s[q]) + cen;
                in_range = (f >= 0.0) && (f <= (double)width1);
                d = (int)floor(f);
                u = (int)ceil(f);
                fpos = f;
                frac = f - (double)d;
            }
            // Both taps of every row are loaded unconditionally, so the
            // 2 * kRows gathers of a projection are in flight together.
            // u == d gives yu == yd and val = yd exactly; outside the range
            // both indices clamp to the edge column, which is the edge_pad
            // value.  Every index lies in [0, width - 1].
            if (!in_range) d = u = (fpos < 0.0) ? 0 : width1;
            const bool use = in_range || a.edge_pad;
            const float* pn = a.proj + (size_t)n * plane + (size_t)i0 * a.width;
            float yd[kRows], yu[kRows];
#pragma unroll
            for (int k = 0; k < kRows; ++k) {
                const bool ok = valid_j && i0 + 4 * k < a.height;
                const float* row = pn + (size_t)(4 * k) * a.width;
                yd[k] = ok ? row[d] : 0.0f;
                yu[k] = ok ? row[u] : 0.0f;
            }
            if (use) {
#pragma unroll
                for (int k = 0; k < kRows; ++k) {
                    const double val =
                        (double)yd[k] + (double)(yu[k] - yd[k]) * frac;
                    if (MODE == kVertGpu)
                        chunk64[k] += val;
                    else
                        chunk32[k] = (float)((double)chunk32[k] + val);
                }
            }
            if ((n + 1) % a.chunk == 0 || n + 1 == a.nproj) {
#pragma unroll
                for (int k = 0; k < kRows; ++k) {
                    if (MODE == kVertGpu) {
                        total[k] = total[k] + (float)chunk64[k];
                        chunk64[k] = 0.0;
                    } else {
                        total[k] = total[k] + chunk32[k];
                        chunk32[k] = 0.0f;
                    }
                }
            }
        }
    }
    if (!valid_j) return;
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const int i = i0 + 4 * k;
        if (i >= a.height) continue;
        float* o = a.out + ((size_t)sl * a.height + i) * a.width + j;
        *o = a.accumulate ? *o + total[k] : total[k];
    }
}

// (img - dark) / flat_dark in float32 (the caller has replaced zeros of
// flat_dark by 1), then -log with values <= 0 set to 1 and flagged.
__device__ __forceinline__ float flat_log_value(float v, bool fd, float dk,
                                                float fdk, int apply_log,
                                                int* flag,
                                                const Log32Table& tab) {
    if (fd) v = (v - dk) / fdk;
    if (apply_log) {
        if (v <= 0.0f) {
            v = 1.0f;
            *flag = 1;
        }
        v = neg_log_f32(v, tab);
    }
    return v;
}

template <typename Tin, int VEC> struct VecOf;
template <> struct VecOf<float, 1> { using type = float; };
template <> struct VecOf<unsigned short, 1> { using type = unsigned short; };
template <> struct VecOf<float, 4> { using type = float4; };
template <> struct VecOf<unsigned short, 4> { using type = ushort4; };

template <typename V>
__device__ __forceinline__ void unpack(const V& x, float (&v)[4]) {
    v[0] = (float)x.x;
    v[1] = (float)x.y;
    v[2] = (float)x.z;
    v[3] = (float)x.w;
}
__device__ __forceinline__ void unpack(float x, float (&v)[1]) { v[0] = x; }
__device__ __forceinline__ void unpack(unsigned short x, float (&v)[1]) {
    v[0] = (float)x;
}

// Projections per pass over one detector row: the dark and flat-dark values
// of a column group are loaded once for all of them.
constexpr int kProjRun = 8;

// VEC = 4: groups of four columns with 16-byte (8 for uint16) loads and
// stores; the caller checks widths, strides and alignment.  Work items are
// (run of kProjRun projections, detector row) pairs.
template <typename Tin, int VEC>
__global__ __launch_bounds__(kThreads) void flat_field_log_kernel(
    const Tin* __restrict__ proj, long long nproj, int height, int width,
    long long proj_stride, long long row_stride,
    const float* __restrict__ dark, const float* __restrict__ flat_dark,
    int apply_log, float* __restrict__ out, int* __restrict__ flag) {
    __shared__ Log32Table s_tab;
    stage_log32_table<kThreads>(&s_tab);
    __syncthreads();
    using V = typename VecOf<Tin, VEC>::type;
    using FV = typename VecOf<float, VEC>::type;
    const int groups = width / VEC;
    const long long items = (nproj + kProjRun - 1) / kProjRun * height;
    const bool fd = dark != nullptr;
    for (long long q = blockIdx.y; q < items; q += gridDim.y) {
        const long long run = q / height;  // block-uniform
        const int iy = (int)(q - run * height);
        const long long n0 = run * kProjRun;
        const int np = (int)min((long long)kProjRun, nproj - n0);
        for (int g = blockIdx.x * kThreads + threadIdx.x; g < groups;
             g += gridDim.x * kThreads) {
            float dk[VEC], fk[VEC];
            if (fd) {
                unpack(reinterpret_cast<const FV*>(dark + (size_t)iy * width)[g],
                       dk);
                unpack(reinterpret_cast<const FV*>(flat_dark +
                                                   (size_t)iy * width)[g],
                       fk);
            } else {
#pragma unroll
                for (int k = 0; k < VEC; ++k) dk[k] = fk[k] = 0.0f;
            }
            for (int p = 0; p < np; ++p) {
                const long long n = n0 + p;
                const V x = reinterpret_cast<const V*>(
                    proj + n * proj_stride + iy * row_stride)[g];
                float v[VEC];
                unpack(x, v);
#pragma unroll
                for (int k = 0; k < VEC; ++k)
                    v[k] = flat_log_value(v[k], fd, dk[k], fk[k], apply_log,
                                          flag, s_tab);
                FV y;
                if constexpr (VEC == 1) {
                    y = v[0];
                } else {
                    y = make_float4(v[0], v[1], v[2], v[3]);
                }
                reinterpret_cast<FV*>(out + (n * height + iy) * width)[g] = y;
            }
        }
    }
}

// In place: columns outside [lo, hi) zeroed (lo < 0: no mask), then the
// factor pi / (nrow - 1) (nrow < 2: none), in the array's precision with
// numpy's order (x * mask, * pi, / (nrow - 1)).
template <typename T>
__global__ __launch_bounds__(kThreads) void vertical_finish_kernel(
    T* __restrict__ data, long long rows, int width, int lo, int hi,
    int nrow) {
    for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
        T* row = data + r * width;
        for (int jx = blockIdx.x * kThreads + threadIdx.x; jx < width;
             jx += gridDim.x * kThreads) {
            T v = row[jx];
            if (lo >= 0) v = v * ((jx >= lo && jx < hi) ? T(1) : T(0));
            if (nrow >= 2) {
                if constexpr (sizeof(T) == 8) {
                    v = v * 3.141592653589793;
                    v = v / (double)(nrow - 1);
                } else {
                    v = v * (float)3.141592653589793;
                    v = v / (float)(nrow - 1);
                }
            }
            row[jx] = v;
        }
    }
}

}  // namespace

int vertical_back_projection(const float* proj, int nproj, int height,
                             int width, const double* angles,
                             const double* xs, const double* ys, int nslice,
                             double center, const double* centers,
                             int edge_pad, int mode, int chunk, int accumulate,
                             float* out, hipStream_t s) {
    VertArgs a{proj, angles, xs, ys, centers, out, nproj, height, width,
               nslice, center, edge_pad, chunk, accumulate};
    const int tiles_x = (width + kPts - 1) / kPts;
    const int tiles = tiles_x * ((height + 4 * kRows - 1) / (4 * kRows));
    const long long blocks = (long long)((tiles + 7) / 8) * 8 * nslice;
    if (blocks > 0x7fffffffLL)
        return set_error(ATH_ERR_UNSUPPORTED, "vertical grid too large");
    const dim3 grid((unsigned)blocks);
    if (mode == kVertGpu)
        hipLaunchKernelGGL(vertical_bp_kernel<kVertGpu>, grid, dim3(kThreads),
                           0, s, a, tiles_x, tiles);
    else
        hipLaunchKernelGGL(vertical_bp_kernel<kVertCpu>, grid, dim3(kThreads),
                           0, s, a, tiles_x, tiles);
    return check_hip(hipGetLastError(), "vertical_bp_kernel launch");
}

template <typename Tin, int VEC>
static void launch_flat_log(const void* proj, int nproj, int height,
                            int width, long long proj_stride,
                            long long row_stride, const float* dark,
                            const float* flat_dark, int apply_log, float* out,
                            int* flag, hipStream_t s) {
    const dim3 grid = row_grid(
        (long long)(nproj + kProjRun - 1) / kProjRun * height, width / VEC,
        kThreads);
    hipLaunchKernelGGL((flat_field_log_kernel<Tin, VEC>), grid,
                       dim3(kThreads), 0, s, static_cast<const Tin*>(proj),
                       (long long)nproj, height, width, proj_stride,
                       row_stride, dark, flat_dark, apply_log, out, flag);
}

int flat_field_log(const void* proj, int in_dtype, int nproj, int height,
                   int width, long long proj_stride, long long row_stride,
                   const float* dark, const float* flat_dark, int apply_log,
                   float* out, int* flag, hipStream_t s) {
    const bool u16 = in_dtype == ATH_U16;
    const size_t in_align = u16 ? 8 : 16;
    const bool vec =
        width % 4 == 0 && row_stride % 4 == 0 && proj_stride % 4 == 0 &&
        reinterpret_cast<uintptr_t>(proj) % in_align == 0 &&
        reinterpret_cast<uintptr_t>(out) % 16 == 0 &&
        (!dark || (reinterpret_cast<uintptr_t>(dark) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(flat_dark) % 16 == 0));
    if (u16) {
        if (vec)
            launch_flat_log<unsigned short, 4>(proj, nproj, height, width,
                                               proj_stride, row_stride, dark,
                                               flat_dark, apply_log, out, flag,
                                               s);
        else
            launch_flat_log<unsigned short, 1>(proj, nproj, height, width,
                                               proj_stride, row_stride, dark,
                                               flat_dark, apply_log, out, flag,
                                               s);
    } else if (vec) {
        launch_flat_log<float, 4>(proj, nproj, height, width, proj_stride,
                                  row_stride, dark, flat_dark, apply_log, out,
                                  flag, s);
    } else {
        launch_flat_log<float, 1>(proj, nproj, height, width, proj_stride,
                                  row_stride, dark, flat_dark, apply_log, out,
                                  flag, s);
    }
    return check_hip(hipGetLastError(), "flat_field_log_kernel launch");
}

int vertical_finish(void* data, int dtype, long long total, int width, int lo,
                    int hi, int nrow, hipStream_t s) {
    if (width <= 0) return ATH_OK;
    const long long rows = total / width;
    const dim3 grid = row_grid(rows, width, kThreads);
    if (dtype == ATH_F64)
        hipLaunchKernelGGL(vertical_finish_kernel<double>, grid,
                           dim3(kThreads), 0, s, static_cast<double*>(data),
                           rows, width, lo, hi, nrow);
    else
        hipLaunchKernelGGL(vertical_finish_kernel<float>, grid, dim3(kThreads),
                           0, s, static_cast<float*>(data), rows, width, lo,
                           hi, nrow);
    return check_hip(hipGetLastError(), "vertical_finish_kernel launch");
}

}  // namespace ath
