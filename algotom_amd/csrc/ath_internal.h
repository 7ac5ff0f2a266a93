// Internal declarations shared by the HIP sources of libalgotom_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/algotom_hip.h"

namespace ath {

// ---- error state (thread-local; read through ath_last_error) -------------
int set_error(int status, const std::string& msg);
int check_hip(hipError_t err, const char* what);
void clear_error();

// ---- launch-state caches (thread-safe; keyed by kernel and device) --------
// Workgroups of `kernel` resident on the current device at once (occupancy
// per CU times the CU count).
int resident_grid(const void* kernel, int threads, size_t lds, int* out);
// hipFuncAttributeMaxDynamicSharedMemorySize = bytes for `kernel`, once per
// kernel and device with the largest size asked for.
int allow_dynamic_lds(const void* kernel, size_t bytes);

// ---- element conversions ---------------------------------------------------
__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ float to_f32(__half v) { return __half2float(v); }
__device__ __forceinline__ float to_f32(__hip_bfloat16 v) {
    return __bfloat162float(v);
}

__device__ __forceinline__ void store_elem(float* p, double v) {
    *p = (float)v;
}
__device__ __forceinline__ void store_elem(double* p, double v) { *p = v; }
__device__ __forceinline__ void store_elem(__half* p, double v) {
    *p = __float2half((float)v);
}
__device__ __forceinline__ void store_elem(__hip_bfloat16* p, double v) {
    *p = __float2bfloat16((float)v);
}

// numpy.pad index maps: i = padded index - pad; returns the source column,
// or -1 where the padded sample is zero (mode "constant").
__host__ __device__ __forceinline__ int pad_source(int i, int n, int mode) {
    if (i >= 0 && i < n) return i;
    switch (mode) {
    case ATH_PAD_EDGE: return i < 0 ? 0 : n - 1;
    case ATH_PAD_WRAP: { int j = i % n; return j < 0 ? j + n : j; }
    case ATH_PAD_REFLECT: {
        if (n == 1) return 0;
        const int p = 2 * (n - 1);
        int j = i % p; if (j < 0) j += p;
        return j < n ? j : p - j; }
    case ATH_PAD_SYMMETRIC: {
        const int p = 2 * n;
        int j = i % p; if (j < 0) j += p;
        return j < n ? j : p - 1 - j; }
    default: return -1;
    }
}

// Launch shape of a rows x width elementwise pass: grid.x spans the width in
// `threads`-wide blocks and grid.y strides over the rows, so splitting a row
// index into its parts is block-uniform (scalar unit, once per row) and no
// lane divides a 64-bit element index.  Kernels loop
// `for (r = blockIdx.y; r < rows; r += gridDim.y)` and
// `for (k = blockIdx.x * threads + threadIdx.x; k < width; k += gridDim.x * threads)`.
inline dim3 row_grid(long long rows, int width, int threads) {
    const long long gx = width > 0 ? (width + threads - 1) / threads : 1;
    long long gy = rows < 1 ? 1 : (rows < 65535 ? rows : 65535);
    if (gx * gy > (1LL << 22)) gy = std::max<long long>(1, (1LL << 22) / gx);
    return dim3((unsigned)std::min<long long>(gx, 1LL << 16), (unsigned)gy);
}

// ---- ramp filter (ramp_filter.hip) ----------------------------------------
int fft_length(int ncol, int pad);
size_t plan_bytes(int fft_len, int nwin);
int plan_build(const double* taps, int nwin, int width, int fft_len,
               void* plan, hipStream_t stream);
int ramp_filter(const void* sino, int in_dtype, int nrow, int nslice,
                int ncol, int pad, int pad_mode, int apply_log,
                const void* plan, int fft_len, int nwin, void* out,
                int out_dtype, hipStream_t stream);

// ---- back-projection (backproject.hip) -------------------------------------
struct BpArgs {
    const void* sino;      // (nrow, nslice, ncol)
    float* recon;          // (ncol, nslice, ncol)
    float* recon2;         // second output of ATH_ACC_F64_DERIV, else null
    const float* angles;   // nrow float32 radians
    int nrow, nslice, ncol;
    double center;         // float32 centre widened to double
    double icenter;        // 0.5 * (ncol - 1)
    int edge_pad;
    int use_mask;
    double mask_r2;        // (ratio * (ncol // 2))^2
    int mask_c;            // ncol // 2
    int apply_scale;
    float pi32, denom;     // float32(pi), float32(nrow - 1)
};

// sb / py <= 0 select the default tile configuration
int back_projection(const BpArgs& a, int sino_dtype, int acc, int sb, int py,
                    hipStream_t stream);
int back_projection_naive(const BpArgs& a, int acc, hipStream_t stream);

// ---- forward projection = adjoint of the back-projection ------------------
// largest FFT length of the 1D ramp filter and of its adjoint
constexpr int kMaxFilterFft = 32768;
constexpr int kMaxFilterAdjointFft = 32768;

struct FpArgs {
    const float* image;    // (ncol, nslice, ncol) = (y, slice, x)
    const float* image_t;  // nullptr, or the same image as (x, slice, y)
    float* sino;           // (nrow, nslice, ncol)
    const float* angles;
    int nrow, nslice, ncol;
    double center, icenter;
    int use_mask;
    double mask_r2;
    int mask_c;
    double scale;          // multiplies every output
};
int forward_projection(const FpArgs& a, hipStream_t stream);
// out(x, j, y) = in(y, j, x) for the (ncol, nslice, ncol) image
int transpose_slices(const float* in, float* out, int nslice, int ncol,
                     hipStream_t stream);
int gaussian_rows(const float* in, float* out, int nrows, int ncol,
                  const double* weights, int radius, hipStream_t s);
size_t circular_plan_bytes(int L);
size_t circular_plan_work_bytes(int width, int L);
size_t circular_filter_work_bytes(int ncol, int pad, int L, int nimg);
int circular_plan_build(const double* g2, int width, int L, void* plan,
                        void* work, hipStream_t s);
int circular_ramp_filter(const float* images, int nimg, int ncol,
                         long long row_stride, long long img_stride, int pad,
                         int pad_mode, const void* plan, int L, void* work,
                         void* out, int out_dtype, long long out_row_stride,
                         long long out_img_stride, double mask_radius,
                         int scale_nrow, hipStream_t s);
size_t dfi_plan_bytes(int width);
int dfi_plan_build(const double* window, int width, void* plan,
                   hipStream_t s);
size_t dfi_workspace_bytes(int nrow_in, int ncol, int nrow, int width);
int dfi_reconstruction(const void* sino, int dtype, int nrow_in, int nslice,
                       int ncol, long long row_stride, long long slice_stride,
                       double xshift, int pad, int pad_mode, int apply_log,
                       int angle_mode, double row_shift, int flip, int nrow,
                       const void* plan, int width, void* workspace,
                       void* out, int out_dtype, long long out_row_stride,
                       long long out_slice_stride, double mask_r2, int* flag,
                       hipStream_t s);
constexpr int kVertGpu = ATH_VERT_GPU;
constexpr int kVertCpu = ATH_VERT_CPU;
int vertical_back_projection(const float* proj, int nproj, int height,
                             int width, const double* angles,
                             const double* xs, const double* ys, int nslice,
                             double center, const double* centers,
                             int edge_pad, int mode, int chunk, int accumulate,
                             float* out, hipStream_t s);
int flat_field_log(const void* proj, int in_dtype, int nproj, int height,
                   int width, long long proj_stride, long long row_stride,
                   const float* dark, const float* flat_dark, int apply_log,
                   float* out, int* flag, hipStream_t s);
int vertical_finish(void* data, int dtype, long long total, int width, int lo,
                    int hi, int nrow, hipStream_t s);
int extend_sinogram(const float* sino, int nrow, int nslice, int ncol,
                    int overlap, int side, int apply_log, float* out,
                    hipStream_t s);
int flat_field_correction(const void* proj, int in_dtype, long long nouter,
                          int height, int width, long long outer_stride,
                          long long row_stride, const void* dark,
                          const void* denom, int field_dtype, void* out,
                          hipStream_t s);
int minus_log(const float* in, float* out, long long n, hipStream_t s);
int zoom_linear(const float* in, int h, int w, float* out, int oh, int ow,
                hipStream_t s);
size_t slice_metrics_workspace(int nimg, int nrows, int ncols, int metric);
int slice_metrics(const void* images, int dtype, int nimg, int nrows,
                  int ncols, long long stride, double nmin, double nmax,
                  int metric, double* out, void* workspace, hipStream_t s);
int gaussian_axis(const float* in, float* out, int nouter, int length,
                  int ninner, const double* weights, int radius,
                  hipStream_t s);
// scipy.ndimage.shift of float32 rows along the columns (order 3, mode
// "nearest"), optionally flipped left-right and -logged first
int shift_rows(const float* src, long long nrows, int ncol, int flip,
               int apply_log, double sh, float* dst, hipStream_t s);
// stitch_image of image 1 (img1) and image 2 (img2, flipped at read when
// flip2), nslice interleaved slices of nrow_180 rows each
int stitch_images(const float* img1, const float* img2, int flip2,
                  int nrow_180, int nslice, int ncol, int overlap, int side,
                  int norm, int width, const void* w1, const void* w2,
                  int wdtype, float* factors, float* out, hipStream_t s);
int ramp_filter_adjoint(const float* in, int nrow, int nslice, int ncol,
                        int pad, int pad_mode, const void* plan, int fft_len,
                        int nwin, float* out, hipStream_t stream);
int back_projection_batched(const ath_bp_job* jobs, int njobs,
                            long long total, int sino_dtype, int acc, int sb,
                            int py, int nblocks, int* counter,
                            hipStream_t stream);
int ramp_filter_batched(const ath_filter_job* jobs, int njobs,
                        long long total_rows, int in_dtype, int out_dtype,
                        int fft_len, int pad_mode, int apply_log,
                        hipStream_t stream);

}  // namespace ath
