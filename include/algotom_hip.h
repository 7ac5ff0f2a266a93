/*
 * algotom_hip.h -- C ABI of the MI355X (gfx950) filtered back-projection path.
 *
 * Every entry point returns an int status (ATH_OK = 0, negative on error);
 * the message of the last error on the calling thread is read with
 * ath_last_error().  Nothing is thrown across the ABI.  All array arguments
 * are DEVICE pointers owned by the caller; the library allocates nothing and
 * copies nothing between host and device.  Work is enqueued on `stream`
 * (a hipStream_t, NULL = the default stream) and is asynchronous.
 *
 * Layouts follow algotom's own: a sinogram stack is (nrow, nslice, ncol) =
 * (angle, slice, detector column), C-contiguous; a 2D sinogram is the
 * nslice = 1 case.  Reconstructions are (ncol, nslice, ncol) = (y, slice, x),
 * i.e. algotom's recon[y, j, x] (reconstruction.py:477-481).
 */
#ifndef ALGOTOM_HIP_H
#define ALGOTOM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ATH_OK 0
#define ATH_ERR_INVALID_ARG (-1)
#define ATH_ERR_HIP (-2)
#define ATH_ERR_UNSUPPORTED (-3)
#define ATH_ERR_NO_DEVICE (-4)
#define ATH_ERR_WORKSPACE (-5)

/* element types of sinogram / filtered-sinogram buffers */
#define ATH_F32 0
#define ATH_F64 1
#define ATH_F16 2
#define ATH_BF16 3
#define ATH_U16 4   /* raw projections (ath_flat_field_log input only) */

/* numpy.pad modes applied inside the filter kernel */
#define ATH_PAD_EDGE 0
#define ATH_PAD_CONSTANT 1
#define ATH_PAD_REFLECT 2
#define ATH_PAD_SYMMETRIC 3
#define ATH_PAD_WRAP 4
/* numpy.pad's value modes with their defaults (end_values=0,
 * stat_length=None: the statistic of the whole row, the same value on both
 * sides), computed per row of the (logged) input by the filter kernel.
 * Accepted by ath_ramp_filter, ath_ramp_filter_batched and
 * ath_fbp_reconstruction only.  maximum / minimum / median are numpy's
 * exactly, linear_ramp as numpy 2's linspace (float64, rounded once to the
 * row's type); the mean's sum is float64 (numpy: pairwise in the row's
 * type), so a float32 mean may differ from numpy's in the last bit. */
#define ATH_PAD_LINEAR_RAMP 5
#define ATH_PAD_MAXIMUM 6
#define ATH_PAD_MEAN 7
#define ATH_PAD_MEDIAN 8
#define ATH_PAD_MINIMUM 9

/* back-projection accumulation (which reference kernel's arithmetic):
 *  ATH_ACC_F64    float64 register sum, one rounding at the end
 *                 (__back_projection_gpu_kernel, reconstruction.py:168,182)
 *  ATH_ACC_F32SEQ float32 storage after every angle, float64 interpolation
 *                 (__back_projection_gpu_chunk_kernel :256 and
 *                  back_projection_cpu :351)
 *  ATH_ACC_F32    float32 interpolation (one FMA) and float32 sum; detector
 *                 position still float64 (throughput mode, not bit-exact) */
#define ATH_ACC_F64 0
#define ATH_ACC_F32SEQ 1
#define ATH_ACC_F32 2
/*  ATH_ACC_DERIV  sum over angles of s[d+1] - s[d] at d = floor f: the
 *                 derivative of the back-projection with respect to the
 *                 rotation centre (float64 sum; no reference counterpart) */
#define ATH_ACC_DERIV 3
/*  ATH_ACC_F64_DERIV  ATH_ACC_F64 and ATH_ACC_DERIV in one pass (two
 *                 outputs, ath_back_projection_with_derivative only) */
#define ATH_ACC_F64_DERIV 4

/* One stack of a batched launch (ath_ramp_filter_batched): rows of
 * `sino` (nrow, nslice, ncol) are filtered into `out`; `plan` comes from
 * ath_filter_plan_build for width ncol + 2*pad and the batch's fft_len.
 * row_begin = sum of nrow*nslice over the preceding jobs. */
typedef struct {
    const void* sino;
    void* out;
    const void* plan;
    int nrow, nslice, ncol, pad, nwin, reserved;
    long long row_begin;
} ath_filter_job;

/* One stack of a batched back-projection (ath_back_projection_batched).
 * tile_begin = sum of ath_bp_job_tiles(...) over the preceding jobs. */
typedef struct {
    const void* sino;      /* filtered (nrow, nslice, ncol) */
    float* recon;          /* (ncol, nslice, ncol) */
    const float* angles;   /* nrow float32 radians */
    int nrow, nslice, ncol;
    float center;
    double mask_radius;    /* < 0: no mask */
    int apply_scale, edge_pad;
    long long tile_begin;
} ath_bp_job;

/* Library version (major*10000 + minor*100 + patch). No reference
 * counterpart. */
int ath_version(void);

/* Copies the calling thread's last error message (NUL-terminated, truncated
 * to len) into buf; returns the status code of that error (ATH_OK if none).
 * No reference counterpart: replaces Python exceptions raised by numba. */
int ath_last_error(char* buf, size_t len);

/* Number of visible HIP devices. Replaces numba's cuda.is_available()
 * (reconstruction.py:420). */
int ath_device_count(int* count);

/* FFT length used by the ramp filter for ncol columns and `pad` padding:
 * the smallest power of two L >= max(2*(ncol+2*pad) - pad, 64).  The filter
 * evaluates the reference's length-W circular convolution exactly as a
 * linear convolution with the doubled kernel [g, g] on this length.  The
 * filter runs up to L = 32768 (ncol <= 16159 at pad 150);
 * ATH_ERR_UNSUPPORTED above.  ath_ramp_filter_adjoint (with pad <= 700) and
 * the circular (BPF) filter run up to 32768 too. */
int ath_filter_fft_length(int ncol, int pad);

/* Device bytes of a filter plan holding `nwin` kernel spectra. */
size_t ath_filter_plan_bytes(int fft_len, int nwin);

/* Builds a filter plan on the device: twiddles and, for each of `nwin`
 * rows of `taps` (nwin x width float64, device), the half spectrum of the
 * doubled real kernel.  taps[r] = Re(ifftshift(ifft(ifftshift(R_r)))) for
 * row r of algotom's ramp window R (make_2d_ramp_window,
 * reconstruction.py:86-115); ath_filter_taps_from_window computes it.
 * Replaces the per-call window product of apply_ramp_filter
 * (reconstruction.py:148-152). */
int ath_filter_plan_build(const double* taps, int nwin, int width,
                          int fft_len, void* plan, size_t plan_bytes,
                          void* stream);

/* HOST helper: real convolution taps of a ramp-window row (host arrays).
 * win_re/win_im: width values of one row of make_2d_ramp_window (in its
 * fftshift order); taps_out: width doubles.  Long-double DFT. */
int ath_filter_taps_from_window(const double* win_re, const double* win_im,
                                int width, double* taps_out);

/* Ramp filter of every row of a (nrow, nslice, ncol) stack:
 * optional -log (apply_log, reconstruction.py:431-436; values must already
 * be > 0), numpy.pad by `pad` columns (pad_mode), windowed ramp filter
 * (circular convolution of width ncol + 2*pad), real part, crop
 * [pad, pad + ncol).  Replaces apply_ramp_filter (reconstruction.py:118-153)
 * and its 3D fan-out (:460-475).  in_dtype: ATH_F32 or ATH_F64;
 * out_dtype: any ATH_* element type.  nwin is 1 (one window for all rows)
 * or nrow (one window per angle row). */
int ath_ramp_filter(const void* sino, int in_dtype, int nrow, int nslice,
                    int ncol, int pad, int pad_mode, int apply_log,
                    const void* plan, int fft_len, int nwin, void* out,
                    int out_dtype, void* stream);

/* Linear-interpolation back-projection of a (nrow, nslice, ncol) filtered
 * stack into recon (ncol, nslice, ncol) float32.  angles: nrow float32
 * radians (device); center: rotation centre (rounded to float32 as the
 * reference does).  edge_pad: out-of-range positions add the edge sample
 * (reconstruction.py:183-189).  mask_radius >= 0 multiplies by
 * make_circle_mask (utility.py:160-182) with that radius; apply_scale
 * multiplies by pi/(nrow-1) in float32 (reconstruction.py:493).
 * Replaces back_projection_gpu (:193-226, acc ATH_ACC_F64),
 * back_projection_gpu_chunk (:267-304) and back_projection_cpu (:307-359)
 * (acc ATH_ACC_F32SEQ).  sino_dtype: ATH_F32, or ATH_F16 / ATH_BF16 with
 * acc ATH_ACC_F32 (reduced-precision storage, float32 accumulation). */
int ath_back_projection(const void* sino, int sino_dtype, int nrow,
                        int nslice, int ncol, const float* angles,
                        float center, int edge_pad, int acc,
                        double mask_radius, int apply_scale, float* recon,
                        void* stream);

/* The back-projection of ath_back_projection with acc ATH_ACC_F64 into
 * recon and, in the same pass over the same positions and staged samples,
 * its derivative with respect to the rotation centre (acc ATH_ACC_DERIV)
 * into deriv (ncol, nslice, ncol) float32, both with the mask and scale;
 * each output is bit-identical to its separate call.  sino: float32.  No
 * reference counterpart: the forward pass of differentiable FBP when the
 * centre needs a gradient, saving the derivative's own pass. */
int ath_back_projection_with_derivative(const float* sino, int nrow,
                                        int nslice, int ncol,
                                        const float* angles, float center,
                                        int edge_pad, double mask_radius,
                                        int apply_scale, float* recon,
                                        float* deriv, void* stream);

/* ath_back_projection with an explicit tile configuration: slice_batch
 * slices per workgroup (1, 2, 4; 8 for float16/bfloat16) and tile_rows
 * image rows per 64-column tile (16, 32, 64).  Used by the autotuner;
 * returns ATH_ERR_UNSUPPORTED for a configuration that is not compiled.
 * No reference counterpart (numba's block=(16,16) argument,
 * reconstruction.py:193-194, is the nearest). */
int ath_back_projection_tiled(const void* sino, int sino_dtype, int nrow,
                              int nslice, int ncol, const float* angles,
                              float center, int edge_pad, int acc,
                              double mask_radius, int apply_scale,
                              float* recon, int slice_batch, int tile_rows,
                              void* stream);

/* Number of work items (tiles) one stack contributes to a batched
 * back-projection with the given tile configuration. */
long long ath_bp_job_tiles(int ncol, int nslice, int slice_batch,
                           int tile_rows);

/* Back-projects many stacks of different widths, angle counts and centres
 * in one persistent launch: `nblocks` workgroups (0 = as many as fit on the
 * device at once) take (stack, tile) items from a device work counter until
 * all total_tiles items are done.  jobs: device array of njobs descriptors,
 * ordered by decreasing per-tile cost (nrow) for load balance.  counter:
 * one device int, reset by a kernel inside this call (no memset, so the
 * call can be captured in a HIP graph).  Each stack gets exactly the
 * arithmetic of ath_back_projection.  No reference counterpart: the
 * reference loops over slices in Python (reconstruction.py:460-490) and over
 * datasets in SLURM jobs. */
int ath_back_projection_batched(const ath_bp_job* jobs, int njobs,
                                long long total_tiles, int sino_dtype,
                                int acc, int slice_batch, int tile_rows,
                                int nblocks, int* counter, void* stream);

/* Ramp filter of many stacks of different widths sharing one FFT length,
 * one launch: a workgroup per row, its stack found from row_begin.
 * total_rows = sum of nrow*nslice.  Same arithmetic as ath_ramp_filter. */
int ath_ramp_filter_batched(const ath_filter_job* jobs, int njobs,
                            long long total_rows, int in_dtype,
                            int out_dtype, int fft_len, int pad_mode,
                            int apply_log, void* stream);

/* Same contract, one thread per output pixel reading global memory directly
 * (the reference kernels' structure).  Kept as an independent second
 * implementation for cross-checks; float32 sinograms only. */
int ath_back_projection_naive(const float* sino, int nrow, int nslice,
                              int ncol, const float* angles, float center,
                              int edge_pad, int acc, double mask_radius,
                              int apply_scale, float* recon, void* stream);

/* Adjoint (transpose) of ath_back_projection: a forward projection with
 * the back-projector's own linear-interpolation weights.  image:
 * (ncol, nslice, ncol) float32; sino: (nrow, nslice, ncol) float32 output,
 * sino = scale * B^T (mask * image) with B the back-projection of
 * reconstruction.py:156-190 and mask as in ath_back_projection
 * (mask_radius < 0: none).  Deterministic gather, float64 sums.  No
 * reference counterpart (the backward pass of differentiable FBP). */
int ath_forward_projection(const float* image, int nrow, int nslice,
                           int ncol, const float* angles, float center,
                           double mask_radius, double scale, float* sino,
                           void* stream);

/* Bytes of the workspace ath_forward_projection_ws needs: one float32 copy
 * of the (ncol, nslice, ncol) image; 0 for an invalid geometry. */
size_t ath_forward_projection_workspace_bytes(int nslice, int ncol);

/* ath_forward_projection with a caller workspace of at least
 * ath_forward_projection_workspace_bytes(nslice, ncol): the image is first
 * transposed per slice into it, so the angles whose rays run closer to the
 * y axis stage contiguous memory too.  Bit-identical to
 * ath_forward_projection; faster where the image does not fit the L2 and
 * last-level caches (1.26x at 16 x 7192^2).  ATH_ERR_WORKSPACE when the
 * workspace is missing or too small.  No reference counterpart (as
 * ath_forward_projection). */
int ath_forward_projection_ws(const float* image, int nrow, int nslice,
                              int ncol, const float* angles, float center,
                              double mask_radius, double scale,
                              void* workspace, size_t workspace_bytes,
                              float* sino, void* stream);

/* Adjoint of ath_ramp_filter for float32 rows: circular correlation with
 * the taps (plan built by ath_filter_plan_build from the REVERSED taps
 * g[(W - n) mod W] with fft_len >= 2 * (ncol + 2 * pad)), then the
 * transpose of the pad index map pad_mode (edge samples collect the padded
 * copies).  No reference counterpart. */
int ath_ramp_filter_adjoint(const float* in, int nrow, int nslice, int ncol,
                            int pad, int pad_mode, const void* plan,
                            int fft_len, int nwin, float* out, void* stream);

/* Device workspace bytes of a conversion: op 0 = ath_convert_sinogram_
 * 360_to_180 (the overlap means, plus two shifted images when the overlap
 * has a sub-pixel part), op 1 = ath_extend_sinogram (one shifted copy of
 * the input for a sub-pixel overlap, else 0).  0 for invalid arguments
 * (op 0 never needs 0 bytes). */
size_t ath_conversion_workspace_bytes(int op, int nrow, int nslice, int ncol,
                                      double overlap);

/* 360-degree -> 180-degree conversion of a half-acquisition (offset axis)
 * sinogram stack: rows [0, nrow/2] (image 1) and the left-right flipped
 * last nrow/2 + 1 rows (image 2) stitched over `overlap` columns (side 0:
 * overlap on the left of image 1, 1: on the right).  overlap_int =
 * round(overlap + 1e-3); a sub-pixel remainder shifts image 1 by it (side
 * 1; its negative on side 0) and image 2 by the opposite amount with
 * scipy.ndimage.shift's cubic spline (mode "nearest") before stitching.
 * norm != 0 scales image 2 by the ratio of the overlap means.  wei1/wei2:
 * caller weight matrices (nrow/2 + 1, ncol) of wei_dtype ATH_F32 or ATH_F64
 * shared by every slice, or both NULL for the linear ramp of
 * make_weight_matrix.  sino360: (nrow, nslice, ncol); out: (nrow/2 + 1,
 * nslice, width) with width >= 2*ncol - overlap_int (extra columns repeat
 * the edge); work: ath_conversion_workspace_bytes(0, ...) bytes.  Replaces
 * convert_sinogram_360_to_180 / stitch_image / make_weight_matrix
 * (algotom/prep/conversion.py:318-364, :76-144, :45-73). */
int ath_convert_sinogram_360_to_180(const float* sino360, int nrow,
                                    int nslice, int ncol, double overlap,
                                    int side, int norm, int width,
                                    const void* wei1, const void* wei2,
                                    int wei_dtype, void* work,
                                    size_t work_bytes, float* out,
                                    void* stream);

/* extend_sinogram (algotom/prep/conversion.py:395-454): sino (nrow,
 * nslice, ncol) float32 -> out (nrow, nslice, 2*ncol - overlap_int)
 * float32: optional -log, the sub-pixel shift (scipy.ndimage.shift, cubic,
 * mode "nearest") by overlap - overlap_int (side 1; its negative on side
 * 0), the overlap columns weighted by numpy.linspace(1, 0, overlap_int)
 * (side 1, last columns) or linspace(0, 1, overlap_int) (side 0, first
 * columns), edge padding on the other side, times 2.  work:
 * ath_conversion_workspace_bytes(1, ...) bytes (NULL when that is 0). */
int ath_extend_sinogram(const float* sino, int nrow, int nslice, int ncol,
                        double overlap, int side, int apply_log, void* work,
                        size_t work_bytes, float* out, void* stream);

/* flat_field_correction (algotom/prep/correction.py:45-98) without its
 * artifact-removal options: out = (float32(proj) - dark) / denom, or
 * float32(proj) / denom when dark is NULL.  proj element (n, i, j) at
 * proj[n*outer_stride + i*row_stride + j] (float32, float64 or ATH_U16;
 * n < nouter, i < height, j < width); dark / denom (height, width) of
 * field_dtype ATH_F32 or ATH_F64, which is also the type of the arithmetic
 * and of the contiguous out, as numpy's promotion makes it.  denom is the
 * caller's ratio*flat - dark (or ratio*flat) with zeros replaced as the
 * reference does (:78-98). */
int ath_flat_field_correction(const void* proj, int in_dtype,
                              long long nouter, int height, int width,
                              long long outer_stride, long long row_stride,
                              const void* dark, const void* denom,
                              int field_dtype, void* out, void* stream);

/* out = -log(in) for n float32 values (in == out allowed), the log taken
 * in float64 and rounded.  Replaces the apply-log step of
 * bpf_reconstruction (reconstruction.py:634-639), which back-projects the
 * logged sinogram before any filtering.  Values <= 0 must have been
 * replaced by the caller (the reference sets them to 1 with a warning). */
int ath_minus_log(const float* in, float* out, long long n, void* stream);

/* ---- BPF: 2D circular ramp filter (reconstruction.py:494-560) ---------- */

/* Bytes of a circular-filter plan for FFT length fft_len (twiddles + the
 * 2D spectrum of the tiled kernel, float64); 0 if fft_len is not a power
 * of two in [64, 32768].  fft_len comes from ath_filter_fft_length(ncol,
 * pad), the same rule as the 1D filter. */
size_t ath_circular_filter_plan_bytes(int fft_len);

/* Device scratch bytes ath_circular_filter_plan_build needs. */
size_t ath_circular_filter_plan_work_bytes(int width, int fft_len);

/* Builds the plan of the W x W real kernel g2 (device, row-major float64),
 * W = ncol + 2*pad: g2 = Re(ifft2(ifftshift(R2))) of the circular ramp
 * window R2 (make_circular_ramp_window, reconstruction.py:494-521), so that
 * the filter equals apply_circular_ramp_filter's spectral product
 * (:555-559). */
int ath_circular_filter_plan_build(const double* kernel2d, int width,
                                   int fft_len, void* plan, size_t plan_bytes,
                                   void* work, size_t work_bytes,
                                   void* stream);

/* Device workspace bytes of ath_circular_ramp_filter for nimg images. */
size_t ath_circular_ramp_filter_workspace_bytes(int ncol, int pad,
                                                int fft_len, int nimg);

/* apply_circular_ramp_filter (reconstruction.py:524-560) of nimg square
 * float32 images (ncol x ncol; element (j, y, x) at
 * images[j*img_stride + y*row_stride + x]): numpy.pad by `pad` in
 * `pad_mode` on both axes, 2D circular convolution with the plan's kernel,
 * crop; float64 arithmetic.  Then, as bpf_reconstruction does
 * (:663-684): the circle mask of radius mask_radius (< 0: none) and, if
 * scale_nrow >= 2, the factor pi/(scale_nrow - 1) -- in float64 for
 * out_dtype ATH_F64 (the 2D branch), after rounding to float32 for ATH_F32
 * (the 3D branch).  Output element (j, y, x) at
 * out[j*out_img_stride + y*out_row_stride + x]. */
int ath_circular_ramp_filter(const float* images, int nimg, int ncol,
                             long long row_stride, long long img_stride,
                             int pad, int pad_mode, const void* plan,
                             int fft_len, void* workspace,
                             size_t workspace_bytes, void* out, int out_dtype,
                             long long out_row_stride,
                             long long out_img_stride, double mask_radius,
                             int scale_nrow, void* stream);

/* ---- DFI: direct Fourier inversion (reconstruction.py:687-907) -------- */

/* DFI angle handling (__dfi_handle_angles, reconstruction.py:731-746) */
#define ATH_DFI_NO_ANGLES 0 /* angles=None: the sinogram is used as it is */
#define ATH_DFI_ANGLES 1    /* shift by the first angle along the angle axis */
#define ATH_DFI_FOLD_360 2  /* a ~360-degree scan: fold to nrow//2 + 1 rows,
                               then shift */

/* Bytes of a DFI plan for padded width `width` (odd; the Bluestein FFT
 * length, the smallest power of two >= 2*width - 1, must be <= 32768, so
 * width <= 16383); 0 if unsupported. */
size_t ath_dfi_plan_bytes(int width);

/* Builds the plan: twiddles, Bluestein chirp and kernel spectrum for the
 * length-`width` DFT, the smoothing window (device float64, `width`
 * values = make_smoothing_window(filter_name, width), reconstruction.py
 * :877-878; NULL = none) and the cubic / quintic B-spline inverse-filter
 * taps of scipy.ndimage's spline prefilter. */
int ath_dfi_plan_build(const double* window, int width, void* plan,
                       size_t plan_bytes, void* stream);

/* Device workspace bytes of ath_dfi_reconstruction (independent of the
 * slice count: slices are processed one after another). */
size_t ath_dfi_workspace_bytes(int nrow_in, int ncol, int nrow, int width);

/* dfi_reconstruction (reconstruction.py:809-907) below its Python argument
 * handling, for nslice sinograms; sinogram j's element (angle i, column c)
 * at sino[j*slice_stride + i*row_stride + c] (float32 or float64).
 *   __dfi_handle_sinogram (:749-781): odd width ncol1 (edge column added
 *     when ncol is even), scipy.ndimage.shift by xshift = (ncol1-1)/2 -
 *     center (order 3, mode "nearest") in the input's precision; with
 *     angle_mode != ATH_DFI_NO_ANGLES the angle handling (:731-746) with
 *     row_shift = b_ang/step and flip = angles[-1] < angles[0], giving nrow
 *     rows; numpy.pad by `pad` columns in pad_mode (width = ncol1 + 2*pad);
 *   apply_log (:862-867): values <= 0 set to 1 (*flag = 1, the caller
 *     warns), then -log in the input's precision;
 *   __dfi_single_slice (:784-806): row FFT (ifftshift / fftshift), window,
 *     order-5 map_coordinates (mode "reflect") at generate_mapping_coordinate
 *     (:687-728), circle mask, inverse 2D FFT, fftshift, real part, crop to
 *     ncol x ncol (float64 arithmetic throughout);
 *   ratio mask (:896-904): zero where x^2 + y^2 > mask_r2 about ncol//2
 *     (mask_r2 < 0: none).
 * Output element (y, j, x) at out[j*out_slice_stride + y*out_row_stride +
 * x], float64 (the reference's result type) or float32. */
int ath_dfi_reconstruction(const void* sino, int in_dtype, int nrow_in,
                           int nslice, int ncol, long long row_stride,
                           long long slice_stride, double xshift, int pad,
                           int pad_mode, int apply_log, int angle_mode,
                           double row_shift, int flip, int nrow,
                           const void* plan, int width, void* workspace,
                           size_t workspace_bytes, void* out, int out_dtype,
                           long long out_row_stride,
                           long long out_slice_stride, double mask_r2,
                           int* flag, void* stream);

/* ---- vertical slices (algotom/rec/vertrec.py) ------------------------- */

/* typing of the vertical back-projection (which reference kernel) */
#define ATH_VERT_GPU 0  /* float32 trig/points/position, float64 chunk sums */
#define ATH_VERT_CPU 1  /* float64 trig/position, float32 chunk images */

/* Flat-field correction and -log of raw projections (vertrec.py:572-583):
 * out = -log((proj - dark) / flat_dark) for proj[n*proj_stride +
 * i*row_stride + j] (n < nproj, i < height, j < width; float32 or
 * ATH_U16), out contiguous (nproj, height, width) float32.  dark /
 * flat_dark: (height, width) float32 (flat - dark with zeros replaced by
 * 1, as the reference does), or both NULL for no correction.  With
 * apply_log, values <= 0 are set to 1 before the log and *flag is set to 1
 * (the reference's warning); flag may be NULL only without apply_log. */
int ath_flat_field_log(const void* proj, int in_dtype, int nproj, int height,
                       int width, long long proj_stride, long long row_stride,
                       const float* dark, const float* flat_dark,
                       int apply_log, float* out, int* flag, void* stream);

/* Vertical back-projection of proj (nproj, height, width) float32 onto
 * nslice lines of width points (xs, ys: (nslice, width) float64, relative
 * to the image centre): out[s][i][j] = sum over projections of proj[n][i]
 * interpolated at x_sj cos(theta_n) + y_sj sin(theta_n) + center, summed
 * in chunks of `chunk` projections as the reference's chunk loop does
 * (vertrec.py:584-620).  mode ATH_VERT_GPU reproduces
 * vertical_back_projection_gpu / _gpu_chunk (:176-335), ATH_VERT_CPU
 * vertical_back_projection_cpu / _cpu_chunk (:53-173).  accumulate: add
 * to out instead of overwriting (for projection stacks fed in pieces of
 * whole chunks).  centers: NULL, or one rotation centre per line (device,
 * float64) used instead of `center` -- every candidate centre of
 * find_center_vertical_slice (vertrec.py:1264-1428) in one launch. */
int ath_vertical_back_projection(const float* proj, int nproj, int height,
                                 int width, const double* angles,
                                 const double* xs, const double* ys,
                                 int nslice, double center,
                                 const double* centers, int edge_pad,
                                 int mode, int chunk, int accumulate,
                                 float* out, void* stream);

/* In place on nslice*height*width values (ATH_F32 or ATH_F64): columns
 * outside [mask_lo, mask_hi) set to zero (mask_lo < 0: no mask), then
 * times pi / (nrow - 1) (nrow < 2: no scale) -- the masking and scale of
 * vertical_reconstruction (vertrec.py:686-695). */
int ath_vertical_finish(void* data, int dtype, long long total, int width,
                        int mask_lo, int mask_hi, int nrow, void* stream);

/* Gaussian smoothing along rows (nrows x ncol float32), as
 * scipy.ndimage.gaussian_filter1d(x, sigma, axis=-1) with mode "reflect":
 * weights (2*radius+1 float64, device) are the normalised Gaussian the
 * caller computed; sums in float64.  radius <= 64.  Used by the centre
 * search (reconstruction.py:1315-1316). */
int ath_gaussian_filter_rows(const float* in, float* out, int nrows,
                             int ncol, const double* weights, int radius,
                             void* stream);

/* Order-1 zoom of a 2D float32 image (nrow x ncol -> out_nrow x out_ncol),
 * as scipy.ndimage.zoom(x, z, order=1, mode="nearest") with its
 * corner-aligned grid: float64 bilinear weights, float32 output.  Replaces
 * the coarse-search zoom of find_center_based_slice_metric
 * (reconstruction.py:1434). */
int ath_zoom_linear(const float* in, int nrow, int ncol, float* out,
                    int out_nrow, int out_ncol, void* stream);

/* Centre-search metrics of nimg images (nrows x ncols, ATH_F32 or ATH_F64,
 * image k at images + k*stride elements):
 *   metric 0  histogram entropy of clip((img - nmin)/(nmax - nmin), 0, 1)
 *             (reconstruction.py:1211-1220, 1243-1246), evaluated in the
 *             image's precision (nmin/nmax rounded to float32 for float32
 *             images);
 *   metric 1  mean gradient magnitude of the float32 image (:1223-1228;
 *             nrows, ncols >= 2);
 *   metric 2  vertrec's autocorrelation metric (vertrec.py:1255-1261): the
 *             mean over rows of sum(row^2), the lag-0 peak of each row's
 *             circular autocorrelation.
 * Sums follow numpy's order (pairwise within 8192-value chunks).  out: nimg
 * float64 (device); workspace: caller device memory of at least
 * ath_slice_metrics_workspace_bytes(...) bytes (may be NULL when that is 0). */
int ath_slice_metrics(const void* images, int dtype, int nimg, int nrows,
                      int ncols, long long stride, double nmin, double nmax,
                      int metric, double* out, void* workspace,
                      size_t workspace_bytes, void* stream);

/* Device workspace bytes ath_slice_metrics needs (per-block partial
 * histograms or per-chunk partial sums); 0 when none is needed or for
 * invalid arguments. */
size_t ath_slice_metrics_workspace_bytes(int nimg, int nrows, int ncols,
                                         int metric);

/* scipy.ndimage.gaussian_filter1d(x, sigma, axis=1) of a contiguous
 * (nouter, length, ninner) float32 array (mode "reflect", float64 sums,
 * float32 out; weights as for ath_gaussian_filter_rows).  Two calls make
 * ndi.gaussian_filter(projections, (0, sigma, sigma)) of
 * find_center_vertical_slice (vertrec.py:1331-1332). */
int ath_gaussian_filter_axis(const float* in, float* out, int nouter,
                             int length, int ninner, const double* weights,
                             int radius, void* stream);

/* Device workspace bytes needed by ath_fbp_reconstruction. */
size_t ath_fbp_workspace_bytes(int nrow, int nslice, int ncol, int pad,
                               int nwin, int filt_dtype);

/* Whole FBP of a (nrow, nslice, ncol) stack: filter (log, pad, windowed
 * ramp) into the workspace, then back-projection with mask and scale into
 * recon (ncol, nslice, ncol).  taps: nwin x (ncol + 2*pad) float64 (device).
 * Replaces fbp_reconstruction (reconstruction.py:362-493) below its Python
 * argument handling. */
int ath_fbp_reconstruction(const void* sino, int in_dtype, int nrow,
                           int nslice, int ncol, const float* angles,
                           float center, double mask_radius,
                           const double* taps, int nwin, int pad,
                           int pad_mode, int apply_log, int acc,
                           int filt_dtype, void* workspace,
                           size_t workspace_bytes, float* recon,
                           void* stream);

/* Largest workspace ath_fbp_reconstruction_noalloc can use (64 MiB): e.g.
 * one 3601 x 4096 float32 sinogram with its plan, or a 1801 x 8 x 1024
 * stack. */
#define ATH_FBP_NOALLOC_MAX_BYTES ((size_t)64 << 20)

/* ath_fbp_reconstruction with no workspace argument; it allocates nothing.
 * No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a fixed device array of
 * ATH_FBP_NOALLOC_MAX_BYTES in the library, one per GPU, held until the
 * process ends; a call whose ath_fbp_workspace_bytes(...) exceeds it
 * returns ATH_ERR_WORKSPACE.  Every call on one GPU shares that array, so
 * calls that overlap on one GPU (other streams or threads) race; calls
 * ordered on one stream do not.  Each call writes every workspace byte it
 * reads, so nothing carries over from the previous call. */
int ath_fbp_reconstruction_noalloc(const void* sino, int in_dtype, int nrow,
                                   int nslice, int ncol, const float* angles,
                                   float center, double mask_radius,
                                   const double* taps, int nwin, int pad,
                                   int pad_mode, int apply_log, int acc,
                                   int filt_dtype, float* recon,
                                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ALGOTOM_HIP_H */
