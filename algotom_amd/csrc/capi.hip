// extern "C" entry points of libalgotom_hip (declared in include/algotom_hip.h).
// Argument checks mirror the reference's own (reconstruction.py:412-419);
// nothing here allocates device memory or copies between host and device.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "ath_internal.h"

namespace ath {

namespace {
thread_local int g_status = ATH_OK;
thread_local std::string g_message;
}  // namespace

int set_error(int status, const std::string& msg) {
    g_status = status;
    g_message = msg;
    return status;
}

void clear_error() {
    g_status = ATH_OK;
    g_message.clear();
}

int check_hip(hipError_t err, const char* what) {
    if (err == hipSuccess) return ATH_OK;
    return set_error(ATH_ERR_HIP, std::string(what) + ": " +
                                      hipGetErrorName(err) + " (" +
                                      hipGetErrorString(err) + ")");
}

namespace {
struct LaunchKey {
    const void* kernel;
    int device, threads;
    size_t lds;
    bool operator<(const LaunchKey& o) const {
        return std::tie(kernel, device, threads, lds) <
               std::tie(o.kernel, o.device, o.threads, o.lds);
    }
};
std::mutex g_launch_mutex;
std::map<LaunchKey, int> g_resident;
std::map<std::pair<const void*, int>, size_t> g_dyn_lds;
}  // namespace

int resident_grid(const void* kernel, int threads, size_t lds, int* out) {
    int dev = 0;
    int st = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (st != ATH_OK) return st;
    const LaunchKey key{kernel, dev, threads, lds};
    {
        std::lock_guard<std::mutex> lock(g_launch_mutex);
        auto it = g_resident.find(key);
        if (it != g_resident.end()) {
            *out = it->second;
            return ATH_OK;
        }
    }
    int per_cu = 0, cus = 0;
    st = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                       &per_cu, kernel, threads, lds),
                   "occupancy query");
    if (st != ATH_OK) return st;
    st = check_hip(hipDeviceGetAttribute(
                       &cus, hipDeviceAttributeMultiprocessorCount, dev),
                   "CU count");
    if (st != ATH_OK) return st;
    const int n = std::max(1, per_cu) * std::max(1, cus);
    std::lock_guard<std::mutex> lock(g_launch_mutex);
    g_resident[key] = n;
    *out = n;
    return ATH_OK;
}

int allow_dynamic_lds(const void* kernel, size_t bytes) {
    int dev = 0;
    int st = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (st != ATH_OK) return st;
    std::lock_guard<std::mutex> lock(g_launch_mutex);
    size_t& done = g_dyn_lds[{kernel, dev}];
    if (done >= bytes) return ATH_OK;
    st = check_hip(hipFuncSetAttribute(
                       kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                       (int)bytes),
                   "hipFuncSetAttribute(max dynamic LDS)");
    if (st == ATH_OK) done = bytes;
    return st;
}

namespace {

bool valid_dtype(int t) {
    return t == ATH_F32 || t == ATH_F64 || t == ATH_F16 || t == ATH_BF16;
}

size_t dtype_bytes(int t) {
    return t == ATH_F64 ? 8 : (t == ATH_F32 ? 4 : 2);
}

int check_geometry(int nrow, int nslice, int ncol) {
    if (nrow < 1 || nslice < 1 || ncol < 1)
        return set_error(ATH_ERR_INVALID_ARG,
                         "nrow, nslice and ncol must be positive");
    return ATH_OK;
}

int fill_bp_args(BpArgs& a, const void* sino, int nrow, int nslice, int ncol,
                 const float* angles, float center, int edge_pad,
                 double mask_radius, int apply_scale, float* recon) {
    if (!sino || !angles || !recon)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    int st = check_geometry(nrow, nslice, ncol);
    if (st != ATH_OK) return st;
    // The reference checks 0 <= center < ncol in Python (reconstruction.py
    // :412-413); its kernels accept any value, as this boundary does.
    if (!std::isfinite(center))
        return set_error(ATH_ERR_INVALID_ARG, "center must be finite");
    if (apply_scale && nrow < 2)
        return set_error(ATH_ERR_INVALID_ARG,
                         "the pi/(nrow-1) scale needs at least two angles");
    a.sino = sino;
    a.recon = recon;
    a.recon2 = nullptr;
    a.angles = angles;
    a.nrow = nrow;
    a.nslice = nslice;
    a.ncol = ncol;
    a.center = (double)center;
    a.icenter = 0.5 * (double)(ncol - 1);
    a.edge_pad = edge_pad ? 1 : 0;
    a.use_mask = mask_radius >= 0.0 ? 1 : 0;
    a.mask_r2 = mask_radius * mask_radius;
    a.mask_c = ncol / 2;
    a.apply_scale = apply_scale ? 1 : 0;
    a.pi32 = (float)M_PI;
    a.denom = (float)(nrow - 1);
    return ATH_OK;
}

}  // namespace
}  // namespace ath

using namespace ath;

namespace ath {

// The fused filter + back-projection on a workspace of workspace_bytes:
// plan (aligned to 256 bytes), then the filtered stack.  Every byte it
// reads is written first in the same call (the plan by plan_build, the
// filtered stack by the filter), so no part needs clearing.
int fbp_into_workspace(const void* sino, int in_dtype, int nrow, int nslice,
                       int ncol, const float* angles, float center,
                       double mask_radius, const double* taps, int nwin,
                       int pad, int pad_mode, int apply_log, int acc,
                       int filt_dtype, void* workspace,
                       size_t workspace_bytes, float* recon, hipStream_t s) {
    if (!workspace || !taps)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    const size_t need = ath_fbp_workspace_bytes(nrow, nslice, ncol, pad, nwin,
                                                filt_dtype);
    if (need == 0)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry or dtype");
    if (workspace_bytes < need)
        return set_error(ATH_ERR_WORKSPACE, "workspace too small");
    if (filt_dtype != ATH_F32 && acc != ATH_ACC_F32)
        return set_error(ATH_ERR_UNSUPPORTED,
                         "reduced-precision filtered storage needs "
                         "acc = ATH_ACC_F32");
    BpArgs a;
    int st = fill_bp_args(a, sino, nrow, nslice, ncol, angles, center, 0,
                          mask_radius, 1, recon);
    if (st != ATH_OK) return st;
    const int L = ath_filter_fft_length(ncol, pad);
    if (L < 0) return L;
    const size_t plan = plan_bytes(L, nwin);
    char* ws = static_cast<char*>(workspace);
    void* filt = ws + ((plan + 255) & ~(size_t)255);
    st = plan_build(taps, nwin, ncol + 2 * pad, L, ws, s);
    if (st != ATH_OK) return st;
    st = ath_ramp_filter(sino, in_dtype, nrow, nslice, ncol, pad, pad_mode,
                         apply_log, ws, L, nwin, filt, filt_dtype, s);
    if (st != ATH_OK) return st;
    a.sino = filt;
    return back_projection(a, filt_dtype, acc, 0, 0, s);
}

namespace {
// Workspace of ath_fbp_reconstruction_noalloc: one copy per GPU, loaded with
// the library's code object and held until the process ends.  Host code
// never passes `scratch` itself to a kernel (on the host it names a stand-in,
// not device memory): it uses the address hipGetSymbolAddress returns on
// each device.
__device__ __align__(256) unsigned char scratch[ATH_FBP_NOALLOC_MAX_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int noalloc_scratch(void** out) {
    int dev = 0;
    int st = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (st != ATH_OK) return st;
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(ATH_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        st = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                       "hipGetSymbolAddress(scratch)");
        if (st != ATH_OK) return st;
        if (!p)
            return set_error(ATH_ERR_HIP,
                             "hipGetSymbolAddress(scratch) gave no address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return ATH_OK;
}

}  // namespace ath

extern "C" {

int ath_version(void) { return 100; }

int ath_last_error(char* buf, size_t len) {
    if (buf && len > 0) {
        std::snprintf(buf, len, "%s", g_message.c_str());
    }
    return g_status;
}

int ath_device_count(int* count) {
    clear_error();
    if (!count) return set_error(ATH_ERR_INVALID_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        (void)hipGetLastError();
        return set_error(ATH_ERR_NO_DEVICE,
                         std::string("no HIP device: ") + hipGetErrorName(e));
    }
    *count = n;
    return ATH_OK;
}

int ath_filter_fft_length(int ncol, int pad) {
    clear_error();
    if (ncol < 1 || pad < 0)
        return set_error(ATH_ERR_INVALID_ARG, "ncol must be > 0, pad >= 0");
    const int L = fft_length(ncol, pad);
    if (L < 0 || L > kMaxFilterFft)
        return set_error(ATH_ERR_UNSUPPORTED,
                         "ramp filter: padded width needs an FFT above 32768");
    return L;
}

size_t ath_filter_plan_bytes(int fft_len, int nwin) {
    if (fft_len < 64 || (fft_len & (fft_len - 1)) || nwin < 1) return 0;
    return plan_bytes(fft_len, nwin);
}

int ath_filter_plan_build(const double* taps, int nwin, int width,
                          int fft_len, void* plan, size_t plan_bytes_given,
                          void* stream) {
    clear_error();
    if (!taps || !plan) return set_error(ATH_ERR_INVALID_ARG, "null pointer");
    if (nwin < 1 || width < 1)
        return set_error(ATH_ERR_INVALID_ARG, "nwin and width must be > 0");
    if (fft_len < 64 || fft_len > kMaxFilterFft || (fft_len & (fft_len - 1)))
        return set_error(ATH_ERR_INVALID_ARG,
                         "fft_len must be a power of two in [64, 32768]");
    if (plan_bytes_given < plan_bytes(fft_len, nwin))
        return set_error(ATH_ERR_WORKSPACE, "filter plan buffer too small");
    return plan_build(taps, nwin, width, fft_len, plan,
                      static_cast<hipStream_t>(stream));
}

int ath_filter_taps_from_window(const double* win_re, const double* win_im,
                                int width, double* taps_out) {
    clear_error();
    if (!win_re || !win_im || !taps_out || width < 1)
        return set_error(ATH_ERR_INVALID_ARG, "bad arguments");
    // R' = ifftshift(R); r' = ifft(R'); g = Re(ifftshift(r'))
    const int W = width, h = W / 2;
    std::vector<long double> cw(W), sw(W);
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int k = 0; k < W; ++k) {
        cw[k] = std::cos(two_pi * (long double)k / (long double)W);
        sw[k] = std::sin(two_pi * (long double)k / (long double)W);
    }
    for (int m = 0; m < W; ++m) {
        const int n = (m + h) % W;  // ifftshift of the inverse transform
        long double acc = 0.0L;
        for (int k = 0; k < W; ++k) {
            const int src = (k + h) % W;  // ifftshift of the window
            const int t = (int)(((long long)k * n) % W);
            acc += (long double)win_re[src] * cw[t] -
                   (long double)win_im[src] * sw[t];
        }
        taps_out[m] = (double)(acc / (long double)W);
    }
    return ATH_OK;
}

int ath_ramp_filter(const void* sino, int in_dtype, int nrow, int nslice,
                    int ncol, int pad, int pad_mode, int apply_log,
                    const void* plan, int fft_len, int nwin, void* out,
                    int out_dtype, void* stream) {
    clear_error();
    if (!sino || !plan || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    int st = check_geometry(nrow, nslice, ncol);
    if (st != ATH_OK) return st;
    if (pad < 0) return set_error(ATH_ERR_INVALID_ARG, "pad must be >= 0");
    if (pad_mode < ATH_PAD_EDGE || pad_mode > ATH_PAD_MINIMUM)
        return set_error(ATH_ERR_INVALID_ARG, "unsupported pad mode");
    if (!valid_dtype(out_dtype) ||
        (in_dtype != ATH_F32 && in_dtype != ATH_F64))
        return set_error(ATH_ERR_INVALID_ARG, "bad dtype");
    if (nwin != 1 && nwin != nrow)
        return set_error(ATH_ERR_INVALID_ARG, "nwin must be 1 or nrow");
    const long long width = (long long)ncol + 2LL * pad;
    if (fft_len < 64 || fft_len > kMaxFilterFft ||
        (fft_len & (fft_len - 1)) || (long long)fft_len < 2 * width - pad)
        return set_error(ATH_ERR_INVALID_ARG,
                         "fft_len must be a power of two >= 2*(ncol+2*pad)-pad "
                         "(see ath_filter_fft_length)");
    return ramp_filter(sino, in_dtype, nrow, nslice, ncol, pad, pad_mode,
                       apply_log, plan, fft_len, nwin, out, out_dtype,
                       static_cast<hipStream_t>(stream));
}

int ath_back_projection(const void* sino, int sino_dtype, int nrow,
                        int nslice, int ncol, const float* angles,
                        float center, int edge_pad, int acc,
                        double mask_radius, int apply_scale, float* recon,
                        void* stream) {
    clear_error();
    BpArgs a;
    int st = fill_bp_args(a, sino, nrow, nslice, ncol, angles, center,
                          edge_pad, mask_radius, apply_scale, recon);
    if (st != ATH_OK) return st;
    return back_projection(a, sino_dtype, acc, 0, 0,
                           static_cast<hipStream_t>(stream));
}

int ath_back_projection_with_derivative(const float* sino, int nrow,
                                        int nslice, int ncol,
                                        const float* angles, float center,
                                        int edge_pad, double mask_radius,
                                        int apply_scale, float* recon,
                                        float* deriv, void* stream) {
    clear_error();
    if (!deriv) return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    BpArgs a;
    int st = fill_bp_args(a, sino, nrow, nslice, ncol, angles, center,
                          edge_pad, mask_radius, apply_scale, recon);
    if (st != ATH_OK) return st;
    a.recon2 = deriv;
    return back_projection(a, ATH_F32, ATH_ACC_F64_DERIV, 0, 0,
                           static_cast<hipStream_t>(stream));
}

int ath_back_projection_tiled(const void* sino, int sino_dtype, int nrow,
                              int nslice, int ncol, const float* angles,
                              float center, int edge_pad, int acc,
                              double mask_radius, int apply_scale,
                              float* recon, int slice_batch, int tile_rows,
                              void* stream) {
    clear_error();
    BpArgs a;
    int st = fill_bp_args(a, sino, nrow, nslice, ncol, angles, center,
                          edge_pad, mask_radius, apply_scale, recon);
    if (st != ATH_OK) return st;
    if (slice_batch < 1 || tile_rows < 4 || tile_rows % 4)
        return set_error(ATH_ERR_INVALID_ARG,
                         "slice_batch must be >= 1, tile_rows a multiple of 4");
    return back_projection(a, sino_dtype, acc, slice_batch, tile_rows / 4,
                           static_cast<hipStream_t>(stream));
}

long long ath_bp_job_tiles(int ncol, int nslice, int slice_batch,
                           int tile_rows) {
    if (ncol < 1 || nslice < 1 || slice_batch < 1 || tile_rows < 4 ||
        tile_rows % 4)
        return 0;
    const long long tx = (ncol + 63) / 64, ty = (ncol + tile_rows - 1) / tile_rows;
    return tx * ty * ((nslice + slice_batch - 1) / slice_batch);
}

int ath_back_projection_batched(const ath_bp_job* jobs, int njobs,
                                long long total_tiles, int sino_dtype,
                                int acc, int slice_batch, int tile_rows,
                                int nblocks, int* counter, void* stream) {
    clear_error();
    if (!jobs || !counter)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (njobs < 1 || total_tiles < 1 || total_tiles > 0x7fffffffLL)
        return set_error(ATH_ERR_INVALID_ARG,
                         "need 1 <= njobs and 1 <= total_tiles < 2^31");
    if (slice_batch < 1 || tile_rows < 4 || tile_rows % 4)
        return set_error(ATH_ERR_INVALID_ARG,
                         "slice_batch must be >= 1, tile_rows a multiple of 4");
    return back_projection_batched(jobs, njobs, total_tiles, sino_dtype, acc,
                                   slice_batch, tile_rows / 4, nblocks,
                                   counter, static_cast<hipStream_t>(stream));
}

int ath_ramp_filter_batched(const ath_filter_job* jobs, int njobs,
                            long long total_rows, int in_dtype,
                            int out_dtype, int fft_len, int pad_mode,
                            int apply_log, void* stream) {
    clear_error();
    if (!jobs) return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (njobs < 1 || total_rows < 1 || total_rows > 0x7fffffffLL)
        return set_error(ATH_ERR_INVALID_ARG,
                         "need 1 <= njobs and 1 <= total_rows < 2^31");
    if (pad_mode < ATH_PAD_EDGE || pad_mode > ATH_PAD_MINIMUM)
        return set_error(ATH_ERR_INVALID_ARG, "unsupported pad mode");
    if (!valid_dtype(out_dtype) ||
        (in_dtype != ATH_F32 && in_dtype != ATH_F64))
        return set_error(ATH_ERR_INVALID_ARG, "bad dtype");
    if (fft_len < 64 || fft_len > kMaxFilterFft || (fft_len & (fft_len - 1)))
        return set_error(ATH_ERR_INVALID_ARG,
                         "fft_len must be a power of two in [64, 32768]");
    // every job must satisfy fft_len >= 2*(ncol+2*pad) - pad: checked by the
    // caller, who built each job's plan for this fft_len
    return ramp_filter_batched(jobs, njobs, total_rows, in_dtype, out_dtype,
                               fft_len, pad_mode, apply_log,
                               static_cast<hipStream_t>(stream));
}

int ath_back_projection_naive(const float* sino, int nrow, int nslice,
                              int ncol, const float* angles, float center,
                              int edge_pad, int acc, double mask_radius,
                              int apply_scale, float* recon, void* stream) {
    clear_error();
    BpArgs a;
    int st = fill_bp_args(a, sino, nrow, nslice, ncol, angles, center,
                          edge_pad, mask_radius, apply_scale, recon);
    if (st != ATH_OK) return st;
    if (nslice > 65535)
        return set_error(ATH_ERR_UNSUPPORTED, "naive kernel: nslice > 65535");
    return back_projection_naive(a, acc, static_cast<hipStream_t>(stream));
}

namespace {

int forward_projection_entry(const float* image, int nrow, int nslice,
                             int ncol, const float* angles, float center,
                             double mask_radius, double scale, void* work,
                             size_t work_bytes, bool use_work, float* sino,
                             void* stream) {
    clear_error();
    if (!image || !angles || !sino)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    int st = check_geometry(nrow, nslice, ncol);
    if (st != ATH_OK) return st;
    if (!std::isfinite(center) || !std::isfinite(scale))
        return set_error(ATH_ERR_INVALID_ARG, "center and scale must be finite");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    FpArgs a;
    a.image_t = nullptr;
    if (use_work) {
        if (!work ||
            work_bytes < ath_forward_projection_workspace_bytes(nslice, ncol))
            return set_error(ATH_ERR_WORKSPACE,
                             "forward projection: workspace too small");
        float* t = static_cast<float*>(work);
        st = transpose_slices(image, t, nslice, ncol, s);
        if (st != ATH_OK) return st;
        a.image_t = t;
    }
    a.image = image;
    a.sino = sino;
    a.angles = angles;
    a.nrow = nrow;
    a.nslice = nslice;
    a.ncol = ncol;
    a.center = (double)center;
    a.icenter = 0.5 * (double)(ncol - 1);
    a.use_mask = mask_radius >= 0.0 ? 1 : 0;
    a.mask_r2 = mask_radius * mask_radius;
    a.mask_c = ncol / 2;
    a.scale = scale;
    return forward_projection(a, s);
}

}  // namespace

int ath_forward_projection(const float* image, int nrow, int nslice,
                           int ncol, const float* angles, float center,
                           double mask_radius, double scale, float* sino,
                           void* stream) {
    return forward_projection_entry(image, nrow, nslice, ncol, angles,
                                    center, mask_radius, scale, nullptr, 0,
                                    false, sino, stream);
}

size_t ath_forward_projection_workspace_bytes(int nslice, int ncol) {
    if (nslice <= 0 || ncol <= 0) return 0;
    return (size_t)ncol * (size_t)nslice * (size_t)ncol * sizeof(float);
}

int ath_forward_projection_ws(const float* image, int nrow, int nslice,
                              int ncol, const float* angles, float center,
                              double mask_radius, double scale,
                              void* workspace, size_t workspace_bytes,
                              float* sino, void* stream) {
    return forward_projection_entry(image, nrow, nslice, ncol, angles,
                                    center, mask_radius, scale, workspace,
                                    workspace_bytes, true, sino, stream);
}

int ath_ramp_filter_adjoint(const float* in, int nrow, int nslice, int ncol,
                            int pad, int pad_mode, const void* plan,
                            int fft_len, int nwin, float* out, void* stream) {
    clear_error();
    if (!in || !plan || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    int st = check_geometry(nrow, nslice, ncol);
    if (st != ATH_OK) return st;
    if (pad < 0) return set_error(ATH_ERR_INVALID_ARG, "pad must be >= 0");
    if (pad_mode < ATH_PAD_EDGE || pad_mode > ATH_PAD_WRAP)
        return set_error(ATH_ERR_INVALID_ARG, "unsupported pad mode");
    if (nwin != 1 && nwin != nrow)
        return set_error(ATH_ERR_INVALID_ARG, "nwin must be 1 or nrow");
    const long long width = (long long)ncol + 2LL * pad;
    if (fft_len < 64 || fft_len > kMaxFilterAdjointFft ||
        (fft_len & (fft_len - 1)) || (long long)fft_len < 2 * width)
        return set_error(ATH_ERR_INVALID_ARG,
                         "adjoint filter: fft_len must be a power of two "
                         ">= 2*(ncol+2*pad), at most 32768");
    return ramp_filter_adjoint(in, nrow, nslice, ncol, pad, pad_mode, plan,
                               fft_len, nwin, out,
                               static_cast<hipStream_t>(stream));
}

size_t ath_conversion_workspace_bytes(int op, int nrow, int nslice, int ncol,
                                      double overlap) {
    if (nrow < 1 || nslice < 1 || ncol < 1 || (op != 0 && op != 1))
        return 0;
    const int ov = (int)std::nearbyint(overlap + 1.0e-3);
    const bool sub = overlap - (double)ov != 0.0;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    if (op == 1)
        return sub ? al((size_t)nrow * nslice * ncol * sizeof(float)) : 0;
    const size_t img = (size_t)(nrow / 2 + 1) * nslice * ncol * sizeof(float);
    return al(2 * (size_t)nslice * sizeof(float)) + (sub ? 2 * al(img) : 0);
}

int ath_convert_sinogram_360_to_180(const float* sino360, int nrow,
                                    int nslice, int ncol, double overlap,
                                    int side, int norm, int width,
                                    const void* wei1, const void* wei2,
                                    int wei_dtype, void* work,
                                    size_t work_bytes, float* out,
                                    void* stream) {
    clear_error();
    if (!sino360 || !out || !work)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    int st = check_geometry(nrow, nslice, ncol);
    if (st != ATH_OK) return st;
    if (nrow < 2) return set_error(ATH_ERR_INVALID_ARG, "nrow must be >= 2");
    if (!std::isfinite(overlap))
        return set_error(ATH_ERR_INVALID_ARG, "overlap must be finite");
    // conversion.py:107-108: the integer part rounds overlap + 1e-3
    const int ov = (int)std::nearbyint(overlap + 1.0e-3);
    const double sub = overlap - (double)ov;
    if (ov < 1 || ov > ncol)
        return set_error(ATH_ERR_INVALID_ARG, "overlap must be in [1, ncol]");
    if (side != 0 && side != 1)
        return set_error(ATH_ERR_INVALID_ARG, "side must be 0 or 1");
    if ((long long)width < 2LL * ncol - ov)
        return set_error(ATH_ERR_INVALID_ARG,
                         "width must be >= 2*ncol - overlap");
    if ((wei1 == nullptr) != (wei2 == nullptr))
        wei1 = wei2 = nullptr;  // stitch_image uses the ramp then (:119)
    if (wei1 && wei_dtype != ATH_F32 && wei_dtype != ATH_F64)
        return set_error(ATH_ERR_INVALID_ARG,
                         "weight matrices must be float32 or float64");
    if (work_bytes < ath_conversion_workspace_bytes(0, nrow, nslice, ncol,
                                                    overlap))
        return set_error(ATH_ERR_WORKSPACE, "workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int n180 = nrow / 2 + 1;
    char* ws = static_cast<char*>(work);
    float* factors = reinterpret_cast<float*>(ws);
    const float* top = sino360;
    const float* bot = sino360 + (size_t)(nrow - n180) * nslice * ncol;
    if (sub == 0.0)
        return stitch_images(top, bot, 1, n180, nslice, ncol, ov, side, norm,
                             width, wei1, wei2, wei_dtype, factors, out, s);
    // conversion.py:109-117: image 1 shifted by +sub (side 1) or -sub
    // (side 0), the flipped image 2 by the opposite amount
    const size_t img = (size_t)n180 * nslice * ncol;
    const size_t off1 = (2 * (size_t)nslice * sizeof(float) + 255) & ~(size_t)255;
    const size_t off2 = off1 + ((img * sizeof(float) + 255) & ~(size_t)255);
    float* t1 = reinterpret_cast<float*>(ws + off1);
    float* t2 = reinterpret_cast<float*>(ws + off2);
    const double s1 = side == 1 ? sub : -sub;
    st = shift_rows(top, (long long)n180 * nslice, ncol, 0, 0, s1, t1, s);
    if (st != ATH_OK) return st;
    st = shift_rows(bot, (long long)n180 * nslice, ncol, 1, 0, -s1, t2, s);
    if (st != ATH_OK) return st;
    return stitch_images(t1, t2, 0, n180, nslice, ncol, ov, side, norm, width,
                         wei1, wei2, wei_dtype, factors, out, s);
}

int ath_flat_field_log(const void* proj, int in_dtype, int nproj, int height,
                       int width, long long proj_stride, long long row_stride,
                       const float* dark, const float* flat_dark,
                       int apply_log, float* out, int* flag, void* stream) {
    clear_error();
    if (!proj || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if ((dark == nullptr) != (flat_dark == nullptr))
        return set_error(ATH_ERR_INVALID_ARG,
                         "dark and flat_dark must both be given or both NULL");
    if (apply_log && !flag)
        return set_error(ATH_ERR_INVALID_ARG, "apply_log needs a flag");
    if (in_dtype != ATH_F32 && in_dtype != ATH_U16)
        return set_error(ATH_ERR_INVALID_ARG,
                         "projections must be float32 or uint16");
    if (nproj < 1 || height < 1 || width < 1 || row_stride < width ||
        proj_stride < (long long)(height - 1) * row_stride + width)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    return flat_field_log(proj, in_dtype, nproj, height, width, proj_stride,
                          row_stride, dark, flat_dark, apply_log, out, flag,
                          static_cast<hipStream_t>(stream));
}

int ath_vertical_back_projection(const float* proj, int nproj, int height,
                                 int width, const double* angles,
                                 const double* xs, const double* ys,
                                 int nslice, double center,
                                 const double* centers, int edge_pad,
                                 int mode, int chunk, int accumulate,
                                 float* out, void* stream) {
    clear_error();
    if (!proj || !angles || !xs || !ys || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (nproj < 1 || height < 1 || width < 1 || nslice < 1 || chunk < 1)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    if (nslice > 65535 || (height + 31) / 32 > 65535)
        return set_error(ATH_ERR_UNSUPPORTED, "too many slices or rows");
    if (mode != ATH_VERT_GPU && mode != ATH_VERT_CPU)
        return set_error(ATH_ERR_INVALID_ARG, "bad vertical mode");
    if (!std::isfinite(center))
        return set_error(ATH_ERR_INVALID_ARG, "center must be finite");
    return vertical_back_projection(proj, nproj, height, width, angles, xs, ys,
                                    nslice, center, centers, edge_pad, mode,
                                    chunk, accumulate, out,
                                    static_cast<hipStream_t>(stream));
}

int ath_vertical_finish(void* data, int dtype, long long total, int width,
                        int mask_lo, int mask_hi, int nrow, void* stream) {
    clear_error();
    if (!data) return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (dtype != ATH_F32 && dtype != ATH_F64)
        return set_error(ATH_ERR_INVALID_ARG, "float32 or float64 data only");
    if (total < 0 || width < 1)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    if (total == 0) return ATH_OK;
    return vertical_finish(data, dtype, total, width, mask_lo, mask_hi, nrow,
                           static_cast<hipStream_t>(stream));
}

int ath_extend_sinogram(const float* sino, int nrow, int nslice, int ncol,
                        double overlap, int side, int apply_log, void* work,
                        size_t work_bytes, float* out, void* stream) {
    clear_error();
    if (!sino || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (nrow < 1 || nslice < 1 || ncol < 1)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    if (!std::isfinite(overlap))
        return set_error(ATH_ERR_INVALID_ARG, "overlap must be finite");
    const int ov = (int)std::nearbyint(overlap + 1.0e-3);
    const double sub = overlap - (double)ov;
    if (ov < 1 || ov > ncol)
        return set_error(ATH_ERR_INVALID_ARG, "overlap must be in [1, ncol]");
    if (side != 0 && side != 1)
        return set_error(ATH_ERR_INVALID_ARG, "side must be 0 or 1");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (sub == 0.0)
        return extend_sinogram(sino, nrow, nslice, ncol, ov, side, apply_log,
                               out, s);
    if (!work || work_bytes < ath_conversion_workspace_bytes(
                                  1, nrow, nslice, ncol, overlap))
        return set_error(ATH_ERR_WORKSPACE, "workspace too small");
    // conversion.py:435-453: (log), then the shift by +sub (side 1) or
    // -sub (side 0), then the weights and the padding
    float* t = static_cast<float*>(work);
    const int st = shift_rows(sino, (long long)nrow * nslice, ncol, 0,
                              apply_log, side == 1 ? sub : -sub, t, s);
    if (st != ATH_OK) return st;
    return extend_sinogram(t, nrow, nslice, ncol, ov, side, 0, out, s);
}

int ath_flat_field_correction(const void* proj, int in_dtype,
                              long long nouter, int height, int width,
                              long long outer_stride, long long row_stride,
                              const void* dark, const void* denom,
                              int field_dtype, void* out, void* stream) {
    clear_error();
    if (!proj || !denom || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (in_dtype != ATH_F32 && in_dtype != ATH_F64 && in_dtype != ATH_U16)
        return set_error(ATH_ERR_INVALID_ARG,
                         "projections must be float32, float64 or uint16");
    if (field_dtype != ATH_F32 && field_dtype != ATH_F64)
        return set_error(ATH_ERR_INVALID_ARG, "fields must be float32/64");
    if (nouter < 1 || height < 1 || width < 1 || row_stride < width ||
        outer_stride < (long long)(height - 1) * row_stride + width)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    return flat_field_correction(proj, in_dtype, nouter, height, width,
                                 outer_stride, row_stride, dark, denom,
                                 field_dtype, out,
                                 static_cast<hipStream_t>(stream));
}

int ath_minus_log(const float* in, float* out, long long n, void* stream) {
    clear_error();
    if (!in || !out) return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (n < 0) return set_error(ATH_ERR_INVALID_ARG, "negative length");
    if (n == 0) return ATH_OK;
    return minus_log(in, out, n, static_cast<hipStream_t>(stream));
}

static bool circular_len_ok(int L) {
    return L >= 64 && L <= 32768 && (L & (L - 1)) == 0;
}

size_t ath_circular_filter_plan_bytes(int fft_len) {
    return circular_len_ok(fft_len) ? circular_plan_bytes(fft_len) : 0;
}

size_t ath_circular_filter_plan_work_bytes(int width, int fft_len) {
    if (width < 1 || !circular_len_ok(fft_len)) return 0;
    return circular_plan_work_bytes(width, fft_len);
}

int ath_circular_filter_plan_build(const double* kernel2d, int width,
                                   int fft_len, void* plan, size_t plan_bytes,
                                   void* work, size_t work_bytes,
                                   void* stream) {
    clear_error();
    if (!kernel2d || !plan || !work)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (width < 1 || !circular_len_ok(fft_len) || fft_len < width)
        return set_error(ATH_ERR_INVALID_ARG,
                         "circular filter: fft_len must be a power of two in "
                         "[64, 32768] and at least the width");
    if (plan_bytes < circular_plan_bytes(fft_len) ||
        work_bytes < circular_plan_work_bytes(width, fft_len))
        return set_error(ATH_ERR_WORKSPACE,
                         "circular filter plan: buffer too small");
    return circular_plan_build(kernel2d, width, fft_len, plan, work,
                               static_cast<hipStream_t>(stream));
}

size_t ath_circular_ramp_filter_workspace_bytes(int ncol, int pad,
                                                int fft_len, int nimg) {
    if (ncol < 1 || pad < 0 || nimg < 1 || !circular_len_ok(fft_len))
        return 0;
    return circular_filter_work_bytes(ncol, pad, fft_len, nimg);
}

int ath_circular_ramp_filter(const float* images, int nimg, int ncol,
                             long long row_stride, long long img_stride,
                             int pad, int pad_mode, const void* plan,
                             int fft_len, void* workspace,
                             size_t workspace_bytes, void* out, int out_dtype,
                             long long out_row_stride,
                             long long out_img_stride, double mask_radius,
                             int scale_nrow, void* stream) {
    clear_error();
    if (!images || !plan || !workspace || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (nimg < 1 || ncol < 1 || pad < 0 || nimg > 65535)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    if (row_stride < ncol || out_row_stride < ncol)
        return set_error(ATH_ERR_INVALID_ARG, "row stride below ncol");
    if (pad_mode < ATH_PAD_EDGE || pad_mode > ATH_PAD_WRAP)
        return set_error(ATH_ERR_INVALID_ARG, "bad pad mode");
    if (out_dtype != ATH_F32 && out_dtype != ATH_F64)
        return set_error(ATH_ERR_INVALID_ARG,
                         "circular filter output must be float32 or float64");
    if (!circular_len_ok(fft_len) || fft_len < fft_length(ncol, pad))
        return set_error(ATH_ERR_INVALID_ARG,
                         "circular filter: fft_len must be a power of two, "
                         "at least ath_filter_fft_length(ncol, pad) and at "
                         "most 32768");
    if (workspace_bytes < circular_filter_work_bytes(ncol, pad, fft_len, nimg))
        return set_error(ATH_ERR_WORKSPACE,
                         "circular filter: workspace too small");
    if (scale_nrow == 1)
        return set_error(ATH_ERR_INVALID_ARG, "scale needs nrow >= 2");
    return circular_ramp_filter(images, nimg, ncol, row_stride, img_stride,
                                pad, pad_mode, plan, fft_len, workspace, out,
                                out_dtype, out_row_stride, out_img_stride,
                                mask_radius, scale_nrow,
                                static_cast<hipStream_t>(stream));
}

size_t ath_dfi_plan_bytes(int width) { return dfi_plan_bytes(width); }

int ath_dfi_plan_build(const double* window, int width, void* plan,
                       size_t plan_bytes, void* stream) {
    clear_error();
    const size_t need = dfi_plan_bytes(width);
    if (need == 0)
        return set_error(ATH_ERR_UNSUPPORTED,
                         "DFI: the padded width must be odd and at most 16383");
    if (!plan) return set_error(ATH_ERR_INVALID_ARG, "null plan pointer");
    if (plan_bytes < need)
        return set_error(ATH_ERR_WORKSPACE, "DFI plan buffer too small");
    return dfi_plan_build(window, width, plan,
                          static_cast<hipStream_t>(stream));
}

size_t ath_dfi_workspace_bytes(int nrow_in, int ncol, int nrow, int width) {
    if (nrow_in < 1 || ncol < 1 || nrow < 1 || width < 1) return 0;
    return dfi_workspace_bytes(nrow_in, ncol, nrow, width);
}

int ath_dfi_reconstruction(const void* sino, int in_dtype, int nrow_in,
                           int nslice, int ncol, long long row_stride,
                           long long slice_stride, double xshift, int pad,
                           int pad_mode, int apply_log, int angle_mode,
                           double row_shift, int flip, int nrow,
                           const void* plan, int width, void* workspace,
                           size_t workspace_bytes, void* out, int out_dtype,
                           long long out_row_stride,
                           long long out_slice_stride, double mask_r2,
                           int* flag, void* stream) {
    clear_error();
    if (!sino || !plan || !workspace || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (apply_log && !flag)
        return set_error(ATH_ERR_INVALID_ARG, "apply_log needs a flag pointer");
    if (in_dtype != ATH_F32 && in_dtype != ATH_F64)
        return set_error(ATH_ERR_INVALID_ARG,
                         "DFI input must be float32 or float64");
    if (out_dtype != ATH_F32 && out_dtype != ATH_F64)
        return set_error(ATH_ERR_INVALID_ARG,
                         "DFI output must be float32 or float64");
    if (nrow_in < 2 || nslice < 1 || ncol < 2 || pad < 0)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    if (row_stride < ncol || out_row_stride < ncol)
        return set_error(ATH_ERR_INVALID_ARG, "row stride below ncol");
    if (pad_mode < ATH_PAD_EDGE || pad_mode > ATH_PAD_WRAP)
        return set_error(ATH_ERR_INVALID_ARG, "bad pad mode");
    const int ncol1 = ncol + (ncol % 2 == 0 ? 1 : 0);
    if (width != ncol1 + 2 * pad || dfi_plan_bytes(width) == 0)
        return set_error(ATH_ERR_INVALID_ARG,
                         "DFI: width must be the odd width + 2*pad and at "
                         "most 16383");
    if (angle_mode == ATH_DFI_NO_ANGLES || angle_mode == ATH_DFI_ANGLES) {
        if (nrow != nrow_in)
            return set_error(ATH_ERR_INVALID_ARG, "nrow must equal nrow_in");
    } else if (angle_mode == ATH_DFI_FOLD_360) {
        if (nrow != nrow_in / 2 + 1)
            return set_error(ATH_ERR_INVALID_ARG,
                             "folded nrow must be nrow_in // 2 + 1");
    } else {
        return set_error(ATH_ERR_INVALID_ARG, "bad angle mode");
    }
    if (!(xshift == xshift) || !(row_shift == row_shift) ||
        fabs(xshift) > 1e6 || fabs(row_shift) > 1e9)
        return set_error(ATH_ERR_INVALID_ARG, "shift out of range");
    if (workspace_bytes < dfi_workspace_bytes(nrow_in, ncol, nrow, width))
        return set_error(ATH_ERR_WORKSPACE, "DFI: workspace too small");
    return dfi_reconstruction(sino, in_dtype, nrow_in, nslice, ncol,
                              row_stride, slice_stride, xshift, pad, pad_mode,
                              apply_log, angle_mode, row_shift, flip, nrow,
                              plan, width, workspace, out, out_dtype,
                              out_row_stride, out_slice_stride, mask_r2, flag,
                              static_cast<hipStream_t>(stream));
}

int ath_gaussian_filter_rows(const float* in, float* out, int nrows,
                             int ncol, const double* weights, int radius,
                             void* stream) {
    clear_error();
    if (!in || !out || !weights)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (nrows < 1 || ncol < 1 || radius < 0)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    return gaussian_rows(in, out, nrows, ncol, weights, radius,
                         static_cast<hipStream_t>(stream));
}

int ath_zoom_linear(const float* in, int nrow, int ncol, float* out,
                    int out_nrow, int out_ncol, void* stream) {
    clear_error();
    if (!in || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (nrow < 1 || ncol < 1 || out_nrow < 1 || out_ncol < 1)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    return zoom_linear(in, nrow, ncol, out, out_nrow, out_ncol,
                       static_cast<hipStream_t>(stream));
}

size_t ath_slice_metrics_workspace_bytes(int nimg, int nrows, int ncols,
                                         int metric) {
    if (nimg < 1 || nrows < 1 || ncols < 1 || metric < 0 || metric > 2 ||
        (long long)nrows * ncols > (1LL << 30))
        return 0;
    return slice_metrics_workspace(nimg, nrows, ncols, metric);
}

int ath_slice_metrics(const void* images, int dtype, int nimg, int nrows,
                      int ncols, long long stride, double nmin, double nmax,
                      int metric, double* out, void* workspace,
                      size_t workspace_bytes, void* stream) {
    clear_error();
    if (!images || !out)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (dtype != ATH_F32 && dtype != ATH_F64)
        return set_error(ATH_ERR_INVALID_ARG, "float32 or float64 images");
    if (nimg < 1 || nrows < 1 || ncols < 1 ||
        stride < (long long)nrows * ncols)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    if (nimg > 65535)
        return set_error(ATH_ERR_UNSUPPORTED, "more than 65535 images a call");
    if ((long long)nrows * ncols > (1LL << 30))
        return set_error(ATH_ERR_UNSUPPORTED, "image above 2^30 pixels");
    if (metric < 0 || metric > 2)
        return set_error(ATH_ERR_INVALID_ARG, "metric must be 0, 1 or 2");
    if (metric == 1 && (nrows < 2 || ncols < 2))
        return set_error(ATH_ERR_INVALID_ARG,
                         "gradient needs at least 2 pixels per axis");
    if (metric == 0 && !(nmax > nmin))
        return set_error(ATH_ERR_INVALID_ARG, "Empty image!!!");
    const size_t need = slice_metrics_workspace(nimg, nrows, ncols, metric);
    if (need > 0 && (!workspace || workspace_bytes < need))
        return set_error(ATH_ERR_WORKSPACE, "slice metrics: workspace too small");
    return slice_metrics(images, dtype, nimg, nrows, ncols, stride, nmin, nmax,
                         metric, out, workspace,
                         static_cast<hipStream_t>(stream));
}

int ath_gaussian_filter_axis(const float* in, float* out, int nouter,
                             int length, int ninner, const double* weights,
                             int radius, void* stream) {
    clear_error();
    if (!in || !out || !weights)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    if (nouter < 1 || length < 1 || ninner < 1 || radius < 0)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry");
    if (nouter > 65535 && ninner > 1)
        return set_error(ATH_ERR_UNSUPPORTED, "more than 65535 outer lines");
    return gaussian_axis(in, out, nouter, length, ninner, weights, radius,
                         static_cast<hipStream_t>(stream));
}

size_t ath_fbp_workspace_bytes(int nrow, int nslice, int ncol, int pad,
                               int nwin, int filt_dtype) {
    if (nrow < 1 || nslice < 1 || ncol < 1 || pad < 0 || nwin < 1 ||
        !valid_dtype(filt_dtype))
        return 0;
    const int L = fft_length(ncol, pad);
    if (L < 0) return 0;
    const size_t plan = plan_bytes(L, nwin);
    const size_t plan_al = (plan + 255) & ~(size_t)255;
    return plan_al + (size_t)nrow * nslice * ncol * dtype_bytes(filt_dtype);
}

int ath_fbp_reconstruction(const void* sino, int in_dtype, int nrow,
                           int nslice, int ncol, const float* angles,
                           float center, double mask_radius,
                           const double* taps, int nwin, int pad,
                           int pad_mode, int apply_log, int acc,
                           int filt_dtype, void* workspace,
                           size_t workspace_bytes, float* recon,
                           void* stream) {
    clear_error();
    if (!workspace)
        return set_error(ATH_ERR_INVALID_ARG, "null device pointer");
    return ath::fbp_into_workspace(sino, in_dtype, nrow, nslice, ncol, angles,
                                   center, mask_radius, taps, nwin, pad,
                                   pad_mode, apply_log, acc, filt_dtype,
                                   workspace, workspace_bytes, recon,
                                   static_cast<hipStream_t>(stream));
}

int ath_fbp_reconstruction_noalloc(const void* sino, int in_dtype, int nrow,
                                   int nslice, int ncol, const float* angles,
                                   float center, double mask_radius,
                                   const double* taps, int nwin, int pad,
                                   int pad_mode, int apply_log, int acc,
                                   int filt_dtype, float* recon,
                                   void* stream) {
    clear_error();
    const size_t need = ath_fbp_workspace_bytes(nrow, nslice, ncol, pad, nwin,
                                                filt_dtype);
    if (need == 0)
        return set_error(ATH_ERR_INVALID_ARG, "bad geometry or dtype");
    if (need > ATH_FBP_NOALLOC_MAX_BYTES)
        return set_error(ATH_ERR_WORKSPACE,
                         "ath_fbp_reconstruction_noalloc: this geometry needs " +
                             std::to_string(need) +
                             " workspace bytes, above the fixed " +
                             std::to_string(ATH_FBP_NOALLOC_MAX_BYTES));
    void* ws = nullptr;
    const int st = ath::noalloc_scratch(&ws);
    if (st != ATH_OK) return st;
    return ath::fbp_into_workspace(sino, in_dtype, nrow, nslice, ncol, angles,
                                   center, mask_radius, taps, nwin, pad,
                                   pad_mode, apply_log, acc, filt_dtype, ws,
                                   ATH_FBP_NOALLOC_MAX_BYTES, recon,
                                   static_cast<hipStream_t>(stream));
}

}  // extern "C"
