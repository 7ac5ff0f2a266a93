"""ctypes binding of libalgotom_hip.so (the C ABI in include/algotom_hip.h).

torch is imported before the library is opened so that both share the one
HIP runtime torch has loaded; torch supplies device memory and streams.
There is no fallback: if the library is missing or no HIP device is
visible, every entry raises.
"""

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libalgotom_hip.so")

ATH_OK = 0
ATH_ERR_INVALID_ARG = -1
ATH_ERR_HIP = -2
ATH_ERR_UNSUPPORTED = -3
ATH_ERR_NO_DEVICE = -4
ATH_ERR_WORKSPACE = -5
FBP_NOALLOC_MAX_BYTES = 64 << 20  # ATH_FBP_NOALLOC_MAX_BYTES

ATH_F32, ATH_F64, ATH_F16, ATH_BF16, ATH_U16 = 0, 1, 2, 3, 4
VERT_GPU, VERT_CPU = 0, 1
DFI_NO_ANGLES, DFI_ANGLES, DFI_FOLD_360 = 0, 1, 2
PAD_MODES = {"edge": 0, "constant": 1, "reflect": 2, "symmetric": 3,
             "wrap": 4}
# every numpy.pad mode the 1D ramp filter takes: the index-map modes above
# and the value modes (numpy's defaults: end value 0, whole-row statistics)
FILTER_PAD_MODES = dict(PAD_MODES, linear_ramp=5, maximum=6, mean=7,
                        median=8, minimum=9)
ACC_F64, ACC_F32SEQ, ACC_F32, ACC_DERIV, ACC_F64_DERIV = 0, 1, 2, 3, 4

# (name, restype, argtypes) of every exported entry point
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_F = ctypes.c_float
_S = ctypes.c_size_t
SIGNATURES = (
    ("ath_version", _I, []),
    ("ath_last_error", _I, [ctypes.c_char_p, _S]),
    ("ath_device_count", _I, [ctypes.POINTER(_I)]),
    ("ath_filter_fft_length", _I, [_I, _I]),
    ("ath_filter_plan_bytes", _S, [_I, _I]),
    ("ath_filter_plan_build", _I, [_P, _I, _I, _I, _P, _S, _P]),
    ("ath_filter_taps_from_window", _I, [_P, _P, _I, _P]),
    ("ath_ramp_filter", _I, [_P, _I, _I, _I, _I, _I, _I, _I, _P, _I, _I, _P,
                             _I, _P]),
    ("ath_back_projection", _I, [_P, _I, _I, _I, _I, _P, _F, _I, _I, _D, _I,
                                 _P, _P]),
    ("ath_back_projection_with_derivative", _I, [_P, _I, _I, _I, _P, _F,
                                                 _I, _D, _I, _P, _P, _P]),
    ("ath_back_projection_tiled", _I, [_P, _I, _I, _I, _I, _P, _F, _I, _I,
                                       _D, _I, _P, _I, _I, _P]),
    ("ath_bp_job_tiles", ctypes.c_longlong, [_I, _I, _I, _I]),
    ("ath_back_projection_batched", _I, [_P, _I, ctypes.c_longlong, _I, _I,
                                         _I, _I, _I, _P, _P]),
    ("ath_ramp_filter_batched", _I, [_P, _I, ctypes.c_longlong, _I, _I, _I,
                                     _I, _I, _P]),
    ("ath_back_projection_naive", _I, [_P, _I, _I, _I, _P, _F, _I, _I, _D,
                                       _I, _P, _P]),
    ("ath_forward_projection", _I, [_P, _I, _I, _I, _P, _F, _D, _D, _P, _P]),
    ("ath_forward_projection_workspace_bytes", _S, [_I, _I]),
    ("ath_forward_projection_ws", _I,
     [_P, _I, _I, _I, _P, _F, _D, _D, _P, _S, _P, _P]),
    ("ath_ramp_filter_adjoint", _I, [_P, _I, _I, _I, _I, _I, _P, _I, _I, _P,
                                     _P]),
    ("ath_conversion_workspace_bytes", _S, [_I, _I, _I, _I, _D]),
    ("ath_convert_sinogram_360_to_180", _I, [_P, _I, _I, _I, _D, _I, _I, _I,
                                             _P, _P, _I, _P, _S, _P, _P]),
    ("ath_minus_log", _I, [_P, _P, ctypes.c_longlong, _P]),
    ("ath_extend_sinogram", _I, [_P, _I, _I, _I, _D, _I, _I, _P, _S, _P,
                                 _P]),
    ("ath_flat_field_correction", _I, [_P, _I, ctypes.c_longlong, _I, _I,
                                       ctypes.c_longlong, ctypes.c_longlong,
                                       _P, _P, _I, _P, _P]),
    ("ath_flat_field_log", _I, [_P, _I, _I, _I, _I, ctypes.c_longlong,
                                ctypes.c_longlong, _P, _P, _I, _P, _P, _P]),
    ("ath_vertical_back_projection", _I, [_P, _I, _I, _I, _P, _P, _P, _I, _D,
                                          _P, _I, _I, _I, _I, _P, _P]),
    ("ath_vertical_finish", _I, [_P, _I, ctypes.c_longlong, _I, _I, _I, _I,
                                 _P]),
    ("ath_circular_filter_plan_bytes", _S, [_I]),
    ("ath_circular_filter_plan_work_bytes", _S, [_I, _I]),
    ("ath_circular_filter_plan_build", _I, [_P, _I, _I, _P, _S, _P, _S, _P]),
    ("ath_circular_ramp_filter_workspace_bytes", _S, [_I, _I, _I, _I]),
    ("ath_circular_ramp_filter", _I, [_P, _I, _I, ctypes.c_longlong,
                                      ctypes.c_longlong, _I, _I, _P, _I, _P,
                                      _S, _P, _I, ctypes.c_longlong,
                                      ctypes.c_longlong, _D, _I, _P]),
    ("ath_dfi_plan_bytes", _S, [_I]),
    ("ath_dfi_plan_build", _I, [_P, _I, _P, _S, _P]),
    ("ath_dfi_workspace_bytes", _S, [_I, _I, _I, _I]),
    ("ath_dfi_reconstruction", _I, [_P, _I, _I, _I, _I, ctypes.c_longlong,
                                    ctypes.c_longlong, _D, _I, _I, _I, _I,
                                    _D, _I, _I, _P, _I, _P, _S, _P, _I,
                                    ctypes.c_longlong, ctypes.c_longlong, _D,
                                    _P, _P]),
    ("ath_gaussian_filter_rows", _I, [_P, _P, _I, _I, _P, _I, _P]),
    ("ath_zoom_linear", _I, [_P, _I, _I, _P, _I, _I, _P]),
    ("ath_slice_metrics", _I, [_P, _I, _I, _I, _I, ctypes.c_longlong, _D, _D,
                               _I, _P, _P, _S, _P]),
    ("ath_slice_metrics_workspace_bytes", _S, [_I, _I, _I, _I]),
    ("ath_gaussian_filter_axis", _I, [_P, _P, _I, _I, _I, _P, _I, _P]),
    ("ath_fbp_workspace_bytes", _S, [_I, _I, _I, _I, _I, _I]),
    ("ath_fbp_reconstruction", _I, [_P, _I, _I, _I, _I, _P, _F, _D, _P, _I,
                                    _I, _I, _I, _I, _I, _P, _S, _P, _P]),
    ("ath_fbp_reconstruction_noalloc", _I, [_P, _I, _I, _I, _I, _P, _F, _D,
                                            _P, _I, _I, _I, _I, _I, _I, _P,
                                            _P]),
)



class FilterJob(ctypes.Structure):
    """ath_filter_job (include/algotom_hip.h)."""
    _fields_ = [("sino", _P), ("out", _P), ("plan", _P), ("nrow", _I),
                ("nslice", _I), ("ncol", _I), ("pad", _I), ("nwin", _I),
                ("reserved", _I), ("row_begin", ctypes.c_longlong)]


class BpJob(ctypes.Structure):
    """ath_bp_job (include/algotom_hip.h)."""
    _fields_ = [("sino", _P), ("recon", _P), ("angles", _P), ("nrow", _I),
                ("nslice", _I), ("ncol", _I), ("center", _F),
                ("mask_radius", _D), ("apply_scale", _I), ("edge_pad", _I),
                ("tile_begin", ctypes.c_longlong)]


_lib = None
_lock = threading.Lock()


class NativeError(RuntimeError):
    """A failure reported by libalgotom_hip."""


def load():
    """Open libalgotom_hip.so (once); raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            import torch  # noqa: F401  (shares its HIP runtime)
            if not os.path.exists(LIB_PATH):
                raise NativeError(
                    "libalgotom_hip.so is not built (expected at %s); run "
                    "python -m algotom_amd.build" % LIB_PATH)
            lib = ctypes.CDLL(LIB_PATH)
            for name, res, args in SIGNATURES:
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    status = load().ath_last_error(buf, len(buf))
    return status, buf.value.decode(errors="replace")


def check(status, what):
    """Map a negative status to ValueError (bad argument) or NativeError."""
    if status >= 0:
        return status
    _, msg = last_error()
    text = "%s failed (status %d): %s" % (what, status, msg)
    if status == ATH_ERR_INVALID_ARG:
        raise ValueError(text)
    raise NativeError(text)


def device():
    """The current HIP device as a torch.device; raises if there is none."""
    import torch
    if not torch.cuda.is_available():
        raise NativeError("no HIP device is visible: the algotom_amd FBP path "
                          "runs only on an AMD GPU (there is no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(dev=None):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def dtype_code(dtype):
    import torch
    codes = {torch.float32: ATH_F32, torch.float64: ATH_F64,
             torch.float16: ATH_F16, torch.bfloat16: ATH_BF16}
    if dtype not in codes:
        raise ValueError("unsupported element type %s" % dtype)
    return codes[dtype]
