"""The C ABI library without a GPU: loads, exports exactly what
include/algotom_hip.h declares, rejects bad arguments with status codes and
a message, and its host helpers compute what the Python layer computes."""

import ctypes
import os
import re

import numpy as np
import pytest

from algotom_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "algotom_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ath_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = nat.load()
    names = declared_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    bound = sorted(n for n, _, _ in nat.SIGNATURES)
    assert bound == names


def test_noalloc_entry_rejects_geometry_above_its_array():
    """The allocation-free fused entry refuses, before any device work, a
    geometry whose workspace exceeds its fixed array."""
    lib = nat.load()
    # 1801 x 8 x 2560 float32 filtered: 147 MB > 64 MiB
    need = lib.ath_fbp_workspace_bytes(1801, 8, 2560, 150, 1, nat.ATH_F32)
    assert need > nat.FBP_NOALLOC_MAX_BYTES
    st = lib.ath_fbp_reconstruction_noalloc(
        None, nat.ATH_F32, 1801, 8, 2560, None, 1280.0, -1.0, None, 1, 150,
        0, 1, nat.ACC_F32SEQ, nat.ATH_F32, None, None)
    assert st == nat.ATH_ERR_WORKSPACE
    status, msg = nat.last_error()
    assert status == st and "workspace bytes" in msg
    # just inside the array: 1801 x 2 x 1024 passes the size check
    assert lib.ath_fbp_workspace_bytes(1801, 2, 1024, 150, 1,
                                       nat.ATH_F32) <= nat.FBP_NOALLOC_MAX_BYTES


def test_version_and_errors_without_device():
    lib = nat.load()
    assert lib.ath_version() >= 100
    st = lib.ath_back_projection(None, 0, 1, 1, 1, None, 0.0, 0, 0, -1.0, 0,
                                 None, None)
    assert st == nat.ATH_ERR_INVALID_ARG
    status, msg = nat.last_error()
    assert status == st and "null" in msg
    with pytest.raises(ValueError):
        nat.check(st, "ath_back_projection")


def test_fft_length_rule():
    lib = nat.load()
    import kernel_model
    for ncol, pad in [(1, 0), (5, 0), (64, 9), (512, 76), (1024, 150),
                      (2560, 150), (4096, 150), (7192, 150)]:
        assert lib.ath_filter_fft_length(ncol, pad) == \
            kernel_model.fft_length(ncol, pad)
    assert lib.ath_filter_fft_length(9000, 150) == 32768
    assert lib.ath_filter_fft_length(16159, 150) == 32768
    assert lib.ath_filter_fft_length(16160, 150) == nat.ATH_ERR_UNSUPPORTED
    assert lib.ath_filter_fft_length(0, 1) == nat.ATH_ERR_INVALID_ARG


def test_plan_and_workspace_sizes():
    lib = nat.load()
    assert lib.ath_filter_plan_bytes(8192, 1) == 16 * (8192 + 4097)
    assert lib.ath_filter_plan_bytes(100, 1) == 0
    ws = lib.ath_fbp_workspace_bytes(1801, 64, 2560, 150, 1, nat.ATH_F32)
    assert ws >= 1801 * 64 * 2560 * 4
    assert lib.ath_fbp_workspace_bytes(0, 1, 1, 0, 1, 0) == 0


def test_taps_from_window_matches_numpy():
    from algotom_amd.rec import reconstruction as rec
    lib = nat.load()
    for width, name in [(82, "hann"), (664, None), (101, "parzen")]:
        row = np.asarray(rec._ramp_row(width, name))
        re_ = np.ascontiguousarray(row.real)
        im_ = np.ascontiguousarray(row.imag)
        out = np.empty(width)
        st = lib.ath_filter_taps_from_window(
            re_.ctypes.data_as(ctypes.c_void_p),
            im_.ctypes.data_as(ctypes.c_void_p), width,
            out.ctypes.data_as(ctypes.c_void_p))
        assert st == nat.ATH_OK
        ref = rec._taps(row)
        assert np.abs(out - ref).max() <= 1e-15 * np.abs(ref).max() * width


def test_invalid_filter_arguments():
    lib = nat.load()
    dummy = ctypes.c_void_p(16)
    # fft length too short for the width
    st = lib.ath_ramp_filter(dummy, 0, 4, 1, 100, 10, 0, 0, dummy, 128, 1,
                             dummy, 0, None)
    assert st == nat.ATH_ERR_INVALID_ARG
    for mode in (-1, 10):  # pad modes are 0..9
        st = lib.ath_ramp_filter(dummy, 0, 4, 1, 100, 10, mode, 0, dummy, 256,
                                 1, dummy, 0, None)
        assert st == nat.ATH_ERR_INVALID_ARG  # bad pad mode
    st = lib.ath_ramp_filter(dummy, 2, 4, 1, 100, 10, 0, 0, dummy, 256, 1,
                             dummy, 0, None)
    assert st == nat.ATH_ERR_INVALID_ARG  # float16 input
    st = lib.ath_ramp_filter(dummy, 0, 4, 1, 100, 10, 0, 0, dummy, 256, 3,
                             dummy, 0, None)
    assert st == nat.ATH_ERR_INVALID_ARG  # nwin neither 1 nor nrow


def test_forward_projection_ws_rejects_small_workspace():
    # the size check comes before any device work
    lib = nat.load()
    dummy = ctypes.c_void_p(16)
    need = lib.ath_forward_projection_workspace_bytes(3, 50)
    assert need == 50 * 3 * 50 * 4
    assert lib.ath_forward_projection_workspace_bytes(0, 50) == 0
    for work, size in ((dummy, need - 1), (None, need)):
        st = lib.ath_forward_projection_ws(dummy, 10, 3, 50, dummy, 24.5,
                                           -1.0, 1.0, work, size, dummy, None)
        assert st == nat.ATH_ERR_WORKSPACE


def test_dfi_plan_sizes_cover_bluestein_32768():
    lib = nat.load()
    assert lib.ath_dfi_plan_bytes(8191) > 0       # length 16384
    assert lib.ath_dfi_plan_bytes(16383) > lib.ath_dfi_plan_bytes(8191)
    assert lib.ath_dfi_plan_bytes(16385) == 0     # length 65536
    assert lib.ath_dfi_plan_bytes(1000) == 0      # even width


def test_dropin_fails_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from algotom_amd.rec import reconstruction as rec
    with pytest.raises(RuntimeError):
        rec.fbp_reconstruction(np.ones((10, 20), np.float32), 5.0)
    with pytest.raises(ValueError):
        rec.fbp_reconstruction(np.ones((10, 20), np.float32), 20.0)
    with pytest.raises(ValueError):
        rec.fbp_reconstruction(np.ones((10, 20), np.float32), 5.0,
                               angles=np.zeros(3))


def test_every_module_imports_first_in_a_fresh_process():
    import pkgutil
    import subprocess
    import sys
    import algotom_amd
    names = ["algotom_amd." + m.name for m in pkgutil.walk_packages(
        algotom_amd.__path__, prefix="") if not m.name.startswith("_")]
    names += ["algotom_amd.rec." + n for n in ("center", "batch", "autograd",
                                              "volume", "reconstruction")]
    for name in sorted(set(names)):
        res = subprocess.run([sys.executable, "-c", "import " + name],
                             capture_output=True, text=True, cwd=ROOT)
        assert res.returncode == 0, name + ": " + res.stderr[-500:]


def test_slice_metrics_workspace_sizes():
    lib = nat.load()
    # entropy: 64 partial 256-bin histograms per 2560^2 image
    ws = lib.ath_slice_metrics_workspace_bytes
    assert ws(3, 2560, 2560, 0) == 3 * 64 * 1024
    assert ws(1, 100, 100, 0) == 1024
    # sharpness: one float per 8192-pixel chunk (2560^2 = 800 chunks)
    assert ws(2, 2560, 2560, 1) == 2 * 800 * 4
    assert ws(1, 97, 97, 1) == 2 * 4
    assert ws(1, 40, 61, 1) == 4
    # autocorrelation needs none; invalid arguments give 0
    assert ws(5, 40, 61, 2) == 0
    assert ws(0, 97, 97, 1) == 0
    assert ws(1, 97, 97, 3) == 0
