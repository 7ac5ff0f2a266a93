"""GPU parity of the ramp-filter kernel against the oracle.

Bar: float64 output within 1e-12 * max|ref| of the oracle's FP64 filter
(oracle/fbp_np.py ramp_filter, itself pinned by brute-force convolution);
float32 output within one float32 ulp of the rounded oracle.  Covers every
in-kernel pad mode, all windows, explicit pads (0, 1, large), per-row
windows, apply_log, float64 input and every FFT length 64..32768 (32768:
the two-half path, widths 7968..16384 at the default pads).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rows(nrow, nslice, ncol, seed, positive=True):
    rng = np.random.default_rng(seed)
    lo = 0.05 if positive else -1.0
    return rng.uniform(lo, 1.0, (nrow, nslice, ncol)).astype(np.float32)


def _gpu_filter(sino, pad, pad_mode, apply_log, taps, out_dtype):
    import torch
    from algotom_amd import _native as nat
    from algotom_amd.rec import reconstruction as rec
    dev = nat.device()
    t = torch.from_numpy(np.ascontiguousarray(sino)).to(dev)
    out = rec._filter_stack(t, pad, pad_mode, apply_log, taps, out_dtype)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _oracle(sino, pad, pad_mode, apply_log, ramp_rows):
    from oracle import fbp_np
    x = np.asarray(sino)
    if apply_log:
        x = (-np.log(x.astype(np.float64))).astype(x.dtype)
    return fbp_np.ramp_filter(np.moveaxis(x, 1, 0), ramp_rows, None, pad,
                              pad_mode).transpose(1, 0, 2)


@pytest.mark.parametrize("ncol", [5, 64, 65, 512, 1024, 2560, 4096, 7192,
                                  7967])
def test_filter_f64_widths(ncol):
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    pad = min(int(0.15 * ncol), 150)
    nrow = 7 if ncol > 4000 else 11
    sino = _rows(nrow, 2, ncol, ncol)
    win = fbp_np.ramp_window_row(ncol + 2 * pad, "hann")
    taps = rec._taps(win)[None, :]
    print("ncol", ncol, "pad", pad, flush=True)
    got = _gpu_filter(sino, pad, "edge", False, taps, torch.float64)
    ref = _oracle(sino, pad, "edge", False, win)
    err = np.abs(got - ref).max()
    assert err <= 1e-12 * np.abs(ref).max(), err


@pytest.mark.parametrize("mode", ["edge", "constant", "reflect", "symmetric",
                                  "wrap"])
@pytest.mark.parametrize("pad", [0, 1, 37, 200])
def test_filter_pad_modes(mode, pad):
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    ncol = 90
    sino = _rows(6, 3, ncol, pad)
    win = fbp_np.ramp_window_row(ncol + 2 * pad, "hamming")
    got = _gpu_filter(sino, pad, mode, False, rec._taps(win)[None, :],
                      torch.float64)
    ref = _oracle(sino, pad, mode, False, win)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("mode", ["linear_ramp", "maximum", "mean",
                                  "median", "minimum"])
@pytest.mark.parametrize("ncol,pad", [(1, 3), (2, 5), (7, 2), (90, 37),
                                      (91, 200), (2560, 150)])
@pytest.mark.parametrize("dtype,apply_log", [("float32", False),
                                             ("float32", True),
                                             ("float64", False)])
def test_filter_value_pad_modes(mode, ncol, pad, dtype, apply_log):
    """numpy.pad's value modes, computed per (logged) row in the kernel.
    One slice's rows are quantised to steps of 1/8 so the median meets
    ties.  Bar: 1e-12 * max|ref|; a float32 mean may sit one float32 ulp
    from numpy's pairwise float32 mean, so that case gets 1e-6."""
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    sino = _rows(5, 3, ncol, 31 * ncol + pad, positive=apply_log)
    sino[:, 1, :] = np.round(sino[:, 1, :] * 8.0) / 8.0
    if apply_log:
        sino[:, 1, :] = np.maximum(sino[:, 1, :], 0.125)
    sino = sino.astype(dtype)
    win = fbp_np.ramp_window_row(ncol + 2 * pad, "hann")
    got = _gpu_filter(sino, pad, mode, apply_log, rec._taps(win)[None, :],
                      torch.float64)
    ref = _oracle(sino, pad, mode, apply_log, win)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    bound = 1e-6 if (mode == "mean" and dtype == "float32") else 1e-12
    print("%s ncol %d pad %d %s log=%s: rel err %.3g" % (
        mode, ncol, pad, dtype, apply_log, err))
    assert err <= bound


@pytest.mark.parametrize("name", [None, "hann", "bartlett", "blackman",
                                  "hamming", "nuttall", "parzen", "triang"])
def test_filter_windows_and_log(name):
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    ncol, pad = 333, 49
    sino = _rows(9, 2, ncol, 3)
    win = fbp_np.ramp_window_row(ncol + 2 * pad, name)
    got = _gpu_filter(sino, pad, "edge", True, rec._taps(win)[None, :],
                      torch.float64)
    ref = _oracle(sino, pad, "edge", True, win)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_filter_per_row_windows():
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    ncol, pad, nrow = 120, 18, 8
    sino = _rows(nrow, 2, ncol, 4)
    names = [None, "hann", "hamming", "blackman", "parzen", "triang",
             "nuttall", "bartlett"]
    wins = np.stack([fbp_np.ramp_window_row(ncol + 2 * pad, n)
                     for n in names])
    got = _gpu_filter(sino, pad, "edge", False, rec._taps(wins),
                      torch.float64)
    for i in range(nrow):
        ref = _oracle(sino[i:i + 1], pad, "edge", False, wins[i])[0]
        assert np.abs(got[i] - ref).max() <= 1e-12 * np.abs(ref).max()


def test_filter_f32_output_within_one_ulp():
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    ncol = 2560
    pad = 150
    sino = _rows(40, 1, ncol, 5)
    win = fbp_np.ramp_window_row(ncol + 2 * pad, "hann")
    got = _gpu_filter(sino, pad, "edge", True, rec._taps(win)[None, :],
                      torch.float32)
    ref = _oracle(sino, pad, "edge", True, win).astype(np.float32)
    ulps = np.abs(got.view(np.int32).astype(np.int64)
                  - ref.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1
    assert (ulps != 0).mean() < 1e-3


def test_filter_f64_input():
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    ncol, pad = 200, 30
    sino = _rows(5, 2, ncol, 6).astype(np.float64) + 1e-9
    win = fbp_np.ramp_window_row(ncol + 2 * pad, "hann")
    got = _gpu_filter(sino, pad, "edge", True, rec._taps(win)[None, :],
                      torch.float64)
    ref = _oracle(sino, pad, "edge", True, win)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_apply_ramp_filter_dropin():
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    rng = np.random.default_rng(8)
    sino = rng.uniform(0, 1, (37, 65)).astype(np.float32)
    got = rec.apply_ramp_filter(sino)
    assert got.dtype == np.float64 and got.shape == sino.shape
    ref = fbp_np.ramp_filter(sino, None, None, None, "edge")
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # a window of the wrong shape is replaced (reconstruction.py:148-149)
    got2 = rec.apply_ramp_filter(sino, ramp_win=np.ones((3, 3)),
                                 filter_name="hann", pad=10)
    ref2 = fbp_np.ramp_filter(sino, None, "hann", 10, "edge")
    assert np.abs(got2 - ref2).max() <= 1e-12 * np.abs(ref2).max()


# FFT length 32768 (the two-half path): the first width past 16384 at the
# default pad, the widest at pad 150 and at pad 0, and every index-map mode
@pytest.mark.parametrize("ncol,pad,mode", [
    (7968, 150, "edge"), (8191, 150, "reflect"), (9000, 37, "constant"),
    (12000, 150, "wrap"), (16159, 150, "symmetric"), (16384, 0, "edge")])
def test_filter_fft_32768(ncol, pad, mode):
    import torch
    import kernel_model as km
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    assert km.fft_length(ncol, pad) == 32768
    sino = _rows(4, 2, ncol, ncol + pad)
    win = fbp_np.ramp_window_row(ncol + 2 * pad, "hann")
    print("ncol", ncol, "pad", pad, mode, flush=True)
    got = _gpu_filter(sino, pad, mode, True, rec._taps(win)[None, :],
                      torch.float64)
    ref = _oracle(sino, pad, mode, True, win)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("  rel err %.3g" % err)
    assert err <= 1e-12


@pytest.mark.parametrize("mode", ["linear_ramp", "maximum", "mean",
                                  "median", "minimum"])
def test_filter_fft_32768_value_modes(mode):
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    ncol, pad = 8500, 150
    sino = _rows(3, 2, ncol, 77)
    sino[:, 1, :] = np.maximum(np.round(sino[:, 1, :] * 8.0) / 8.0, 0.125)
    win = fbp_np.ramp_window_row(ncol + 2 * pad, "hann")
    got = _gpu_filter(sino, pad, mode, True, rec._taps(win)[None, :],
                      torch.float64)
    ref = _oracle(sino, pad, mode, True, win)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("%s: rel err %.3g" % (mode, err))
    assert err <= (1e-6 if mode == "mean" else 1e-12)


def test_filter_fft_32768_per_row_windows_and_f32():
    # per-row kernel spectra (the plan's row stride at M + 1 = 16385) and
    # float32 output within one ulp of the rounded oracle
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import reconstruction as rec
    ncol, pad, nrow = 8000, 150, 3
    sino = _rows(nrow, 2, ncol, 8)
    wins = np.stack([fbp_np.ramp_window_row(ncol + 2 * pad, n)
                     for n in (None, "hann", "parzen")])
    got = _gpu_filter(sino, pad, "edge", False, rec._taps(wins),
                      torch.float64)
    got32 = _gpu_filter(sino, pad, "edge", False, rec._taps(wins),
                        torch.float32)
    for i in range(nrow):
        ref = _oracle(sino[i:i + 1], pad, "edge", False, wins[i])[0]
        assert np.abs(got[i] - ref).max() <= 1e-12 * np.abs(ref).max()
        r32 = ref.astype(np.float32)
        ulps = np.abs(got32[i].view(np.int32).astype(np.int64)
                      - r32.view(np.int32).astype(np.int64))
        assert ulps.max() <= 1
