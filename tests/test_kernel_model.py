"""CPU checks of the kernels' index arithmetic through tests/kernel_model.py
(pass structure, table indices, real split/merge, staging bounds) against
numpy's FFT and the oracle."""

import numpy as np
import pytest

import kernel_model as km
from oracle import fbp_np


@pytest.mark.parametrize("M", [32, 64, 128, 256, 512, 1024, 2048, 4096,
                               8192])
@pytest.mark.parametrize("direction", [-1, 1])
def test_stockham_fft(M, direction):
    rng = np.random.default_rng(M)
    x = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    got = km.lds_fft(x, km.twiddles(2 * M), direction)
    ref = np.fft.fft(x) if direction < 0 else np.fft.ifft(x) * M
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("ncol", [1, 5, 64, 65, 512, 1024, 2560])
def test_filter_row_model_vs_oracle(ncol):
    rng = np.random.default_rng(ncol)
    pad = min(int(0.15 * ncol), 150)
    width = ncol + 2 * pad
    L = km.fft_length(ncol, pad)
    row = rng.uniform(0.1, 1.0, ncol)
    r = fbp_np.ramp_window_row(width, "hann")
    got = km.ramp_filter_row(fbp_np.pad_rows(row, pad, "edge"),
                             fbp_np.filter_kernel(r), pad, L)
    ref = fbp_np.ramp_filter(row[None], r, None, pad)[0]
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_kernel_spectrum_model_vs_rfft():
    g = np.random.default_rng(3).standard_normal(82)
    L = 256
    got = km.kernel_spectrum(g, L)
    ref = np.fft.rfft(np.concatenate([g, g, np.zeros(L - 164)])) / (L // 2)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def test_fft_length_is_wraparound_free():
    # outputs W+pad .. W+pad+ncol-1 of the linear convolution of a length-W
    # row with the length-2W doubled kernel must not alias on length L
    for ncol in range(1, 300, 7):
        for pad in (0, 1, min(int(0.15 * ncol), 150), ncol):
            W = ncol + 2 * pad
            L = km.fft_length(ncol, pad)
            assert L >= W + pad + ncol           # last output index < L
            assert (W + pad) + L > 3 * W - 2     # first aliased index beyond


@pytest.mark.parametrize("ncol,nang,center", [(65, 37, 32.0), (200, 90, 3.0),
                                              (257, 45, 200.7),
                                              (512, 60, 255.5)])
def test_bp_staging_span_bound(ncol, nang, center):
    rng = np.random.default_rng(ncol)
    ang = np.sort(rng.uniform(0, 2 * np.pi, nang))
    worst, span = km.bp_tile_footprints(ncol, ang,
                                        np.float64(np.float32(center)))
    assert worst <= span - 2
