"""BPF oracle (oracle/bpf_np.py) pinned by an independent window
construction, brute-force 2D circular convolution, the tiled-kernel
identity the GPU filter uses, and the reference test's thresholds."""

import numpy as np
import pytest

from oracle import bpf_np, fbp_np


@pytest.mark.parametrize("width", [7, 12, 64, 65, 100])
@pytest.mark.parametrize("name", [None, "hann", "parzen"])
def test_circular_window_vs_map_coordinates(width, name):
    from algotom_amd.util import utility
    from algotom_amd.rec import reconstruction as rec
    ref = utility.transform_1d_window_to_2d(
        bpf_np.ramp_spectrum_abs(width, name), order=1, mode="nearest")
    got = bpf_np.circular_window(width, name)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    # the drop-in's host window (scipy's windows) agrees too
    dropin = rec.make_circular_ramp_window(width, name)
    assert np.abs(dropin - got).max() <= 1e-14 * np.abs(got).max()
    # centre value is the 1d spectrum at its centre (r = 0)
    n = width + 1 if width % 2 == 0 else width
    c = n // 2
    one_d = bpf_np.ramp_spectrum_abs(width, name)
    assert got[c, c] == one_d[c]


def _brute_circular(img, g2, pad, mode):
    p = np.pad(np.asarray(img, np.float64), pad, mode=mode)
    w = p.shape[0]
    out = np.zeros_like(p)
    for y in range(w):
        for x in range(w):
            acc = 0.0
            for u in range(w):
                for v in range(w):
                    acc += p[u, v] * g2[(y - u) % w, (x - v) % w]
            out[y, x] = acc
    n = img.shape[0]
    return out[pad:pad + n, pad:pad + n]


@pytest.mark.parametrize("mode", ["edge", "constant", "reflect",
                                  "symmetric", "wrap"])
def test_filter_is_circular_convolution(mode):
    rng = np.random.default_rng(2)
    img = rng.standard_normal((5, 5)).astype(np.float32)
    pad = 2
    win = bpf_np.circular_window(9, "hann")
    g2 = bpf_np.circular_kernel(win)
    ref = _brute_circular(img, g2, pad, mode)
    got = bpf_np.circular_filter(img, win, pad, mode)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("ncol,pad", [(6, 1), (20, 3), (33, 4), (64, 9)])
def test_tiled_kernel_linear_convolution_identity(ncol, pad):
    """The GPU evaluation: linear convolution with g2 tiled 2 x 2 on an
    L x L grid (L the 1D filter's FFT length, kernel samples beyond L
    dropped), outputs at (W+pad+y, W+pad+x)."""
    from tests import kernel_model as km
    rng = np.random.default_rng(ncol)
    img = rng.standard_normal((ncol, ncol))
    w = ncol + 2 * pad
    win = bpf_np.circular_window(w, "hann")
    g2 = bpf_np.circular_kernel(win)
    L = km.fft_length(ncol, pad)
    K = np.zeros((L, L))
    n = min(2 * w, L)
    idx = np.arange(n) % w
    K[:n, :n] = g2[np.ix_(idx, idx)]
    P = np.zeros((L, L))
    P[:w, :w] = np.pad(img, pad, mode="edge")
    Z = np.real(np.fft.ifft2(np.fft.fft2(P) * np.fft.fft2(K)))
    got = Z[w + pad:w + pad + ncol, w + pad:w + pad + ncol]
    ref = bpf_np.circular_filter(img, win, pad, "edge")
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_reference_thresholds_on_oracle():
    """tests/test_rec/test_reconstruction.py:146-202 on the oracle."""
    from __graft_entry__ import _reference_test_case
    mat, sino, ang, center = _reference_test_case()
    r = bpf_np.bpf(sino, center, apply_log=False, gpu=False)
    assert r.dtype == np.float64
    assert np.median(np.abs(mat - r)) <= 0.2
    r = bpf_np.bpf(sino, center, apply_log=False, gpu=True)
    assert np.max(np.abs(mat - r)) <= 0.1
    stack = np.pad(np.expand_dims(sino, 1), ((0, 0), (2, 2), (0, 0)),
                   mode="edge")
    rs = bpf_np.bpf(stack, center, np.deg2rad(ang), apply_log=False,
                    gpu=False, ratio=0.0)
    assert rs.dtype == np.float32 and rs.shape[1] == 5
    assert np.median(np.abs(mat - rs[:, 0, :])) <= 0.2
    rs = bpf_np.bpf(stack, center, apply_log=False, gpu=True, ratio=0.0)
    assert np.max(np.abs(mat - rs[:, 0, :])) <= 0.1
