"""Pins the oracle's windows, ramp and filter (CPU).

The reference's tests only check that windows differ from ones
(tests/test_rec/test_reconstruction.py:71-86); these checks pin the oracle
against scipy's window vectors (the library the reference calls), closed
forms, an O(W^2) brute-force circular convolution, the reference's literal
fftshift chain, and invariants a dropped term or shifted index would break.
"""

import os

import numpy as np
import pytest
from scipy.signal import windows as sw

from oracle import fbp_np, port

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
NAMES = ["hann", "bartlett", "blackman", "hamming", "nuttall", "parzen",
         "triang"]


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("width", [1, 2, 5, 64, 65, 664, 2860])
def test_windows_closed_form_vs_scipy(name, width):
    ref = getattr(sw, name)(width)
    got = fbp_np.smoothing_window(name, width)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-14


def test_windows_golden_fixture():
    data = np.load(os.path.join(GOLDEN, "windows.npz"))
    for key in data.files:
        name, width = key.rsplit("_", 1)
        got = fbp_np.smoothing_window(name, int(width))
        assert np.abs(got - data[key]).max() <= 1e-14, key


def test_unknown_window_is_ones():
    assert np.array_equal(fbp_np.smoothing_window(None, 7), np.ones(7))
    assert np.array_equal(fbp_np.smoothing_window("foo", 7), np.ones(7))
    for name in NAMES:  # reference test: each window differs from ones
        assert np.mean(np.abs(fbp_np.smoothing_window(name, 64) - 1)) > 0


@pytest.mark.parametrize("width", [7, 8, 82, 664, 2860])
def test_ramlak_closed_form(width):
    h = fbp_np.ramlak_kernel(width)
    n = np.arange(width) - int(np.ceil((width - 1) / 2))
    for k, nk in enumerate(n):
        if nk == 0:
            assert h[k] == 0.25
        elif nk % 2 == 0:
            assert h[k] == 0.0
        else:
            assert h[k] == -1.0 / (nk * np.pi) ** 2
    # sum over all odd n of 1/(n pi)^2 is 1/4 (pi^2/8 identity), so the DC
    # gain equals the part of that series outside the kernel's support
    lo, hi = int(n[0]), int(n[-1])
    odd = np.arange(1, 2_000_001, 2, dtype=np.float64)
    tail = (np.sum(1.0 / (odd[odd > hi] * np.pi) ** 2)
            + np.sum(1.0 / (odd[odd > -lo] * np.pi) ** 2))
    assert abs(h.sum() - tail) < 1e-6


def test_ramp_row_identities():
    width = 664
    r = fbp_np.ramp_window_row(width, None)
    h = fbp_np.ramlak_kernel(width)
    assert np.allclose(np.fft.ifftshift(r), np.fft.fft(h), atol=0, rtol=0)
    # the Ram-Lak spectrum is close to |frequency| (Kak & Slaney): at the
    # Nyquist bin it is ~1/2, and real to rounding up to the linear phase
    mag = np.abs(np.fft.ifftshift(r))
    assert abs(mag[width // 2] - 0.5) < 1e-3
    assert mag[0] < 1e-3


def brute_force_filter(row, ramp_row, pad, mode):
    """O(W^2) circular convolution of the padded row with Re(g)."""
    p = np.pad(row.astype(np.float64), (pad, pad), mode=mode)
    w = p.size
    g = np.real(np.fft.ifftshift(np.fft.ifft(np.fft.ifftshift(ramp_row))))
    out = np.array([sum(p[k] * g[(n - k) % w] for k in range(w))
                    for n in range(w)])
    return out[pad:pad + row.size]


@pytest.mark.parametrize("mode", ["edge", "constant", "reflect", "symmetric",
                                  "wrap"])
@pytest.mark.parametrize("ncol,pad", [(9, 3), (12, 0), (7, 11), (20, 4)])
def test_filter_vs_brute_force(mode, ncol, pad):
    rng = np.random.default_rng(ncol * 100 + pad)
    row = rng.uniform(0.1, 1.0, ncol).astype(np.float32)
    r = fbp_np.ramp_window_row(ncol + 2 * pad, "hann")
    ref = brute_force_filter(row, r, pad, mode)
    got = fbp_np.ramp_filter(row[None], r, None, pad, mode)[0]
    lit = fbp_np.ramp_filter_spectral(row[None], r, pad, mode)[0]
    c = port.ramp_filter(row[None], fbp_np.filter_kernel(r), pad, mode)[0]
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-14 * scale
    assert np.abs(lit - ref).max() <= 1e-14 * scale
    assert np.abs(c - ref).max() <= 1e-14 * scale


@pytest.mark.parametrize("mode", ["linear_ramp", "maximum", "mean",
                                  "median", "minimum"])
@pytest.mark.parametrize("ncol,pad", [(9, 3), (12, 5), (7, 11), (1, 4)])
def test_value_pad_modes_vs_brute_force(mode, ncol, pad):
    """numpy.pad's value modes (the reference passes pad_mode to np.pad,
    reconstruction.py:147): padded in the row's float32 type, then the
    O(W^2) circular convolution; the numpy and C-port filters agree."""
    rng = np.random.default_rng(ncol * 100 + pad + 7)
    row = rng.uniform(-1.0, 1.0, ncol).astype(np.float32)
    r = fbp_np.ramp_window_row(ncol + 2 * pad, "hann")
    p = np.pad(row, (pad, pad), mode=mode).astype(np.float64)
    g = np.real(np.fft.ifftshift(np.fft.ifft(np.fft.ifftshift(r))))
    w = p.size
    ref = np.array([sum(p[k] * g[(n - k) % w] for k in range(w))
                    for n in range(w)])[pad:pad + ncol]
    got = fbp_np.ramp_filter(row[None], r, None, pad, mode)[0]
    c = port.ramp_filter(row[None], fbp_np.filter_kernel(r), pad, mode)[0]
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-14 * scale
    assert np.abs(c - ref).max() <= 1e-14 * scale


def test_value_pad_mode_closed_forms():
    """What numpy.pad's value modes put beside a float32 row: the row's
    max / min / median / mean on both sides, and linear ramps from 0 that
    stop one step short of each edge value."""
    row = np.array([3.0, -1.0, 7.0, 2.0, 2.0, 5.0], dtype=np.float32)
    pad = 4
    p = fbp_np.pad_rows(row[None], pad, "maximum")[0]
    assert p.dtype == np.float32
    assert (p[:pad] == 7.0).all() and (p[-pad:] == 7.0).all()
    assert (fbp_np.pad_rows(row[None], pad, "minimum")[0][:pad] == -1.0).all()
    # even length: mean of the two middle values (2 and 3)
    assert (fbp_np.pad_rows(row[None], pad, "median")[0][-pad:] == 2.5).all()
    assert (fbp_np.pad_rows(row[None], pad, "mean")[0][:pad]
            == np.float32(18.0 / 6.0)).all()
    lr = fbp_np.pad_rows(row[None], pad, "linear_ramp")[0]
    assert np.array_equal(lr[:pad], np.float32([0.0, 0.75, 1.5, 2.25]))
    assert np.array_equal(lr[-pad:], np.float32([3.75, 2.5, 1.25, 0.0]))
    assert np.array_equal(lr[pad:-pad], row)
    # the C port's log-then-pad order: pad values come from the logged row
    x = np.array([[0.5, 0.25, 0.125, 1.0, 0.75]], dtype=np.float32)
    g = fbp_np.filter_kernel(fbp_np.ramp_window_row(5 + 6, None))
    a = port.ramp_filter(x, g, 3, "maximum", apply_log=True)
    lx = (-np.log(x.astype(np.float64))).astype(np.float32)
    b = fbp_np.ramp_filter(lx, fbp_np.ramp_window_row(11, None), None, 3,
                           "maximum")
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


def test_pad_index_matches_numpy_pad():
    rng = np.random.default_rng(1)
    for n in (1, 2, 5, 9):
        row = rng.standard_normal(n)
        for pad in (0, 1, 4, 13):
            for mode in ("edge", "constant", "reflect", "symmetric", "wrap"):
                if mode == "reflect" and n == 1 and pad > 0:
                    continue  # numpy pads a 1-sample reflect with the sample
                assert np.array_equal(fbp_np.pad_rows(row, pad, mode),
                                      np.pad(row, pad, mode=mode))


def test_filter_linearity_and_dc():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((3, 100))
    b = rng.standard_normal((3, 100))
    fa = fbp_np.ramp_filter(a, None, "hann", 15, "wrap")
    fb = fbp_np.ramp_filter(b, None, "hann", 15, "wrap")
    fab = fbp_np.ramp_filter(2.0 * a - 3.0 * b, None, "hann", 15, "wrap")
    assert np.abs(fab - (2.0 * fa - 3.0 * fb)).max() < 1e-12
    # a constant row with edge padding is filtered to c * sum(g)
    c = fbp_np.ramp_filter(np.full((1, 64), 2.5), None, "hann", 9, "edge")
    g = fbp_np.filter_kernel(fbp_np.ramp_window_row(82, "hann"))
    assert np.abs(c - 2.5 * g.sum()).max() < 1e-13


@pytest.mark.parametrize("ncol", [64, 65, 512, 2560])
def test_port_filter_matches_numpy(ncol):
    rng = np.random.default_rng(ncol)
    pad = min(int(0.15 * ncol), 150)
    rows = rng.uniform(0.05, 1.0, (5, ncol)).astype(np.float32)
    r = fbp_np.ramp_window_row(ncol + 2 * pad, "hann")
    a = fbp_np.ramp_filter(rows, r, None, pad)
    b = port.ramp_filter(rows, fbp_np.filter_kernel(r), pad)
    assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()
    # apply_log inside the port equals -log in double rounded to float32
    al = fbp_np.ramp_filter((-np.log(rows.astype(np.float64))).astype(
        np.float32), r, None, pad)
    bl = port.ramp_filter(rows, fbp_np.filter_kernel(r), pad,
                          apply_log=True)
    assert np.abs(al - bl).max() <= 1e-14 * np.abs(al).max()


def test_numpy_log32_within_4ulp_of_rounded():
    """numpy's float32 log (what the reference calls) vs the correctly
    rounded value the oracle and kernel use: at most a few ulp apart."""
    x = np.random.default_rng(3).uniform(1e-3, 2.0, 200000).astype(np.float32)
    a = (-np.log(x)).view(np.int32).astype(np.int64)
    b = (-np.log(x.astype(np.float64))).astype(np.float32).view(
        np.int32).astype(np.int64)
    assert np.abs(a - b).max() <= 4
