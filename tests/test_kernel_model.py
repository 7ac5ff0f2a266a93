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


@pytest.mark.parametrize("rec_bytes,sb,planes", [(8, 1, True), (8, 2, True),
                                                 (8, 4, True), (8, 4, False),
                                                 (4, 8, True), (4, 4, True),
                                                 (4, 8, False)])
def test_bp_record_planes_layout(rec_bytes, sb, planes):
    # staging writes every (angle, column, slice) to its own record of the
    # block; inner_angle reads back the record staged for column d - dlo;
    # each plane of a column starts on a 16-byte boundary when it holds more
    # than one record
    span, na = 74, 3
    p, npl = km.bp_rec_planes(rec_bytes, sb, planes)
    assert p * npl == sb
    if planes and sb * rec_bytes > 16:
        assert p * rec_bytes == 16 and npl > 1
    else:
        assert npl == 1
    seen = set()
    for al in range(na):
        for k in range(span):
            for js in range(sb):
                i = km.bp_rec_stage_index(al, k, js, span, sb, p)
                assert 0 <= i < na * span * sb
                seen.add(i)
                if js % p == 0 and p * rec_bytes == 16:
                    assert (i * rec_bytes) % 16 == 0
                for dlo in (-7, 0, 913):
                    assert km.bp_rec_read_index(al, dlo + k, dlo, js, span,
                                                sb, p) == i
    assert len(seen) == na * span * sb


@pytest.mark.parametrize("ncol,ratio", [(64, 1.0), (101, 0.8), (256, None),
                                        (300, 0.0)])
def test_fp_empty_chunks_hold_no_kept_pixel(ncol, ratio):
    # a chunk the kernel skips may not hold a pixel inside the image and the
    # mask, within one column of its staged window (the reach of a run that
    # leaves the window by a rounding step)
    rng = np.random.default_rng(ncol)
    mc = ncol // 2
    use_mask = ratio is not None
    r2 = (ratio * mc) ** 2 if use_mask else 0.0
    yy, xx = np.mgrid[:ncol, :ncol]
    kept = ((xx - mc) ** 2 + (yy - mc) ** 2 <= r2) if use_mask else \
        np.ones((ncol, ncol), bool)
    skipped = 0
    for _ in range(4000):
        m0 = int(rng.integers(0, ncol))
        m1 = min(m0 + int(rng.integers(1, 9)), ncol) - 1
        ka = int(rng.integers(-80, ncol + 40))
        kb = ka + int(rng.integers(2, 130))
        if not km.fp_chunk_is_empty(ncol, use_mask, r2, mc, m0, m1, ka, kb):
            continue
        skipped += 1
        lo, hi = max(ka - 1, 0), min(kb + 1, ncol - 1)
        if lo <= hi:
            assert not kept[m0:m1 + 1, lo:hi + 1].any()
    assert skipped > 200


@pytest.mark.parametrize("nrow,ncol,nslice,fsb", [(901, 1024, 16, 16),
                                                  (1801, 7192, 16, 16),
                                                  (5, 65, 3, 1),
                                                  (181, 300, 9, 8),
                                                  (7, 64, 1, 1)])
def test_fp_block_remap_is_a_bijection(nrow, ncol, nslice, fsb):
    # the 1-D launch covers every (bin, angle, slice) block exactly once,
    # and the ids one XCD runs are consecutive angle blocks
    nab, nbb = (nrow + 3) // 4, (ncol + 63) // 64
    nsb = (nslice + fsb - 1) // fsb
    total = nab * nbb * nsb
    seen = set()
    for p in range(total):
        b = km.fp_block(p, nrow, ncol, nslice, fsb)
        assert 0 <= b[0] < nbb and 0 <= b[1] < nab and 0 <= b[2] < nsb
        seen.add(b)
    assert len(seen) == total
    if total >= 16:
        b0 = km.fp_block(0, nrow, ncol, nslice, fsb)
        b8 = km.fp_block(8, nrow, ncol, nslice, fsb)
        assert b8[1] == (b0[1] + 1) % nab
