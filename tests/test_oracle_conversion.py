"""Pins the conversion oracle (CPU) with closed forms."""

import numpy as np
import pytest

from oracle import conversion_np as conv


@pytest.mark.parametrize("side", [0, 1])
def test_weights_sum_to_one_over_overlap(side):
    w1, w2 = conv.weight_matrices((3, 20), (3, 20), 7, side)
    if side == 1:
        s = w1[:, -7:].astype(np.float64) + w2[:, :7]
    else:
        s = w2[:, -7:].astype(np.float64) + w1[:, :7]
    assert np.abs(s - 1.0).max() < 1e-7
    assert w1.sum() + w2.sum() == pytest.approx(3 * (40 - 7), rel=1e-6)


@pytest.mark.parametrize("cor", [4.5, 14.5, (6, 0), (5, 1)])
def test_constant_input_stays_constant(cor):
    sino = np.full((21, 20), 0.7, dtype=np.float32)
    out, new_cor = conv.convert_360_to_180(sino, cor)
    assert np.abs(out - 0.7).max() < 1e-6
    ov = 2 * cor + 1 if not isinstance(cor, tuple) and cor <= 9.5 else (
        cor[0] if isinstance(cor, tuple) else 2 * (20 - cor) - 1)
    assert out.shape == (11, 40 - int(ov))
    assert new_cor == 20 - int(ov) / 2.0 - 0.5


def test_non_overlapping_parts_are_copies():
    rng = np.random.default_rng(0)
    sino = rng.uniform(0.5, 1.0, (31, 24)).astype(np.float32)
    ov = 6
    out, _ = conv.convert_360_to_180(sino, (ov, 1), norm=False)
    top = sino[:16]
    bot = np.fliplr(sino[-16:])
    assert np.array_equal(out[:, :24 - ov], top[:, :24 - ov])
    assert np.array_equal(out[:, 24:], bot[:, ov:])


@pytest.mark.parametrize("cor", [20.5, 44.5, (12, 0), (9, 1)])
def test_extend_closed_forms(cor):
    from oracle import conversion_np
    v = np.float32(0.75)
    s = np.full((6, 64), v, dtype=np.float32)
    out, new_cor = conversion_np.extend(s, cor, apply_log=False)
    if isinstance(cor, tuple):
        ov, side = cor
    else:
        side = 0 if cor <= 31.5 else 1
        ov = int(round(2 * cor + 1)) if side == 0 else int(round(2 * (64 - cor) - 1))
    assert out.shape == (6, 128 - ov)
    assert new_cor == 64 - ov / 2.0 - 0.5
    w = np.linspace(1.0, 0.0, ov) if side == 1 else np.linspace(0.0, 1.0, ov)
    ramp = 2 * np.float32(v * w)
    if side == 1:
        assert np.all(out[:, :64 - ov] == 2 * v)       # untouched columns
        assert np.array_equal(out[0, 64 - ov:64], ramp)
        assert np.all(out[:, 64:] == ramp[-1])         # edge padding
    else:
        assert np.all(out[:, 128 - ov - (64 - ov):][:, ov:] == 2 * v)
        assert np.array_equal(out[0, 64 - ov:64], ramp)
        assert np.all(out[:, :64 - ov] == ramp[0])


def test_flat_field_closed_forms():
    from oracle import conversion_np
    rng = np.random.default_rng(0)
    flat = rng.uniform(100, 200, (5, 7)).astype(np.float32)
    dark = np.full((5, 7), 10.0, np.float32)
    k = np.float32(0.25)
    proj = np.stack([(flat - dark) * k + dark] * 3)
    got = conversion_np.flat_field(proj, flat, dark)
    assert np.allclose(got, k, rtol=1e-6)
    # zeros of flat - dark are replaced by its mean
    flat2 = flat.copy()
    flat2[1, 2] = dark[1, 2]
    den = flat2 - dark
    m = np.mean(den)
    got = conversion_np.flat_field(proj, flat2, dark)
    assert np.allclose(got[:, 1, 2], (proj[:, 1, 2] - dark[1, 2]) / m)
    # ratio scales the flat; without dark the projection is divided only
    got = conversion_np.flat_field(proj, flat, dark, ratio=2.0,
                                   use_dark=False)
    assert np.allclose(got, proj / (2.0 * flat))


def _linear_rows(nrow, ncol, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.0, (nrow, 1))
    b = rng.uniform(-0.01, 0.01, (nrow, 1))
    return a, b, (a + b * np.arange(ncol)[None, :]).astype(np.float32)


@pytest.mark.parametrize("cor,side,sub", [(20.3, 0, -0.4), (22.15, 0, 0.3),
                                          (49.2, 1, -0.4), (47.6, 1, 0.8 - 1)])
def test_subpixel_shift_of_linear_rows(cor, side, sub):
    # conversion.py:107-117: both images move by the sub-pixel remainder in
    # opposite directions (scipy's shift: out[j] = in(j - shift)); a cubic
    # spline reproduces a linear row exactly away from the two edges, so
    # the stitched interior is the closed form below (the edge effect of
    # mode "nearest" decays as 0.268^k: 10 columns from an edge it is
    # below 1e-9)
    nrow, ncol = 41, 80
    a, b, sino = _linear_rows(nrow, ncol, 5)
    out, new_cor = conv.convert_360_to_180(sino, cor, norm=False)
    overlap = 2 * cor + 1 if side == 0 else 2 * (ncol - cor) - 1
    ov = int(np.round(overlap + 1e-3))
    assert abs((overlap - ov) - sub) < 1e-9
    assert new_cor == ncol - ov / 2.0 - 0.5
    n180 = nrow // 2 + 1
    j = np.arange(ncol, dtype=np.float64)
    s1 = sub if side == 1 else -sub
    top = a[:n180] + b[:n180] * (j - s1)[None, :]
    # image 2 = fliplr(bottom rows); its column m holds bottom column
    # ncol - 1 - m, and it moves by -s1
    bot = a[-n180:] + b[-n180:] * (ncol - 1 - (j + s1))[None, :]
    w1, w2 = conv.weight_matrices((n180, ncol), (n180, ncol), ov, side)
    width = 2 * ncol - ov
    exp = np.zeros((n180, width))
    cover = np.zeros(width, bool)
    if side == 1:
        exp[:, :ncol] = top * w1
        exp[:, ncol - ov:] += bot * w2
        keep1 = (j >= 10) & (j < ncol - 10)
        cover[:ncol] |= ~keep1
        cover[ncol - ov:] |= ~keep1
    else:
        exp[:, :ncol] = bot * w2
        exp[:, ncol - ov:] += top * w1
        keep = (j >= 10) & (j < ncol - 10)
        cover[:ncol] |= ~keep
        cover[ncol - ov:] |= ~keep
    interior = ~cover
    assert interior.sum() > 20
    assert np.abs(out[:, interior] - exp[:, interior]).max() < 2e-6


@pytest.mark.parametrize("cor", [10.3, 29.2])
def test_subpixel_constant_and_continuity(cor):
    sino = np.full((21, 40), 0.6, dtype=np.float32)
    out, _ = conv.convert_360_to_180(sino, cor)
    assert np.abs(out - 0.6).max() < 1e-6
    # a vanishing fraction gives the integer-overlap result
    rng = np.random.default_rng(2)
    s = rng.uniform(0.5, 1.0, (21, 40)).astype(np.float32)
    side = 0 if cor < 19.5 else 1
    ov = int(np.round((2 * cor + 1 if side == 0 else 2 * (40 - cor) - 1)))
    near, _ = conv.convert_360_to_180(s, (ov + 1e-7, side), norm=False)
    exact, _ = conv.convert_360_to_180(s, (ov, side), norm=False)
    assert np.abs(near - exact).max() < 1e-5


def test_custom_weights_are_used():
    rng = np.random.default_rng(4)
    s = rng.uniform(0.5, 1.0, (21, 30)).astype(np.float32)
    n180 = 11
    w1 = np.full((n180, 30), 0.25)
    w2 = np.full((n180, 30), 0.75)
    out, _ = conv.convert_360_to_180(s, (6, 1), norm=False, wei_mat1=w1,
                                     wei_mat2=w2)
    top = s[:n180].astype(np.float64)
    bot = np.fliplr(s[-n180:]).astype(np.float64)
    assert np.allclose(out[:, :24], 0.25 * top[:, :24], rtol=1e-6)
    assert np.allclose(out[:, 24:30], 0.25 * top[:, 24:] + 0.75 * bot[:, :6],
                       rtol=1e-6)
    assert np.allclose(out[:, 30:], 0.75 * bot[:, 6:], rtol=1e-6)


@pytest.mark.parametrize("cor,side", [(20.3, 0), (43.65, 1)])
def test_extend_subpixel_linear(cor, side):
    nrow, ncol = 8, 64
    a, b, s = _linear_rows(nrow, ncol, 7)
    out, _ = conv.extend(s, cor, apply_log=False)
    overlap = 2 * cor + 1 if side == 0 else 2 * (ncol - cor) - 1
    ov = int(np.round(overlap + 1e-3))
    sub = overlap - ov
    j = np.arange(ncol, dtype=np.float64)
    sh = sub if side == 1 else -sub
    v = a + b * (j - sh)[None, :]
    if side == 1:
        v[:, -ov:] *= np.linspace(1.0, 0.0, ov)[None, :]
        exp = 2 * np.pad(v, ((0, 0), (0, ncol - ov)), mode="edge")
        interior = np.zeros(2 * ncol - ov, bool)
        interior[10:ncol - 10] = True
    else:
        v[:, :ov] *= np.linspace(0.0, 1.0, ov)[None, :]
        exp = 2 * np.pad(v, ((0, 0), (ncol - ov, 0)), mode="edge")
        interior = np.zeros(2 * ncol - ov, bool)
        interior[ncol - ov + 10:2 * ncol - ov - 10] = True
    assert np.abs(out[:, interior] - exp[:, interior]).max() < 2e-6
