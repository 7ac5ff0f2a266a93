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
