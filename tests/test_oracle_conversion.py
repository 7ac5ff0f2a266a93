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
