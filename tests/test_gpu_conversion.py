"""GPU 360 -> 180 conversion against the oracle restatement.

Bar: bit-identical without normalisation; with normalisation the only
difference is the float32 mean (numpy pairwise float32 sum vs float64 sum
here), so within 1e-6 relative.  Both sides, centre form and explicit
(overlap, side), total_width extension, 2D and stacks.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cor", [20.5, 77.5, (31, 0), (17, 1), 0.0])
@pytest.mark.parametrize("norm", [False, True])
def test_convert_360_to_180_matches_oracle(cor, norm):
    from oracle import conversion_np as oconv
    from algotom_amd.prep import conversion
    rng = np.random.default_rng(7)
    sino = rng.uniform(0.2, 1.0, (61, 100)).astype(np.float32)
    got, gc = conversion.convert_sinogram_360_to_180(sino, cor, norm=norm)
    ref, rc = oconv.convert_360_to_180(sino, cor, norm=norm)
    assert gc == rc and got.shape == ref.shape
    if norm:
        assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()
    else:
        assert np.array_equal(got, ref)


def test_convert_stack_and_total_width():
    from oracle import conversion_np as oconv
    from algotom_amd.prep import conversion
    rng = np.random.default_rng(8)
    stack = rng.uniform(0.2, 1.0, (40, 3, 50)).astype(np.float32)
    got, _ = conversion.convert_sinogram_360_to_180(stack, 9.5,
                                                     total_width=120)
    assert got.shape == (21, 3, 120)
    for j in range(3):
        ref, _ = oconv.convert_360_to_180(stack[:, j, :], 9.5,
                                          total_width=120)
        assert np.abs(got[:, j, :] - ref).max() <= 1e-6 * np.abs(ref).max()


def test_subpixel_overlap_is_rejected():
    from algotom_amd.prep import conversion
    with pytest.raises(NotImplementedError):
        conversion.convert_sinogram_360_to_180(
            np.ones((10, 20), np.float32), 3.3)
