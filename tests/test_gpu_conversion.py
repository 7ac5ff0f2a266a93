"""GPU 360 -> 180 conversion against the oracle restatement.

Bar: bit-identical for integer overlaps without normalisation; with
normalisation the only difference is the float32 mean (numpy pairwise
float32 sum vs float64 sum here), and with a sub-pixel overlap the spline
prefilter (a symmetric FIR here, scipy's recursive filter in the oracle,
equal to ~1e-15 before the float32 rounding), so within 1e-6 relative.
Both sides, centre form and explicit (overlap, side), sub-pixel centres,
custom float32/float64 weights, total_width extension, 2D and stacks.
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


@pytest.mark.parametrize("cor", [20.3, 22.15, 77.6, (31.4, 0), (17.7, 1)])
@pytest.mark.parametrize("norm", [False, True])
def test_convert_subpixel_matches_oracle(cor, norm):
    # both images shifted by the sub-pixel remainder with the cubic spline
    # (scipy.ndimage.shift, mode "nearest"); the kernel's FIR prefilter and
    # scipy's recursive one agree to ~1e-15, so the float32 results differ
    # by at most an ulp where a value sits on a rounding boundary
    from oracle import conversion_np as oconv
    from algotom_amd.prep import conversion
    rng = np.random.default_rng(9)
    sino = rng.uniform(0.2, 1.0, (61, 100)).astype(np.float32)
    got, gc = conversion.convert_sinogram_360_to_180(sino, cor, norm=norm)
    ref, rc = oconv.convert_360_to_180(sino, cor, norm=norm)
    assert gc == rc and got.shape == ref.shape
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("cor %r norm %s: max rel %.3g, %d of %d differ"
          % (cor, norm, err, int((got != ref).sum()), got.size))
    assert err <= 1e-6


@pytest.mark.parametrize("wdtype", [np.float32, np.float64])
@pytest.mark.parametrize("cor", [20.5, 20.3])
def test_convert_custom_weights(wdtype, cor):
    from oracle import conversion_np as oconv
    from algotom_amd.prep import conversion
    rng = np.random.default_rng(10)
    sino = rng.uniform(0.2, 1.0, (41, 64)).astype(np.float32)
    n180 = 21
    w1 = rng.uniform(0.0, 1.0, (n180, 64)).astype(wdtype)
    w2 = (1.0 - w1).astype(wdtype)
    got, _ = conversion.convert_sinogram_360_to_180(sino, cor, wei_mat1=w1,
                                                     wei_mat2=w2, norm=False)
    ref, _ = oconv.convert_360_to_180(sino, cor, norm=False, wei_mat1=w1,
                                      wei_mat2=w2)
    if cor == 20.5:
        assert np.array_equal(got, ref)
    else:
        assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()
    # a stack shares the weights across slices
    st = rng.uniform(0.2, 1.0, (41, 3, 64)).astype(np.float32)
    got3, _ = conversion.convert_sinogram_360_to_180(st, cor, wei_mat1=w1,
                                                      wei_mat2=w2, norm=False)
    for j in range(3):
        r, _ = oconv.convert_360_to_180(st[:, j, :], cor, norm=False,
                                        wei_mat1=w1, wei_mat2=w2)
        assert np.abs(got3[:, j, :] - r).max() <= 1e-6 * np.abs(r).max()


def test_convert_subpixel_stack_total_width():
    from oracle import conversion_np as oconv
    from algotom_amd.prep import conversion
    rng = np.random.default_rng(11)
    stack = rng.uniform(0.2, 1.0, (40, 3, 50)).astype(np.float32)
    got, _ = conversion.convert_sinogram_360_to_180(stack, 9.37,
                                                     total_width=120)
    assert got.shape == (21, 3, 120)
    for j in range(3):
        ref, _ = oconv.convert_360_to_180(stack[:, j, :], 9.37,
                                          total_width=120)
        assert np.abs(got[:, j, :] - ref).max() <= 1e-6 * np.abs(ref).max()


def test_convert_full_width_subpixel():
    # a 4096-column scan at the sub-pixel centre 1234.37 (overlap 2469.74:
    # 2470 columns and a -0.26 px shift, stitched width 5722; the shift
    # kernel needs dynamic LDS above 64 KB)
    from oracle import conversion_np as oconv
    from algotom_amd.prep import conversion
    rng = np.random.default_rng(12)
    sino = rng.uniform(0.2, 1.0, (21, 4096)).astype(np.float32)
    got, gc = conversion.convert_sinogram_360_to_180(sino, 1234.37)
    ref, rc = oconv.convert_360_to_180(sino, 1234.37)
    assert gc == rc and got.shape == ref.shape == (11, 2 * 4096 - 2470)
    assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("cor", [20.5, 44.5, (12, 0), (9, 1), (64, 1),
                                 20.3, 44.65, (12.4, 0), (9.7, 1)])
@pytest.mark.parametrize("apply_log", [True, False])
def test_extend_sinogram_vs_oracle(cor, apply_log):
    from oracle import conversion_np
    from algotom_amd.prep import conversion
    rng = np.random.default_rng(3)
    s = rng.uniform(0.2, 1.0, (90, 64)).astype(np.float32)
    got, gc = conversion.extend_sinogram(s.copy(), cor, apply_log)
    ref, rc = conversion_np.extend(s, cor, apply_log)
    assert gc == rc and got.dtype == np.float32
    c = cor[0] if isinstance(cor, tuple) else cor
    exact = float(c) * 2 == int(float(c) * 2)  # no sub-pixel shift

    def same(a, b):
        if exact:
            return np.array_equal(a, b)
        return np.abs(a - b).max() <= 1e-6 * np.abs(b).max()

    assert same(got, ref)
    # a stack: every slice as its own 2D call
    st = rng.uniform(0.2, 1.0, (90, 3, 64)).astype(np.float32)
    got3, _ = conversion.extend_sinogram(st.copy(), cor, apply_log)
    for j in range(3):
        assert same(got3[:, j, :],
                    conversion_np.extend(st[:, j, :], cor, apply_log)[0])


@pytest.mark.parametrize("case", ["f32", "f64", "u16", "row", "nodark",
                                  "zeros"])
def test_flat_field_correction_vs_oracle(case):
    from oracle import conversion_np
    from algotom_amd.prep import correction
    rng = np.random.default_rng(5)
    flat = rng.uniform(1000, 2000, (20, 33)).astype(np.float32)
    dark = rng.uniform(50, 60, (20, 33)).astype(np.float32)
    proj = rng.uniform(60, 1500, (7, 20, 33)).astype(np.float32)
    kw = {}
    if case == "f64":
        flat, dark = flat.astype(np.float64), dark.astype(np.float64)
    if case == "u16":
        proj = proj.astype(np.uint16)
    if case == "row":
        proj, flat, dark = proj[:, 3, :], flat[3], dark[3]
    if case == "nodark":
        kw = dict(ratio=1.5, use_dark=False)
    if case == "zeros":
        flat[2, 5] = dark[2, 5]
        flat[7, 1] = dark[7, 1]
    got = correction.flat_field_correction(proj, flat.copy(), dark.copy(),
                                           **kw)
    ref = conversion_np.flat_field(proj, flat.copy(), dark.copy(), **kw)
    assert got.dtype == ref.dtype and got.shape == ref.shape
    assert np.array_equal(got, ref)
