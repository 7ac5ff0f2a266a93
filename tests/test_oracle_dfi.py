"""Pins the DFI oracle (oracle/dfi_np.py) -- CPU only.

* B-spline pieces against scipy.ndimage 1.15.3, the library the reference
  calls (reconstruction.py:769-779, :741, :797-799): shift order 3 in modes
  "nearest" and "wrap" (bit-identical float32 results) and map_coordinates
  order 5 mode "reflect" (float64 within 1e-13 on lines of >= 48 samples;
  scipy's start-up on shorter lines is covered separately);
* the mathematics: partition of unity, polynomial reproduction, the
  inverse-filter taps against the closed form of the quintic's poles and the
  sampled B-spline they invert;
* generate_mapping_coordinate's geometry at closed-form points;
* end to end, the reference test's thresholds and errors
  (tests/test_rec/test_reconstruction.py:204-237) and a scipy-composed
  restatement of the per-slice pipeline on non-trivial sizes.
"""

import warnings

import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import dfi_np as D


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("s", [0.0, 0.4, -3.7, 17.2, -25.5, 3.0])
def test_shift_nearest_matches_scipy(dtype, s):
    x = np.random.default_rng(0).random((13, 61)).astype(dtype)
    ref = ndi.shift(x, (0, s), order=3, mode="nearest")
    got = D.shift_nearest(x, s, axis=-1)
    assert got.dtype == ref.dtype
    tol = 0.0 if dtype == np.float32 else 1e-14
    assert np.abs(got.astype(np.float64) - ref).max() <= tol


@pytest.mark.parametrize("n", [5, 21, 73])
@pytest.mark.parametrize("s", [0.0, 1.0, -2.0, 0.5, 13.3, -40.7])
def test_shift_wrap_matches_scipy(n, s):
    x = np.random.default_rng(n).random((n, 7)).astype(np.float32)
    ref = ndi.shift(x, (s, 0), order=3, mode="wrap")
    got = D.shift_wrap(x, s, axis=0)
    assert np.array_equal(got, ref)


def test_map_coordinates_matches_scipy():
    rng = np.random.default_rng(1)
    a = rng.random((50, 61)) + 1j * rng.random((50, 61))
    rr = np.concatenate([rng.uniform(0, 49, 3000), [0.0, 49.0, 0.0, 49.0]])
    cc = np.concatenate([rng.uniform(0, 60, 3000), [0.0, 60.0, 60.0, 0.0]])
    ref = ndi.map_coordinates(a, [rr, cc], order=5, mode="reflect")
    got = D.map_coordinates_reflect5(a, rr, cc)
    assert np.abs(got - ref).max() <= 1e-13


def test_map_coordinates_short_axis_startup():
    """On a 37-row axis (the reference test's sinogram height) scipy's
    truncated start-up leaves a difference far below the test tolerances."""
    rng = np.random.default_rng(2)
    a = rng.random((37, 65))
    rr = rng.uniform(0, 36, 2000)
    cc = rng.uniform(0, 64, 2000)
    ref = ndi.map_coordinates(a, [rr, cc], order=5, mode="reflect")
    got = D.map_coordinates_reflect5(a, rr, cc)
    assert np.abs(got - ref).max() <= 1e-6


@pytest.mark.parametrize("order", [3, 5])
def test_bspline_partition_of_unity_and_support(order):
    x = np.linspace(-0.5, 0.5, 101)
    total = sum(D.bspline(x - k, order) for k in range(-4, 5))
    assert np.abs(total - 1.0).max() <= 1e-14
    assert D.bspline(np.array([(order + 1) / 2.0]), order)[0] == 0.0


@pytest.mark.parametrize("order", [3, 5])
def test_inverse_taps_invert_the_sampled_spline(order):
    h = D.inverse_taps(order)
    half = (len(h) - 1) // 2
    b = D.bspline(np.arange(-(order // 2), order // 2 + 1), order)
    conv = np.convolve(h, b)
    centre = half + order // 2
    delta = np.zeros_like(conv)
    delta[centre] = 1.0
    assert np.abs(conv - delta)[order:-order].max() <= 1e-15


def test_quintic_taps_match_pole_closed_form():
    """1/B(q) = C (z2 g_z2 - z1 g_z1) / (z1 z2 (a1 - a2)), g_z[k] =
    z^|k|/(1 - z^2), with a = z + 1/z the roots of a^2 + 26a + 64 = 0."""
    a1, a2 = -13.0 + np.sqrt(105.0), -13.0 - np.sqrt(105.0)
    z1 = 2.0 / (a1 - np.sqrt(a1 * a1 - 4.0))
    z2 = 2.0 / (a2 - np.sqrt(a2 * a2 - 4.0))
    k = np.arange(40)
    g = lambda z: z ** k / (1.0 - z * z)  # noqa: E731
    C = (1 - z1) ** 2 * (1 - z2) ** 2
    h = C * (z2 * g(z2) - z1 * g(z1)) / (z1 * z2 * (a1 - a2))
    ref = D.inverse_taps(5)[64:64 + 40]
    assert np.abs(h - ref).max() <= 1e-15


@pytest.mark.parametrize("order", [3, 5])
def test_spline_reproduces_polynomials(order):
    """Interior values of a spline fitted to samples of a polynomial of
    degree <= order are the polynomial itself."""
    n = 120
    t = np.arange(n, dtype=np.float64) / n
    poly = 0.3 - 1.1 * t + 0.7 * t ** 2 + (0.5 * t ** 3 if order >= 3 else 0)
    c = D.prefilter(poly, order, "half")
    pos = np.linspace(40.0, 80.0, 57)
    got = D.evaluate(c, pos, order, "half")
    tt = pos / n
    want = 0.3 - 1.1 * tt + 0.7 * tt ** 2 + 0.5 * tt ** 3
    assert np.abs(got - want).max() <= 1e-12


def test_mapping_coordinates_geometry():
    """generate_mapping_coordinate at closed-form points: X = c - x,
    Y = y - c.  Centre row (Y = 0): theta = pi (both sides; the right half
    has atan2 = pi, so r is negated), r = 2c - x from 2c down to 0.  Centre
    column (X = 0, off the centre): theta = pi/2 -> (nrow - 1)/2, r = 2c - y.
    Every value is float32 and inside the sinogram spectrum."""
    w, h = 33, 19
    r, th = D.mapping_coordinates(w, h, w, w)
    assert r.dtype == np.float32 and th.dtype == np.float32
    c = (w - 1) // 2
    x = np.arange(w)
    np.testing.assert_array_equal(r[c, :], (2 * c - x).astype(np.float32))
    assert np.all(th[c, :] == np.float32(h - 1))
    np.testing.assert_array_equal(r[:, c], (2 * c - x).astype(np.float32))
    off = np.arange(w) != c  # the centre itself: atan2(0, 0) = 0 -> nrow - 1
    assert np.all(th[off, c] == np.float32((h - 1) / 2.0))
    assert th[c, c] == np.float32(h - 1)
    # the diagonal through the centre: |r - c| = sqrt(2) |y - c|, 45 degrees
    d = np.arange(w) - c
    np.testing.assert_allclose(np.abs(r[x, x] - c), np.float32(
        np.sqrt(2.0) * np.abs(d)).clip(0, c), rtol=0, atol=1e-5)
    assert r.min() >= 0 and r.max() <= w - 1
    assert th.min() >= 0 and th.max() <= h - 1


def _reference_case():
    size = 64
    mat = np.zeros((size + 1, size + 1), dtype=np.float32)
    mat[20:40, 30:45] = np.float32(1.0)
    mat = ndi.gaussian_filter(mat, 2.0)
    angles = np.linspace(0.0, 360.0, 73, dtype=np.float32)
    sino = np.zeros((73, size + 1), dtype=np.float32)
    for i, ang in enumerate(angles):
        sino[i] = np.sum(ndi.rotate(mat, -ang, reshape=False), axis=0)
    return mat, sino, angles, size // 2


def test_reference_thresholds():
    """tests/test_rec/test_reconstruction.py:204-237."""
    mat, sino_360, angles, center = _reference_case()
    sino_180, angles2 = sino_360[:37], angles[:37]
    r1 = D.dfi(sino_180, center, ratio=0.0, apply_log=False)
    assert r1.dtype == np.float64 and np.abs(mat - r1).max() <= 0.1
    r2 = D.dfi(sino_360, center, angles=np.deg2rad(angles), apply_log=False,
               ratio=None)
    assert np.abs(mat - r2).max() <= 0.1
    stack = np.pad(sino_180[:, None, :], ((0, 0), (2, 2), (0, 0)),
                   mode="edge")
    r3 = D.dfi(stack, center, apply_log=False, ratio=0.0,
               angles=np.deg2rad(angles2))
    assert r3.shape == (65, 5, 65) and np.abs(mat - r3[:, 0, :]).max() <= 0.1
    r4 = D.dfi(stack, center, apply_log=False, ratio=0.0)
    assert np.abs(mat - r4[:, 0, :]).max() <= 0.1
    with pytest.raises(ValueError):
        D.dfi(sino_180, center, np.deg2rad(angles), apply_log=False)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        D.dfi(sino_180, center, apply_log=True)
    assert len(w) >= 1
    with pytest.raises(ValueError):
        D.dfi(sino_180, -1, apply_log=False)


@pytest.mark.parametrize("ncol,nrow,center", [(96, 121, 46.3), (81, 60, 41.0)])
def test_pipeline_against_scipy_composition(ncol, nrow, center):
    """handle_sinogram + single_slice against the same steps composed from
    scipy.ndimage (sizes where every spline line has >= 48 samples)."""
    from oracle import fbp_np
    rng = np.random.default_rng(ncol)
    sino = (rng.random((nrow, ncol)) + 0.5).astype(np.float32)
    got_s, _, pad = D.handle_sinogram(sino, None, center, 0.25, "edge")
    s = np.pad(sino, ((0, 0), (0, 1)), mode="edge") if ncol % 2 == 0 else sino
    ncol1 = s.shape[1]
    ref_s = ndi.shift(s, (0, (ncol1 - 1) / 2.0 - center), mode="nearest")
    ref_s = np.pad(ref_s, ((0, 0), (pad, pad)), mode="edge")
    assert np.array_equal(got_s, ref_s)
    w = ref_s.shape[1]
    mask = fbp_np.circle_mask(w, 1.0)
    r_mat, theta_mat = D.mapping_coordinates(w, nrow, w, w)
    window = np.tile(fbp_np.smoothing_window("hann", w), (nrow, 1))
    got = D.single_slice(ref_s, window, mask, r_mat, theta_mat)
    ref = D.scipy_single_slice(ref_s.astype(np.float64), window, mask, r_mat,
                               theta_mat)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    whole = D.scipy_dfi_slice(sino, center, apply_log=False)
    mine = D.dfi(sino, center, ratio=None, apply_log=False)
    assert np.abs(whole - mine).max() <= 1e-12 * np.abs(whole).max()
