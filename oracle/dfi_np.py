"""Oracle for DFI (direct Fourier inversion) -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module, to check results against.

Restates algotom/rec/reconstruction.py:
* ``mapping_coordinates``  :687-728 (generate_mapping_coordinate);
* ``handle_angles``        :731-746 (__dfi_handle_angles);
* ``handle_sinogram``      :749-781 (__dfi_handle_sinogram);
* ``single_slice``         :784-806 (__dfi_single_slice, scipy >= 1.6 branch);
* ``dfi``                  :809-907 (dfi_reconstruction);
with algotom/util/utility.py mapping (map_coordinates over a coordinate
pair) and make_circle_mask.

The reference reaches its spline interpolation through scipy.ndimage
(scipy 1.15.3 in this image): ``shift(order=3, mode="nearest")`` for the
centre shift, ``shift(order=3, mode="wrap")`` for the angle offset and
``map_coordinates(order=5, mode="reflect")`` for the polar-to-Cartesian
gridding.  Those routines implement the published B-spline method (Unser,
Aldroubi & Eden, IEEE TSP 41 (1993) 821-848; Thevenaz, Blu & Unser, IEEE TMI
19 (2000) 739-758): a prefilter (the inverse of the sampled B-spline,
applied along every axis to a symmetric extension of the samples), then
evaluation as sum_k c[k] beta_n(x - k) over the n+1 nearest knots.  Here the
prefilter is the exact inverse filter's impulse response (taps from the
discrete-time Fourier transform of 1 / B(w)) convolved with the extended
samples.  The boundary rules, pinned against scipy.ndimage in
tests/test_oracle_dfi.py:

* mode "nearest": the line is first edge-padded by 12 samples, prefiltered
  with the half-sample symmetric extension (d c b a | a b c d | d c b a),
  and knots past the padded line are clamped to its ends;
* mode "reflect": half-sample symmetric extension for both prefilter and
  knot folding;
* mode "wrap": the prefilter uses the whole-sample symmetric extension
  (d c b | a b c d | c b a), positions outside [0, n - 1] wrap with period
  n - 1, knots fold with the whole-sample rule.

scipy starts its reflect-mode recursion on lines shorter than about 40
samples from a truncated sum, so there it differs from the exact symmetric
extension by up to ~1e-7 (relative); the oracle keeps the exact extension.
An unshifted axis of ``ndimage.shift`` is taken as the identity (evaluating a
spline at its knots returns the samples up to rounding).
"""

import warnings

import numpy as np

from oracle import fbp_np

NEAREST_PAD = 12  # scipy's pre-padding for mode "nearest"


# --------------------------------------------------------------------------
# B-splines
# --------------------------------------------------------------------------
def bspline(t, order):
    """Centred B-spline of degree 3 or 5 from its truncated-power form."""
    t = np.abs(np.asarray(t, dtype=np.float64))
    p = lambda a, n: np.where(a > 0.0, a, 0.0) ** n  # noqa: E731
    if order == 3:
        return (p(2.0 - t, 3) - 4.0 * p(1.0 - t, 3)) / 6.0
    if order == 5:
        return (p(3.0 - t, 5) - 6.0 * p(2.0 - t, 5)
                + 15.0 * p(1.0 - t, 5)) / 120.0
    raise ValueError(order)


def inverse_taps(order, half=64):
    """h[k], k = -half..half: impulse response of 1 / B(w), B(w) = sum_k
    beta_n(k) e^{-iwk}, from its DTFT sampled on 2^14 points (the response
    decays as |z1|^k, |z1| < 0.44, so aliasing is far below rounding)."""
    nfft = 1 << 14
    w = 2.0 * np.pi * np.arange(nfft) / nfft
    m = order // 2
    B = bspline(0.0, order) + sum(
        2.0 * bspline(float(k), order) * np.cos(k * w) for k in range(1, m + 1))
    h = np.real(np.fft.ifft(1.0 / B))
    return np.concatenate([h[-half:], h[:half + 1]])


def _fold(idx, n, rule):
    idx = np.asarray(idx)
    if rule == "clamp":
        return np.clip(idx, 0, n - 1)
    if rule == "half":
        j = np.mod(idx, 2 * n)
        return np.where(j < n, j, 2 * n - 1 - j)
    if rule == "whole":
        if n == 1:
            return np.zeros_like(idx)
        j = np.mod(idx, 2 * n - 2)
        return np.where(j < n, j, 2 * n - 2 - j)
    raise ValueError(rule)


def prefilter(x, order, rule, axis=-1):
    """Spline coefficients along ``axis`` for the symmetric extension
    ``rule`` ("half": scipy reflect/nearest, "whole": scipy mirror/wrap)."""
    x = np.moveaxis(np.asarray(x, dtype=np.float64), axis, -1)
    n = x.shape[-1]
    h = inverse_taps(order)
    half = (len(h) - 1) // 2
    src = _fold(np.arange(-half, n + half), n, rule)
    ext = x[..., src]
    out = np.zeros_like(x)
    for k in range(-half, half + 1):
        out += h[k + half] * ext[..., half + k:half + k + n]
    return np.moveaxis(out, -1, axis)


def evaluate(c, pos, order, rule, axis=-1):
    """sum_k c[k] beta_n(pos - k) along ``axis`` for every pos (1D array),
    knots outside [0, n) mapped by ``rule``."""
    c = np.moveaxis(np.asarray(c, dtype=np.float64), axis, -1)
    n = c.shape[-1]
    pos = np.asarray(pos, dtype=np.float64)
    start = np.floor(pos).astype(np.int64) - order // 2
    out = np.zeros(c.shape[:-1] + pos.shape)
    for j in range(order + 1):
        idx = start + j
        out += bspline(pos - idx, order) * c[..., _fold(idx, n, rule)]
    return np.moveaxis(out, -1, axis)


def shift_nearest(x, s, axis=-1):
    """scipy.ndimage.shift(x, s along axis, order=3, mode="nearest"):
    out[i] = f(i - s)."""
    x = np.moveaxis(np.asarray(x), axis, -1)
    n = x.shape[-1]
    xp = x[..., np.clip(np.arange(-NEAREST_PAD, n + NEAREST_PAD), 0, n - 1)]
    c = prefilter(xp, 3, "half")
    out = evaluate(c, np.arange(n) - s + NEAREST_PAD, 3, "clamp")
    return np.moveaxis(out.astype(x.dtype), -1, axis)


def shift_wrap(x, s, axis=0):
    """scipy.ndimage.shift(x, s along axis, order=3, mode="wrap")."""
    x = np.moveaxis(np.asarray(x), axis, -1)
    n = x.shape[-1]
    c = prefilter(x, 3, "whole")
    pos = np.arange(n) - s
    if n > 1:  # only positions outside [0, n - 1] wrap (period n - 1)
        pos = np.where((pos < 0) | (pos > n - 1), np.mod(pos, n - 1), pos)
    out = evaluate(c, pos, 3, "whole")
    return np.moveaxis(out.astype(x.dtype), -1, axis)


def map_coordinates_reflect5(a, rows, cols):
    """scipy.ndimage.map_coordinates(a, [rows, cols], order=5,
    mode="reflect") for a real or complex 2D array (the prefilter is real,
    so real and imaginary parts go through it together)."""
    c = prefilter(a.real, 5, "half", axis=0)
    c = prefilter(c, 5, "half", axis=1)
    if np.iscomplexobj(a):
        ci = prefilter(prefilter(a.imag, 5, "half", axis=0), 5, "half",
                       axis=1)
        c = c + 1j * ci
    n0, n1 = c.shape
    rows = np.asarray(rows, dtype=np.float64)
    cols = np.asarray(cols, dtype=np.float64)
    s0 = np.floor(rows).astype(np.int64) - 2
    s1 = np.floor(cols).astype(np.int64) - 2
    wy = [bspline(rows - (s0 + i), 5) for i in range(6)]
    wx = [bspline(cols - (s1 + j), 5) for j in range(6)]
    cx = [_fold(s1 + j, n1, "half") for j in range(6)]
    flat = c.ravel()
    out = np.zeros(rows.shape, dtype=c.dtype)
    for i in range(6):
        base = _fold(s0 + i, n0, "half") * n1
        row = np.zeros(rows.shape, dtype=c.dtype)
        for j in range(6):
            row += wx[j] * flat[base + cx[j]]
        out += wy[i] * row
    return out


# --------------------------------------------------------------------------
# DFI (reconstruction.py:687-907)
# --------------------------------------------------------------------------
def mapping_coordinates(width_sino, height_sino, width_rec, height_rec):
    """generate_mapping_coordinate (reconstruction.py:687-728): float32
    (r, theta) sinogram-spectrum coordinates of each Cartesian frequency."""
    xcenter = (width_rec - 1.0) * 0.5
    ycenter = (height_rec - 1.0) * 0.5
    r_max = np.floor(min(xcenter, ycenter))
    x = (np.arange(width_rec)[::-1] - xcenter)[None, :]
    y = (np.arange(height_rec) - ycenter)[:, None]
    x, y = np.broadcast_arrays(x, y)
    r = np.float32(np.clip(np.sqrt(x * x + y * y), 0, r_max))
    theta = np.pi + np.arctan2(y, x)
    r[theta > np.pi] *= -1
    r = np.float32(np.clip(r + r_max, 0, width_sino - 1))
    theta[theta > np.pi] -= np.pi
    theta = np.float32(theta * (height_sino - 1.0) / np.pi)
    return r, theta


def angle_shift(angles):
    """(fold_360, rows, row shift, flip) of __dfi_handle_angles
    (reconstruction.py:731-746), the scalars in numpy's own arithmetic."""
    angles = np.asarray(angles)
    nrow = len(angles)
    t_ang = np.sum(np.abs(np.diff(angles * 180.0 / np.pi)))
    fold = bool(abs(t_ang - 360) < 10)
    rows = nrow // 2 + 1 if fold else nrow
    step = np.mean(np.abs(np.diff(angles)))
    b_ang = angles[0] - (angles[0] // (2 * np.pi)) * (2 * np.pi)
    return fold, rows, float(b_ang / step), bool(angles[-1] < angles[0])


def handle_angles(sino, angles):
    """__dfi_handle_angles on one (nrow, ncol) sinogram."""
    if len(angles) != sino.shape[0]:
        raise ValueError("!!! Number of angles is not the same as the row "
                         "number of the sinogram !!!")
    fold, rows, shift, flip = angle_shift(angles)
    if fold:
        sino = (sino[:rows] + np.fliplr(sino[-rows:])) / 2
    sino_360 = np.vstack((sino[:rows - 1], np.fliplr(sino)))
    sino = shift_wrap(sino_360, shift, axis=0)[:rows]
    if flip:
        sino = np.flipud(np.fliplr(sino))
    return sino


def handle_sinogram(sinogram, angles, center, pad_rate, pad_mode):
    """__dfi_handle_sinogram: odd width, centre shift, angles, padding.
    Returns (sinogram, num_sino, pad) with 3D stacks as (nrow, S, W)."""
    sino = np.squeeze(sinogram)
    if sino.ndim == 3:
        num_sino = sino.shape[1]
        if sino.shape[-1] % 2 == 0:
            sino = np.pad(sino, ((0, 0), (0, 0), (0, 1)), mode="edge")
    else:
        num_sino = 1
        if sino.shape[-1] % 2 == 0:
            sino = np.pad(sino, ((0, 0), (0, 1)), mode="edge")
    ncol1 = sino.shape[-1]
    xshift = (ncol1 - 1) / 2.0 - center
    pad = int(pad_rate * ncol1)
    sino = shift_nearest(sino, xshift, axis=-1)
    if num_sino > 1:
        if angles is not None:
            sino = np.moveaxis(np.asarray(
                [handle_angles(sino[:, i, :], angles)
                 for i in range(num_sino)]), 0, 1)
        sino = fbp_np.pad_rows(sino, pad, pad_mode)
    else:
        if angles is not None:
            sino = handle_angles(sino, angles)
        sino = fbp_np.pad_rows(sino, pad, pad_mode)
    return np.ascontiguousarray(sino), num_sino, pad


def single_slice(sino, window, mask, r_mat, theta_mat):
    """__dfi_single_slice: row spectra, order-5 gridding, inverse 2D FFT."""
    spec = np.fft.fftshift(np.fft.fft(np.fft.ifftshift(
        np.asarray(sino, dtype=np.float64), axes=1)), axes=1)
    if window is not None:
        spec = spec * window
    reg = map_coordinates_reflect5(spec, theta_mat, r_mat) * mask
    return np.real(np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(reg))))


def dfi(sinogram, center, angles=None, ratio=1.0, filter_name="hann",
        pad_rate=0.25, pad_mode="edge", apply_log=True):
    """dfi_reconstruction (reconstruction.py:809-907), float64 result."""
    sinogram = np.asarray(sinogram)
    input_3d = sinogram.ndim == 3
    nrow, ncol = sinogram.shape[0], sinogram.shape[-1]
    if center < 0 or center >= ncol:
        raise ValueError("Center is out of the range [0, {}]".format(ncol - 1))
    sino, num_sino, pad = handle_sinogram(sinogram, angles, center,
                                          pad_rate, pad_mode)
    if apply_log:
        if np.any(sino <= 0.0):
            warnings.warn("!!!Applying logarithm is enabled but there are "
                          "values <= 0.0 in the data!!!")
            sino = sino.copy()
            sino[sino <= 0.0] = np.float32(1.0)
        # numpy's float32 log modelled as correctly rounded
        sino = (-np.log(sino.astype(np.float64))).astype(sino.dtype)
    nrow, ncol2 = sino.shape[0], sino.shape[-1]
    mask = fbp_np.circle_mask(ncol2, 1.0)
    r_mat, theta_mat = mapping_coordinates(ncol2, nrow, ncol2, ncol2)
    window = None
    if filter_name is not None:
        window = np.tile(fbp_np.smoothing_window(filter_name, ncol2),
                         (nrow, 1))
    if num_sino > 1:
        recon = np.moveaxis(np.asarray(
            [single_slice(sino[:, i, :], window, mask, r_mat, theta_mat)
             for i in range(num_sino)]), 0, 1)
        recon = recon[pad:ncol + pad, :, pad:ncol + pad]
    else:
        recon = single_slice(sino, window, mask, r_mat, theta_mat)
        recon = recon[pad:ncol + pad, pad:ncol + pad]
    if ratio is not None:
        if ratio == 0.0:
            ratio = min(center, ncol - center) / (0.5 * ncol)
        m = fbp_np.circle_mask(ncol, ratio)
        recon = recon * (m[:, None, :] if num_sino > 1 else m)
    recon = np.ascontiguousarray(recon)
    if input_3d and num_sino == 1:
        recon = recon[:, None, :]
    return recon


# --------------------------------------------------------------------------
# The same slice through scipy.ndimage itself (the library calls the
# reference makes): the CPU baseline of the DFI bench and a cross-check of
# the restatement above.
# --------------------------------------------------------------------------
def scipy_single_slice(sino, window, mask, r_mat, theta_mat):
    """__dfi_single_slice (reconstruction.py:784-806) with
    scipy.ndimage.map_coordinates."""
    import scipy.ndimage as ndi
    spec = np.fft.fftshift(np.fft.fft(np.fft.ifftshift(
        np.asarray(sino, dtype=np.float64), axes=1)), axes=1)
    if window is not None:
        spec = spec * window
    coords = np.vstack((theta_mat.ravel(), r_mat.ravel()))
    reg = ndi.map_coordinates(spec, coords, order=5,
                              mode="reflect").reshape(r_mat.shape) * mask
    return np.real(np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(reg))))


def scipy_dfi_slice(sino, center, filter_name="hann", pad_rate=0.25,
                    pad_mode="edge", apply_log=True):
    """One 2D sinogram through reconstruction.py:849-895 (angles=None,
    ratio=None) with scipy.ndimage.shift for the centre shift."""
    import scipy.ndimage as ndi
    sino = np.asarray(sino)
    ncol = sino.shape[1]
    if ncol % 2 == 0:
        sino = np.pad(sino, ((0, 0), (0, 1)), mode="edge")
    ncol1 = sino.shape[1]
    pad = int(pad_rate * ncol1)
    sino = ndi.shift(sino, (0, (ncol1 - 1) / 2.0 - center), mode="nearest")
    sino = np.pad(sino, ((0, 0), (pad, pad)), mode=pad_mode)
    if apply_log:
        sino[sino <= 0.0] = np.float32(1.0)
        sino = -np.log(sino)
    nrow, w = sino.shape
    r_mat, theta_mat = mapping_coordinates(w, nrow, w, w)
    window = None
    if filter_name is not None:
        window = np.tile(fbp_np.smoothing_window(filter_name, w), (nrow, 1))
    rec = scipy_single_slice(sino, window, fbp_np.circle_mask(w, 1.0), r_mat,
                             theta_mat)
    return rec[pad:ncol + pad, pad:ncol + pad]
