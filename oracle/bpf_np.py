"""Oracle for BPF (back-projection filtering) -- TEST INFRASTRUCTURE ONLY.

Restates algotom/rec/reconstruction.py:494-521 (make_circular_ramp_window,
with algotom/util/utility.py:925-963 transform_1d_window_to_2d), :524-560
(apply_circular_ramp_filter) and :563-684 (bpf_reconstruction).  The
circular window's radial interpolation is written here with np.interp (the
reference goes through scipy's map_coordinates); the filter is the
reference's own spectral product in complex128 (numpy 1.x semantics); the
back-projection is the oracle's C port with edge padding.

Pinned in tests/test_oracle_bpf.py: the window against the map_coordinates
construction, the filter against a brute-force 2D circular convolution on
tiny images (every pad mode) and against the tiled-kernel linear
convolution, and end to end against the reference test's thresholds
(tests/test_rec/test_reconstruction.py:146-202).
"""

import numpy as np

from oracle import fbp_np, port


def ramp_spectrum_abs(width, filter_name=None):
    """|fftshift(fft(Ram-Lak))| * smoothing window (reconstruction.py
    :510-518)."""
    h = fbp_np.ramlak_kernel(width)
    return np.abs(np.fft.fftshift(np.fft.fft(h))) * \
        fbp_np.smoothing_window(filter_name, width)


def circular_window(width, filter_name=None):
    """make_circular_ramp_window: the 1d spectrum placed at signed radius
    r (float32, clipped to the half width) from the centre, linear
    interpolation along r (the rows of the reference's tiled matrix are
    identical, so its bilinear sampling reduces to this)."""
    win = ramp_spectrum_abs(width, filter_name)
    n = width + 1 if width % 2 == 0 else width
    c = n // 2
    x = (np.arange(n)[::-1] - c).astype(np.float64)[None, :]
    y = (np.arange(n) - c).astype(np.float64)[:, None]
    r = np.float32(np.minimum(np.sqrt(x * x + y * y), c))
    neg = np.arctan2(np.broadcast_to(y, r.shape), np.broadcast_to(x, r.shape)) < 0
    r = np.where(neg, -r, r).astype(np.float32)
    pos = np.float32(np.clip(r + np.float32(c), 0, n - 1)).astype(np.float64)
    out = np.interp(pos.ravel(), np.arange(width, dtype=np.float64),
                    win).reshape(n, n)
    return out[:width, :width]


def circular_filter(img, ramp_win, pad, pad_mode="edge"):
    """apply_circular_ramp_filter (:524-560): pad, fft2, shifted product
    with the window, ifft2, real part, crop -- complex128 throughout."""
    ncol = img.shape[0]
    p = np.pad(np.asarray(img, dtype=np.float64), pad, mode=pad_mode)
    spec = np.fft.fftshift(np.fft.fft2(p)) * ramp_win
    out = np.real(np.fft.ifft2(np.fft.ifftshift(spec)))
    return out[pad:ncol + pad, pad:ncol + pad]


def circular_kernel(ramp_win):
    """g2 with circular_filter(x) = (pad(x) circ-conv g2) cropped."""
    return np.real(np.fft.ifft2(np.fft.ifftshift(ramp_win)))


def bpf(sinogram, center, angles=None, ratio=1.0, filter_name="hann",
        pad=None, pad_mode="edge", apply_log=True, gpu=True):
    """bpf_reconstruction (:563-684): (ncol, ncol) float64 for one
    sinogram, (ncol, nslice, ncol) float32 for a stack."""
    sino = np.array(sinogram, dtype=np.float32, copy=True)
    input_3d = sino.ndim == 3
    if not input_3d:
        sino = sino[:, None, :]
    nrow, nslice, ncol = sino.shape
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, nrow))
    if apply_log:
        sino[sino <= 0.0] = 1.0
        sino = (-np.log(sino.astype(np.float64))).astype(np.float32)
    if pad is None:
        pad = min(int(0.15 * ncol), 150)
    win = circular_window(ncol + 2 * pad, filter_name)
    mask = None
    if ratio is not None:
        if ratio == 0.0:
            ratio = min(center, ncol - center) / (0.5 * ncol)
        mask = fbp_np.circle_mask(ncol, ratio)
    acc = fbp_np.ACC_F64 if (gpu and nslice == 1) else fbp_np.ACC_F32SEQ
    bp = port.backproject(sino, angles, center, edge_pad=True, acc=acc)
    if nslice == 1:
        out = circular_filter(bp[:, 0, :], win, pad, pad_mode)
        if mask is not None:
            out = out * mask
        out = out * np.pi / (nrow - 1)
        return out[:, None, :] if input_3d else out
    out = np.empty((ncol, nslice, ncol), dtype=np.float32)
    for j in range(nslice):
        out[:, j, :] = np.float32(circular_filter(bp[:, j, :], win, pad,
                                                  pad_mode))
        if mask is not None:
            out[:, j, :] = out[:, j, :] * mask
    return out * np.float32(np.pi) / np.float32(nrow - 1)
