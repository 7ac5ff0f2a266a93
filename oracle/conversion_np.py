"""Oracle for the 360 -> 180 degree conversion -- TEST INFRASTRUCTURE ONLY.

NumPy restatement of algotom/prep/conversion.py:45-73 (make_weight_matrix),
:76-144 (stitch_image), :318-364 (convert_sinogram_360_to_180) and
:395-454 (extend_sinogram), including the sub-pixel part of the overlap,
which the reference shifts with scipy.ndimage.shift (cubic spline, mode
"nearest"; scipy is the reference's own dependency and is called here the
same way), and caller weight matrices.  Pinned in
tests/test_oracle_conversion.py by closed forms: constant inputs, the
weight ramp summing to one, the identity of the non-overlapping parts, the
returned centre, a linear ramp shifted exactly by the spline away from the
edges, and the shift's continuity as the overlap's fraction goes to zero.
"""

import numpy as np
from scipy.ndimage import shift as ndi_shift


def weight_matrices(shape1, shape2, overlap, side):
    """make_weight_matrix (conversion.py:45-73)."""
    w1 = np.ones(shape1, dtype=np.float32)
    w2 = np.ones(shape2, dtype=np.float32)
    down = np.linspace(1.0, 0.0, overlap)
    up = 1.0 - down
    if side == 1:
        w1[:, -overlap:] = np.float32(down)
        w2[:, :overlap] = np.float32(up)
    else:
        w2[:, -overlap:] = np.float32(down)
        w1[:, :overlap] = np.float32(up)
    return w1, w2


def stitch(mat1, mat2, overlap, side, norm=True, total_width=None,
           wei_mat1=None, wei_mat2=None):
    """stitch_image (conversion.py:76-144): the sub-pixel remainder of
    overlap shifts both images (:107-115), then the linear-ramp (or the
    caller's) weights over round(overlap + 1e-3) columns."""
    nrow, ncol1 = mat1.shape
    ncol2 = mat2.shape[1]
    ov = int(np.round(overlap + 1.0e-3))
    sub = overlap - ov
    if sub != 0.0:
        if side == 1:
            mat1 = ndi_shift(mat1, (0, sub), mode="nearest")
            mat2 = ndi_shift(mat2, (0, -sub), mode="nearest")
        else:
            mat1 = ndi_shift(mat1, (0, -sub), mode="nearest")
            mat2 = ndi_shift(mat2, (0, sub), mode="nearest")
    if wei_mat1 is None or wei_mat2 is None:
        wei_mat1, wei_mat2 = weight_matrices(mat1.shape, mat2.shape, ov, side)
    width0 = ncol1 + ncol2 - ov
    out = np.zeros((nrow, width0), dtype=np.float32)
    if side == 1:
        if norm:
            f1 = np.mean(mat1[:, -ov:])
            f2 = np.mean(mat2[:, :ov])
            mat2 = mat2 * f1 / f2
        out[:, 0:ncol1] = mat1 * wei_mat1
        out[:, ncol1 - ov:width0] += mat2 * wei_mat2
    else:
        if norm:
            f2 = np.mean(mat2[:, -ov:])
            f1 = np.mean(mat1[:, :ov])
            mat2 = mat2 * f1 / f2
        out[:, 0:ncol2] = mat2 * wei_mat2
        out[:, ncol2 - ov:width0] += mat1 * wei_mat1
    if total_width is not None and total_width > width0:
        out = np.pad(out, ((0, 0), (0, total_width - width0)), mode="edge")
    return out


def _overlap_side(ncol, cor):
    xcenter = (ncol - 1.0) * 0.5
    if isinstance(cor, tuple):
        return cor
    if cor <= xcenter:
        return 2 * cor + 1, 0
    return 2 * (ncol - cor) - 1, 1


def convert_360_to_180(sino, cor, norm=True, total_width=None,
                       wei_mat1=None, wei_mat2=None):
    """convert_sinogram_360_to_180 (conversion.py:318-364)."""
    nrow, ncol = sino.shape
    n180 = nrow // 2 + 1
    top = sino[0:n180, :]
    bot = np.fliplr(sino[-n180:, :])
    overlap, side = _overlap_side(ncol, cor)
    out = stitch(top, bot, overlap, side, norm, total_width, wei_mat1,
                 wei_mat2)
    ov = int(np.round(overlap + 1.0e-3))
    return out, ncol - ov / 2.0 - 0.5


def extend(sino_360, cor, apply_log=True):
    """extend_sinogram (conversion.py:395-454), 2D float32 input; the log in
    float64 rounded (the oracle's convention for the float32 log)."""
    s = np.asarray(sino_360, dtype=np.float32)
    if apply_log:
        s = (-np.log(s.astype(np.float64))).astype(np.float32)
    else:
        s = s.copy()
    nrow, ncol = s.shape
    overlap, side = _overlap_side(ncol, cor)
    ov = int(np.round(overlap + 1.0e-3))
    sub = overlap - ov
    pad = ncol - ov
    if side == 1:
        if sub != 0.0:
            s = ndi_shift(s, (0, sub), mode="nearest")
        s[:, -ov:] = s[:, -ov:] * np.linspace(1.0, 0.0, ov)[None, :]
        out = np.pad(s, ((0, 0), (0, pad)), mode="edge")
    else:
        if sub != 0.0:
            s = ndi_shift(s, (0, -sub), mode="nearest")
        s[:, :ov] = s[:, :ov] * np.linspace(0.0, 1.0, ov)[None, :]
        out = np.pad(s, ((0, 0), (pad, 0)), mode="edge")
    return 2 * out, ncol - ov / 2.0 - 0.5


def flat_field(proj, flat, dark, ratio=1.0, use_dark=True):
    """flat_field_correction (correction.py:45-98) without options."""
    flat = ratio * np.asarray(flat)
    if use_dark:
        den = flat - dark
        if 0.0 in den:
            m = np.mean(den)
            den[den == 0.0] = m if m != 0.0 else 1
        return (np.float32(proj) - dark) / den
    if 0.0 in flat:
        m = np.mean(flat)
        flat[flat == 0.0] = m if m != 0.0 else 1
    return np.float32(proj) / flat
