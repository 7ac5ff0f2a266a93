"""Oracle for the 360 -> 180 degree conversion -- TEST INFRASTRUCTURE ONLY.

NumPy restatement of algotom/prep/conversion.py:45-73 (make_weight_matrix),
:76-144 (stitch_image) and :318-364 (convert_sinogram_360_to_180) for
integer overlaps (no scipy shift).  Pinned in tests/test_oracle_conversion.py
by closed forms: constant inputs, the weight ramp summing to one, the
identity of the non-overlapping parts, and the returned centre.
"""

import numpy as np


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


def stitch(mat1, mat2, overlap, side, norm=True, total_width=None):
    """stitch_image (conversion.py:76-144) for an integer overlap."""
    nrow, ncol1 = mat1.shape
    ncol2 = mat2.shape[1]
    w1, w2 = weight_matrices(mat1.shape, mat2.shape, overlap, side)
    width0 = ncol1 + ncol2 - overlap
    out = np.zeros((nrow, width0), dtype=np.float32)
    if side == 1:
        if norm:
            f1 = np.mean(mat1[:, -overlap:])
            f2 = np.mean(mat2[:, :overlap])
            mat2 = mat2 * f1 / f2
        out[:, 0:ncol1] = mat1 * w1
        out[:, ncol1 - overlap:width0] += mat2 * w2
    else:
        if norm:
            f2 = np.mean(mat2[:, -overlap:])
            f1 = np.mean(mat1[:, :overlap])
            mat2 = mat2 * f1 / f2
        out[:, 0:ncol2] = mat2 * w2
        out[:, ncol2 - overlap:width0] += mat1 * w1
    if total_width is not None and total_width > width0:
        out = np.pad(out, ((0, 0), (0, total_width - width0)), mode="edge")
    return out


def convert_360_to_180(sino, cor, norm=True, total_width=None):
    """convert_sinogram_360_to_180 (conversion.py:318-364)."""
    nrow, ncol = sino.shape
    xcenter = (ncol - 1.0) * 0.5
    n180 = nrow // 2 + 1
    top = sino[0:n180, :]
    bot = np.fliplr(sino[-n180:, :])
    if isinstance(cor, tuple):
        overlap, side = cor
    elif cor <= xcenter:
        overlap, side = 2 * cor + 1, 0
    else:
        overlap, side = 2 * (ncol - cor) - 1, 1
    ov = int(np.round(overlap + 1.0e-3))
    assert ov == overlap, "integer overlaps only"
    out = stitch(top, bot, ov, side, norm, total_width)
    return out, ncol - ov / 2.0 - 0.5


def extend(sino_360, cor, apply_log=True):
    """extend_sinogram (conversion.py:395-454) for an integer overlap,
    2D float32 input; log in float64 rounded (the oracle's convention)."""
    s = np.asarray(sino_360, dtype=np.float32)
    if apply_log:
        s = (-np.log(s.astype(np.float64))).astype(np.float32)
    else:
        s = s.copy()
    nrow, ncol = s.shape
    xc = (ncol - 1.0) * 0.5
    if isinstance(cor, tuple):
        overlap, side = cor
    elif cor <= xc:
        overlap, side = 2 * cor + 1, 0
    else:
        overlap, side = 2 * (ncol - cor) - 1, 1
    ov = int(np.round(overlap + 1.0e-3))
    pad = ncol - ov
    if side == 1:
        s[:, -ov:] = s[:, -ov:] * np.linspace(1.0, 0.0, ov)[None, :]
        out = np.pad(s, ((0, 0), (0, pad)), mode="edge")
    else:
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
