"""Half-acquisition sinogram conversion -- drop-in for
algotom.prep.conversion.convert_sinogram_360_to_180 (algotom 1.6.0,
algotom/prep/conversion.py:318-364) on the GPU.

Integer overlaps (centre of rotation at x.5, or an explicit integer
overlap) run in one HIP kernel pair; a sub-pixel overlap needs scipy's
spline shift (conversion.py:108-115) and custom weight matrices are host
arrays: both raise NotImplementedError rather than falling back to the CPU.
"""

import numpy as np

from algotom_amd import _host
from algotom_amd import _native as nat


def _overlap_side(ncol, cor):
    # conversion.py:346-354
    xcenter = (ncol - 1.0) * 0.5
    if isinstance(cor, tuple):
        return cor
    if cor <= xcenter:
        return 2 * cor + 1, 0
    return 2 * (ncol - cor) - 1, 1


def convert_sinogram_360_to_180(sino_360, cor, wei_mat1=None, wei_mat2=None,
                                norm=True, total_width=None):
    """Convert a 360-degree sinogram to a 180-degree sinogram.

    Parameters
    ----------
    sino_360 : array_like
        2D (nrow, ncol) or 3D (nrow, nslice, ncol) 360-degree sinogram(s);
        numpy or a torch tensor on the GPU.
    cor : float or tuple of float
        Center-of-rotation or (Overlap_area, overlap_side).
    wei_mat1, wei_mat2 : None
        Custom weighting matrices are not supported on the GPU path.
    norm : bool, optional
        Enable/disable normalization before stitching.
    total_width : int, optional
        Final width of the stitched image.

    Returns
    -------
    sino_stitch : array_like
        Converted sinogram(s), float32.
    cor : float
        Updated center-of-rotation referred to the converted sinogram.
    """
    import torch
    from algotom_amd.rec import reconstruction as rec
    if wei_mat1 is not None or wei_mat2 is not None:
        raise NotImplementedError("custom weight matrices are not supported")
    shape = tuple(sino_360.shape)
    two_d = len(shape) == 2
    nrow, ncol = shape[0], shape[-1]
    overlap, side = _overlap_side(ncol, cor)
    overlap_int = int(np.round(overlap + 1.0e-3))
    if overlap - overlap_int != 0.0:
        raise NotImplementedError(
            "sub-pixel overlap %r needs a spline shift; use a centre at x.5 "
            "or an integer overlap" % (overlap,))
    width0 = 2 * ncol - overlap_int
    width = width0 if (total_width is None or total_width < width0) \
        else int(total_width)
    dev = nat.device()
    as_numpy = not rec._is_tensor(sino_360)
    src = rec._to_device(sino_360, dev, allow_f64=False)
    src = src.reshape(nrow, -1, ncol)
    nslice = src.shape[1]
    out = torch.empty((nrow // 2 + 1, nslice, width), dtype=torch.float32,
                      device=dev)
    factors = torch.empty(2 * nslice, dtype=torch.float32, device=dev)
    lib = nat.load()
    nat.check(lib.ath_convert_sinogram_360_to_180(
        nat.ptr(src), nrow, nslice, ncol, overlap_int, int(side),
        int(bool(norm)), width, nat.ptr(factors), nat.ptr(out),
        nat.stream_ptr(dev)), "ath_convert_sinogram_360_to_180")
    new_cor = ncol - overlap_int / 2.0 - 0.5
    if two_d:
        out = out[:, 0, :]
    if as_numpy:
        out = _host.to_host(out)
    return out, new_cor


def extend_sinogram(sino_360, cor, apply_log=True):
    """Extend a 360-degree sinogram with an offset centre of rotation for
    reconstruction (conversion.py:395-454): (log), linear weights over the
    overlap, edge padding to 2*ncol - overlap columns, times 2.

    Parameters
    ----------
    sino_360 : array_like
        2D (nrow, ncol) or 3D (nrow, nslice, ncol) 360-degree sinogram(s);
        numpy or a torch tensor on the GPU.
    cor : float or tuple of float
        Center-of-rotation or (Overlap_area, overlap_side).
    apply_log : bool, optional
        Apply the logarithm function if True.

    Returns
    -------
    sino_pad : array_like
        Extended sinogram(s), float32.
    cor : float
        Updated center-of-rotation referred to the converted sinogram.
    """
    import torch
    from algotom_amd.rec import reconstruction as rec
    shape = tuple(sino_360.shape)
    two_d = len(shape) == 2
    nrow, ncol = shape[0], shape[-1]
    overlap, side = _overlap_side(ncol, cor)
    overlap_int = int(np.round(overlap + 1.0e-3))
    if overlap - overlap_int != 0.0:
        raise NotImplementedError(
            "sub-pixel overlap %r needs a spline shift; use a centre at x.5 "
            "or an integer overlap" % (overlap,))
    dev = nat.device()
    as_numpy = not rec._is_tensor(sino_360)
    src = rec._to_device(sino_360, dev, allow_f64=False).reshape(nrow, -1,
                                                                 ncol)
    nslice = src.shape[1]
    out = torch.empty((nrow, nslice, 2 * ncol - overlap_int),
                      dtype=torch.float32, device=dev)
    nat.check(nat.load().ath_extend_sinogram(
        nat.ptr(src), nrow, nslice, ncol, overlap_int, int(side),
        int(bool(apply_log)), nat.ptr(out), nat.stream_ptr(dev)),
        "ath_extend_sinogram")
    new_cor = ncol - overlap_int / 2.0 - 0.5
    if two_d:
        out = out[:, 0, :]
    if as_numpy:
        out = _host.to_host(out)
    return out, new_cor
