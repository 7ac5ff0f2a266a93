"""Half-acquisition sinogram conversion -- drop-in for
algotom.prep.conversion.convert_sinogram_360_to_180 and extend_sinogram
(algotom 1.6.0, algotom/prep/conversion.py:318-364, :395-454) on the GPU.

Every step runs in HIP kernels: the sub-pixel part of the overlap shifts
the two images with scipy.ndimage.shift's cubic spline (mode "nearest",
conversion.py:109-117 and :435-453), the overlap means normalise image 2,
and the stitch applies make_weight_matrix's linear ramp or the caller's
weight matrices.  Nothing falls back to the CPU.
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


def _weights_device(wei, shape, dev):
    """A caller's weight matrix on the device (float32 stays float32, any
    other type becomes float64, as numpy's product with a float32 image
    would promote it)."""
    import torch
    from algotom_amd.rec import reconstruction as rec
    if rec._is_tensor(wei):
        t = wei.detach()
        t = t if t.dtype in (torch.float32, torch.float64) else t.double()
    else:
        a = np.asarray(wei)
        a = a if a.dtype in (np.float32, np.float64) else a.astype(np.float64)
        t = torch.from_numpy(np.ascontiguousarray(a))
    if tuple(t.shape) != tuple(shape):
        raise ValueError("weight matrices must have the shape of the images "
                         "being stitched %s, got %s" % (tuple(shape),
                                                        tuple(t.shape)))
    return t.to(dev).contiguous()


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
    wei_mat1 : array_like, optional
        Weighting matrix used for the 1st haft of the sinogram, shape
        (nrow // 2 + 1, ncol), shared by every slice of a stack.
    wei_mat2 : array_like, optional
        Weighting matrix used for the 2nd haft of the sinogram.
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
    shape = tuple(sino_360.shape)
    two_d = len(shape) == 2
    nrow, ncol = shape[0], shape[-1]
    overlap, side = _overlap_side(ncol, cor)
    overlap_int = int(np.round(overlap + 1.0e-3))
    width0 = 2 * ncol - overlap_int
    width = width0 if (total_width is None or total_width < width0) \
        else int(total_width)
    dev = nat.device()
    as_numpy = not rec._is_tensor(sino_360)
    src = rec._to_device(sino_360, dev, allow_f64=False)
    src = src.reshape(nrow, -1, ncol)
    nslice = src.shape[1]
    n180 = nrow // 2 + 1
    w1 = w2 = None
    wdt = nat.ATH_F32
    if wei_mat1 is not None and wei_mat2 is not None:
        w1 = _weights_device(wei_mat1, (n180, ncol), dev)
        w2 = _weights_device(wei_mat2, (n180, ncol), dev)
        if w1.dtype != w2.dtype:
            w1, w2 = w1.double(), w2.double()
        wdt = nat.dtype_code(w1.dtype)
    out = torch.empty((n180, nslice, width), dtype=torch.float32, device=dev)
    lib = nat.load()
    nwork = lib.ath_conversion_workspace_bytes(0, nrow, nslice, ncol,
                                               float(overlap))
    work = torch.empty(max(nwork, 1), dtype=torch.uint8, device=dev)
    nat.check(lib.ath_convert_sinogram_360_to_180(
        nat.ptr(src), nrow, nslice, ncol, float(overlap), int(side),
        int(bool(norm)), width, nat.ptr(w1) if w1 is not None else None,
        nat.ptr(w2) if w2 is not None else None, wdt, nat.ptr(work), nwork,
        nat.ptr(out), nat.stream_ptr(dev)), "ath_convert_sinogram_360_to_180")
    new_cor = ncol - overlap_int / 2.0 - 0.5
    if two_d:
        out = out[:, 0, :]
    if as_numpy:
        out = _host.to_host(out)
    return out, new_cor


def extend_sinogram(sino_360, cor, apply_log=True):
    """Extend a 360-degree sinogram with an offset centre of rotation for
    reconstruction (conversion.py:395-454): (log), the sub-pixel shift,
    linear weights over the overlap, edge padding to 2*ncol - overlap
    columns, times 2.

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
    dev = nat.device()
    as_numpy = not rec._is_tensor(sino_360)
    src = rec._to_device(sino_360, dev, allow_f64=False).reshape(nrow, -1,
                                                                 ncol)
    nslice = src.shape[1]
    lib = nat.load()
    out = torch.empty((nrow, nslice, 2 * ncol - overlap_int),
                      dtype=torch.float32, device=dev)
    nwork = lib.ath_conversion_workspace_bytes(1, nrow, nslice, ncol,
                                               float(overlap))
    work = (torch.empty(nwork, dtype=torch.uint8, device=dev)
            if nwork else None)
    nat.check(lib.ath_extend_sinogram(
        nat.ptr(src), nrow, nslice, ncol, float(overlap), int(side),
        int(bool(apply_log)), nat.ptr(work) if work is not None else None,
        nwork, nat.ptr(out), nat.stream_ptr(dev)), "ath_extend_sinogram")
    new_cor = ncol - overlap_int / 2.0 - 0.5
    if two_d:
        out = out[:, 0, :]
    if as_numpy:
        out = _host.to_host(out)
    return out, new_cor
