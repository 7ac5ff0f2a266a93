"""Chunked FBP of large volumes, including half-acquisition (offset axis)
360-degree scans -- the chunk loop of algotom's full-reconstruction CLIs
(examples/.../full_reconstruction_cli.py:169-243,
full_reconstruction_360_cli.py:252-276) kept on the GPU.

A chunk of slices is converted 360 -> 180 degrees (ath_convert_sinogram_
360_to_180), filtered and back-projected; only one chunk's stitched and
filtered sinograms exist at a time.  Output goes to a caller array (device
tensor, or a host numpy / memory-mapped array, copied chunk by chunk).  With
both the input and the output on the host, chunk k+1 uploads and chunk k-1
downloads while chunk k computes.
"""

import numpy as np

from algotom_amd import _host
from algotom_amd import _native as nat
from algotom_amd.prep import conversion
from algotom_amd.rec import reconstruction as rec


def reconstruct_volume_360(sino_360, cor, chunk=16, angles=None, ratio=1.0,
                           filter_name="hann", pad=None, pad_mode="edge",
                           apply_log=True, precision=None, out=None,
                           total_width=None):
    """FBP of a (nrow_360, nslice, ncol) half-acquisition stack.

    Returns (recon, new_cor) with recon of shape (width, nslice, width),
    width = max(2*ncol - overlap, total_width).  ``angles`` are those of the
    converted 180-degree sinogram (default: the reference's
    deg2rad(linspace(0, 180, nrow_360 // 2 + 1))).
    """
    import torch
    dev = nat.device()
    nrow, nslice, ncol = tuple(sino_360.shape)
    overlap, side = conversion._overlap_side(ncol, cor)
    ov = int(np.round(overlap + 1.0e-3))
    width = 2 * ncol - ov
    if total_width is not None and total_width > width:
        width = int(total_width)
    new_cor = ncol - ov / 2.0 - 0.5
    n180 = nrow // 2 + 1
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, n180))
    ang = rec._angles_device(angles, dev)

    def run(part):
        stitched, _ = conversion.convert_sinogram_360_to_180(
            part, cor, total_width=width)
        return rec.fbp_reconstruction(stitched, new_cor, ang, ratio=ratio,
                                      filter_name=filter_name, pad=pad,
                                      pad_mode=pad_mode, apply_log=apply_log,
                                      precision=precision)

    ranges = [(j0, min(nslice, j0 + chunk)) for j0 in range(0, nslice, chunk)]
    if out is not None and not rec._is_tensor(out) \
            and not rec._is_tensor(sino_360):
        # host in, host out: transfers overlap the computation
        src = np.asarray(sino_360)
        _host.stream_chunks(src, out, ranges, run, torch.float32, dev)
        return out, new_cor
    if out is None:
        out = torch.empty((width, nslice, width), dtype=torch.float32,
                          device=dev)
    for j0, j1 in ranges:
        part = sino_360[:, j0:j1, :]
        if not rec._is_tensor(part):
            part = torch.from_numpy(np.ascontiguousarray(part,
                                                         dtype=np.float32))
        r = run(part.to(dev, torch.float32).contiguous())
        if rec._is_tensor(out):
            out[:, j0:j1, :] = r
        else:
            out[:, j0:j1, :] = rec._to_host(r)
        del r
    return out, new_cor
