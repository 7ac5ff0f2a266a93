"""Flat-field correction on the GPU -- drop-in for
algotom.prep.correction.flat_field_correction (algotom 1.6.0,
algotom/prep/correction.py:45-131) without its artifact-removal options.

The 2D denominator (ratio * flat - dark, zeros replaced by its mean or by
1, :77-98) is prepared on the host with numpy, exactly as the reference
does; the correction of every projection / sinogram row runs in one HIP
kernel in the type numpy's promotion gives (float32 when flat and dark are
float32, float64 otherwise).  Zinger and stripe removal options are other
modules' methods: passing one raises NotImplementedError.
"""

import numpy as np

from algotom_amd import _host
from algotom_amd import _native as nat


def flat_field_correction(proj, flat, dark, ratio=1.0, use_dark=True,
                          **options):
    """Perform flat-field correction.

    Parameters
    ----------
    proj : array_like
        3D or 2D array. Projection images or a sinogram image (numpy,
        float32 / float64 / uint16, or a torch tensor on the GPU).
    flat : array_like
        2D or 1D array. Flat-field image or a single row of it.
    dark : array_like
        2D or 1D array. Dark-field image or a single row of it.
    ratio : float
        Ratio between exposure time used for recording projections
        and exposure time used for recording flat field.
    use_dark : bool
        Subtracting dark field if True.
    options : dict, optional
        Artifact-removal methods; only None values are accepted here.

    Returns
    -------
    array_like
        3D or 2D array. Corrected projections or corrected sinograms.
    """
    import torch
    from algotom_amd.rec import reconstruction as rec
    for opt in options.values():
        if isinstance(opt, dict):
            raise NotImplementedError(
                "artifact-removal options of flat_field_correction are not "
                "on the GPU path")
        if opt is not None:
            raise ValueError("\n Please use the dictionary format: "
                             "options={'method': 'filter_name', 'para1': "
                             "parameter_1, 'para2': parameter_2}")
    flat = ratio * np.asarray(flat)
    dark_h = np.asarray(dark) if use_dark else None
    denom = flat - dark_h if use_dark else flat
    if use_dark:
        denom = np.array(denom)
    if 0.0 in denom:
        nmean = np.mean(denom)
        denom[denom == 0.0] = nmean if nmean != 0.0 else 1
    field = np.result_type(np.float32, denom, *(
        [dark_h] if dark_h is not None else []))
    if field not in (np.float32, np.float64):
        field = np.float64
    dev = nat.device()
    as_numpy = not rec._is_tensor(proj)
    if as_numpy:
        p = np.asarray(proj)
        if p.dtype not in (np.float32, np.float64, np.uint16):
            p = p.astype(np.float32)
        p = np.ascontiguousarray(p)
        if p.dtype == np.uint16:
            pt = torch.from_numpy(p.view(np.int16)).to(dev).view(torch.uint16)
        else:
            pt = torch.from_numpy(p).to(dev)
    else:
        pt = proj.to(dev).contiguous()
        if pt.dtype not in (torch.float32, torch.float64, torch.uint16):
            pt = pt.to(torch.float32)
    shape = tuple(pt.shape)
    width = shape[-1]
    if denom.ndim == 1:
        height, nouter = 1, int(np.prod(shape[:-1]))
        fshape = (1, width)
    else:
        height = denom.shape[-2]
        nouter = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
        fshape = (height, width)
    if np.broadcast_shapes(denom.shape, shape) != shape:
        raise ValueError("flat/dark fields do not broadcast to the data")
    tdt = torch.float64 if field == np.float64 else torch.float32
    den_t = torch.from_numpy(np.ascontiguousarray(
        np.broadcast_to(denom, fshape), dtype=field)).to(dev)
    dark_t = None
    if dark_h is not None:
        dark_t = torch.from_numpy(np.ascontiguousarray(
            np.broadcast_to(dark_h, fshape), dtype=field)).to(dev)
    out = torch.empty(shape, dtype=tdt, device=dev)
    code = {torch.float32: nat.ATH_F32, torch.float64: nat.ATH_F64,
            torch.uint16: nat.ATH_U16}[pt.dtype]
    nat.check(nat.load().ath_flat_field_correction(
        nat.ptr(pt), code, nouter, height, width, height * width, width,
        nat.ptr(dark_t) if dark_t is not None else None, nat.ptr(den_t),
        nat.dtype_code(tdt), nat.ptr(out), nat.stream_ptr(dev)),
        "ath_flat_field_correction")
    return _host.to_host(out) if as_numpy else out
