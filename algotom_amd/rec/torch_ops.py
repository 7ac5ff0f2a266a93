"""The device path as registered torch operators (``torch.ops.algotom_amd``).

The drop-in functions call libalgotom_hip through ctypes on raw device
pointers, which the dispatcher cannot see.  These operators wrap the same
calls with ``torch.library.custom_op`` and give each a fake (meta)
implementation, so the dispatcher, FakeTensor tracing and ``torch.compile``
treat them as ordinary functional ops on device tensors:

    algotom_amd::fbp(sinogram, angles, center, ratio, filter_name, pad,
                     pad_mode, apply_log, gpu)
        fbp_reconstruction (reconstruction.py:362-493) of a (nrow, nslice,
        ncol) float32/float64 device stack -> (ncol, nslice, ncol) float32.
        Unlike the drop-in it never writes to its input: values <= 0 are
        replaced in a copy (the drop-in's in-place fix of the caller's
        array, reconstruction.py:431-435, is a side effect an op may not
        have).  ratio < 0 means None (no mask).
    algotom_amd::ramp_filter(sinogram, pad, pad_mode, filter_name,
                             apply_log) -> float32 (nrow, nslice, ncol)
        apply_ramp_filter's arithmetic (reconstruction.py:118-153); -log
        fused when apply_log (values must be > 0).
    algotom_amd::back_projection(filtered, angles, center, acc, mask_radius,
                                 apply_scale, edge_pad)
        the back-projection of ath_back_projection, (ncol, nslice, ncol).
    algotom_amd::forward_projection(image, angles, center, mask_radius,
                                    scale)
        its adjoint (ath_forward_projection_ws), (nrow, nslice, ncol).

Angles are a float tensor (any device); the ops compute on the device of
their first tensor argument.  Importing this module registers the ops.
"""

import numpy as np
import torch

from algotom_amd import _native as nat
from algotom_amd.rec import reconstruction as rec

_LIB = "algotom_amd"


def _check_device(t, name):
    if t.device.type != "cuda":
        raise ValueError("algotom_amd::%s runs on the GPU; got a %s tensor"
                         % (name, t.device.type))


@torch.library.custom_op(_LIB + "::ramp_filter", mutates_args=())
def ramp_filter(sinogram: torch.Tensor, pad: int, pad_mode: str,
                filter_name: str, apply_log: bool) -> torch.Tensor:
    _check_device(sinogram, "ramp_filter")
    nrow, nslice, ncol = sinogram.shape
    name = filter_name or None
    taps = rec._window_taps(None, nrow, ncol + 2 * pad, name)
    return rec._filter_stack(sinogram.contiguous(), pad, pad_mode, apply_log,
                             taps, torch.float32)


@ramp_filter.register_fake
def _(sinogram, pad, pad_mode, filter_name, apply_log):
    return sinogram.new_empty(sinogram.shape, dtype=torch.float32)


@torch.library.custom_op(_LIB + "::back_projection", mutates_args=())
def back_projection(filtered: torch.Tensor, angles: torch.Tensor,
                    center: float, acc: int, mask_radius: float,
                    apply_scale: bool, edge_pad: bool) -> torch.Tensor:
    _check_device(filtered, "back_projection")
    return rec._backproject_stack(filtered.contiguous(), angles, center, acc,
                                  edge_pad=edge_pad, mask_radius=mask_radius,
                                  apply_scale=apply_scale)


@back_projection.register_fake
def _(filtered, angles, center, acc, mask_radius, apply_scale, edge_pad):
    nrow, nslice, ncol = filtered.shape
    return filtered.new_empty((ncol, nslice, ncol), dtype=torch.float32)


@torch.library.custom_op(_LIB + "::forward_projection", mutates_args=())
def forward_projection(image: torch.Tensor, angles: torch.Tensor,
                       center: float, mask_radius: float,
                       scale: float) -> torch.Tensor:
    _check_device(image, "forward_projection")
    ncol, nslice, _ = image.shape
    dev = image.device
    ang = rec._angles_device(angles, dev)
    img = image.contiguous().to(torch.float32)
    out = torch.empty((ang.numel(), nslice, ncol), dtype=torch.float32,
                      device=dev)
    lib = nat.load()
    nbytes = lib.ath_forward_projection_workspace_bytes(nslice, ncol)
    work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    nat.check(lib.ath_forward_projection_ws(
        nat.ptr(img), ang.numel(), nslice, ncol, nat.ptr(ang),
        float(np.float32(center)), float(mask_radius), float(scale),
        nat.ptr(work), nbytes, nat.ptr(out), nat.stream_ptr(dev)),
        "ath_forward_projection_ws")
    return out


@forward_projection.register_fake
def _(image, angles, center, mask_radius, scale):
    ncol, nslice, _ = image.shape
    return image.new_empty((angles.shape[0], nslice, ncol),
                           dtype=torch.float32)


@torch.library.custom_op(_LIB + "::fbp", mutates_args=())
def fbp(sinogram: torch.Tensor, angles: torch.Tensor, center: float,
        ratio: float, filter_name: str, pad: int, pad_mode: str,
        apply_log: bool, gpu: bool) -> torch.Tensor:
    _check_device(sinogram, "fbp")
    nrow, nslice, ncol = sinogram.shape
    if center < 0 or center >= ncol:
        raise ValueError("Center is out of the range [0, {}]".format(ncol - 1))
    if angles.shape[0] != nrow:
        raise ValueError("!!!Number of angles is not the same as the row "
                         "number of the sinogram!!!")
    sino = sinogram.contiguous()
    if apply_log and bool((sino <= 0.0).any()):
        sino = sino.clone()
        rec._clamp_nonpositive(sino)    # warns, as the reference does
    return rec.fbp_reconstruction(sino, center, angles,
                                  ratio=None if ratio < 0 else ratio,
                                  filter_name=filter_name or None,
                                  pad=pad, pad_mode=pad_mode,
                                  apply_log=apply_log, gpu=gpu)


@fbp.register_fake
def _(sinogram, angles, center, ratio, filter_name, pad, pad_mode,
      apply_log, gpu):
    nrow, nslice, ncol = sinogram.shape
    return sinogram.new_empty((ncol, nslice, ncol), dtype=torch.float32)
