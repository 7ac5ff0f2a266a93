"""Filtered back-projection on AMD Instinct MI355X -- drop-in for algotom's
``algotom.rec.reconstruction`` FBP functions (algotom 1.6.0).

Same names, arguments, defaults, return shapes, warnings and errors as the
reference; the filter and back-projection run as hand-written HIP kernels
(libalgotom_hip.so) and nothing falls back to the CPU: without a HIP device
every compute entry raises ``RuntimeError``.

* numpy input gives numpy output (host copies in and out, as the reference's
  numba launches do); a torch tensor on the GPU stays on the GPU and a
  tensor comes back.
* ``gpu`` keeps its meaning of "which reference kernel": ``gpu=True`` on a
  single sinogram reproduces ``__back_projection_gpu_kernel`` (float64
  accumulation, reconstruction.py:156-190); a stack with ``gpu=True``
  reproduces ``__back_projection_gpu_chunk_kernel`` (:229-264) and
  ``gpu=False`` reproduces ``back_projection_cpu`` (:307-359), both float32
  storage after every angle.  All of them run on the GPU.
* ``block`` and ``ncore`` are accepted and ignored: tiles are fixed by the
  kernel and there is no CPU worker pool.
* keyword-only ``precision`` (not in the reference): ``None`` keeps the
  reference arithmetic; ``"fp32"`` interpolates and accumulates in float32
  (float64 detector positions); ``"fp16"`` / ``"bf16"`` also store the
  filtered sinogram in 16 bits.
"""

import functools
import math
import threading
import warnings

import numpy as np
from scipy.signal import windows as _windows

from algotom_amd import _cache
from algotom_amd import _host
from algotom_amd import _native as nat
from algotom_amd.rec import autotune

_WINDOWS = ("hann", "bartlett", "blackman", "hamming", "nuttall", "parzen",
            "triang")


# --------------------------------------------------------------------------
# windows (host, float64; reconstruction.py:50-115)
# --------------------------------------------------------------------------
def make_smoothing_window(filter_name, width):
    """Make a 1d smoothing window (reconstruction.py:50-83).

    Parameters
    ----------
    filter_name : {"hann", "bartlett", "blackman", "hamming", "nuttall",
                   "parzen", "triang"}
        Window function; anything else gives ones.
    width : int
        Width of the window.

    Returns
    -------
    array_like
        1D float64 array (scipy.signal.windows, symmetric).
    """
    if filter_name in _WINDOWS:
        return getattr(_windows, filter_name)(width)
    return np.ones(width)


def _ramp_kernel(width):
    # discrete Ram-Lak on n = k - ceil((W-1)/2): 1/4 at 0, 0 at even n,
    # -1/(n pi)^2 at odd n (reconstruction.py:106-111)
    n = np.arange(0.0, width) - np.ceil((width - 1.0) / 2)
    h = np.zeros(width)
    h[n == 0.0] = 0.25
    odd = np.mod(n, 2) == 1.0
    h[odd] = -1.0 / (n[odd] * np.pi) ** 2
    return h


@functools.lru_cache(maxsize=64)
def _ramp_row(width, filter_name):
    row = np.fft.fftshift(np.fft.fft(_ramp_kernel(width)))
    row = row * make_smoothing_window(filter_name, width)
    row.setflags(write=False)
    return row


def make_2d_ramp_window(height, width, filter_name=None):
    """Make the 2d ramp window in Fourier space (reconstruction.py:86-115).

    Parameters
    ----------
    height : int
        Height of the window.
    width : int
        Width of the window.
    filter_name : {None, "hann", "bartlett", "blackman", "hamming",
                   "nuttall", "parzen", "triang"}
        Name of a smoothing window used.

    Returns
    -------
    complex ndarray
        2D array, every row the windowed ramp spectrum (fftshift order).
    """
    return np.tile(_ramp_row(int(width), filter_name), (height, 1))


def _taps(rows, outer_shift=True):
    """Real circular-convolution taps equivalent to multiplying the shifted
    spectrum by ``rows`` and keeping the real part (reconstruction.py
    :150-152): Re(ifftshift(ifft(ifftshift(R)))) per row.  With
    outer_shift=False, Re(ifft(ifftshift(R))): the form without the final
    ifftshift, which vertrec's filter after back-projection uses with the
    magnitude |R| (vertrec.py:621-627)."""
    rows = np.asarray(rows)
    g = np.fft.ifft(np.fft.ifftshift(rows, axes=-1), axis=-1)
    if outer_shift:
        g = np.fft.ifftshift(g, axes=-1)
    return np.ascontiguousarray(np.real(g), dtype=np.float64)


# --------------------------------------------------------------------------
# device plumbing
# --------------------------------------------------------------------------
def _is_tensor(x):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


def _to_device(x, dev, allow_f64=True):
    """Contiguous float32/float64 tensor on ``dev`` (no copy if already so)."""
    import torch
    if _is_tensor(x):
        t = x
    else:
        a = np.asarray(x)
        if a.dtype not in (np.float32, np.float64):
            a = a.astype(np.float32)
        t = torch.from_numpy(np.ascontiguousarray(a))
    if t.dtype not in (torch.float32, torch.float64) or (
            t.dtype == torch.float64 and not allow_f64):
        t = t.to(torch.float32)
    return t.to(dev).contiguous()


def _angles_device(angles, dev):
    import torch
    if _is_tensor(angles):
        return angles.detach().to(dev, torch.float32).contiguous()
    a = np.ascontiguousarray(np.asarray(angles, dtype=np.float64)
                             .astype(np.float32))
    return torch.from_numpy(a).to(dev)


_plan_cache = _cache.DeviceCache(32)


def _filter_plan(taps, dev, fft_len):
    """Device filter plan for ``taps`` (nwin, W) on FFT length ``fft_len``;
    cached per device (thread-safe LRU of 32 plans)."""
    import torch
    nwin, width = taps.shape
    key = (dev.index, width, fft_len, taps.tobytes())

    def build():
        lib = nat.load()
        nbytes = lib.ath_filter_plan_bytes(fft_len, nwin)
        if nbytes == 0:
            raise ValueError("ramp filter width %d is too large" % width)
        plan = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        taps_d = torch.from_numpy(np.ascontiguousarray(taps)).to(dev)
        nat.check(lib.ath_filter_plan_build(nat.ptr(taps_d), nwin, width,
                                            fft_len, nat.ptr(plan), nbytes,
                                            nat.stream_ptr(dev)),
                  "ath_filter_plan_build")
        return (plan, nwin, taps_d)

    return _plan_cache.get(key, build, lambda e: (e[0], e[2]))


def _window_taps(ramp_win, nrow, width, filter_name):
    """Taps (nwin, width) for the window the reference would use."""
    if ramp_win is None or np.shape(ramp_win) != (nrow, width):
        # reconstruction.py:148-149: regenerate when the shape mismatches
        return _taps(_ramp_row(width, filter_name))[None, :]
    win = np.asarray(ramp_win)
    if np.all(win == win[0:1]):
        return _taps(win[0])[None, :]
    return _taps(win)


def _clamp_nonpositive_device(sinogram, sino_t):
    """reconstruction.py:431-435 for an input already on the device as
    ``sino_t``: the test runs there; if any value is <= 0, warn and set
    those values to 1 in ``sino_t`` and, in place, in the caller's array."""
    bad = sino_t <= 0.0
    if bool(bad.any()):
        warnings.warn("!!!Applying logarithm is enabled but "
                      "there are values <= 0.0 in the data!!!")
        sino_t[bad] = 1.0
        if _is_tensor(sinogram):
            if sinogram.data_ptr() != sino_t.data_ptr():
                sinogram[sinogram <= 0.0] = 1.0
        else:
            sinogram[sinogram <= 0.0] = np.float32(1.0)


def _clamp_nonpositive(sinogram):
    """reconstruction.py:431-435: warn and set values <= 0 to 1, in place."""
    if _is_tensor(sinogram):
        bad = sinogram <= 0.0
        if bool(bad.any()):
            warnings.warn("!!!Applying logarithm is enabled but "
                          "there are values <= 0.0 in the data!!!")
            sinogram[bad] = 1.0
    elif np.any(sinogram <= 0.0):
        warnings.warn("!!!Applying logarithm is enabled but "
                      "there are values <= 0.0 in the data!!!")
        sinogram[sinogram <= 0.0] = np.float32(1.0)


def _filter_stack(sino_t, pad, pad_mode, apply_log, taps, out_dtype):
    """Run the ramp-filter kernel on a (nrow, nslice, ncol) device tensor."""
    import torch
    if pad_mode not in nat.FILTER_PAD_MODES:
        raise NotImplementedError(
            "pad_mode %r is not supported by the GPU filter (supported: %s)"
            % (pad_mode, ", ".join(nat.FILTER_PAD_MODES)))
    dev = sino_t.device
    nrow, nslice, ncol = sino_t.shape
    lib = nat.load()
    fft_len = nat.check(lib.ath_filter_fft_length(ncol, pad),
                        "ath_filter_fft_length")
    plan, nwin, _ = _filter_plan(taps, dev, fft_len)
    out = torch.empty((nrow, nslice, ncol), dtype=out_dtype, device=dev)
    nat.check(lib.ath_ramp_filter(nat.ptr(sino_t), nat.dtype_code(sino_t.dtype),
                                  nrow, nslice, ncol, int(pad),
                                  nat.FILTER_PAD_MODES[pad_mode],
                                  int(bool(apply_log)),
                                  nat.ptr(plan), fft_len, nwin, nat.ptr(out),
                                  nat.dtype_code(out_dtype),
                                  nat.stream_ptr(dev)),
              "ath_ramp_filter")
    return out


def _backproject_stack(filt_t, angles, center, acc, edge_pad=False,
                       mask_radius=-1.0, apply_scale=False):
    import torch
    dev = filt_t.device
    nrow, nslice, ncol = filt_t.shape
    ang = _angles_device(angles, dev)
    if ang.numel() != nrow:
        raise ValueError("!!!Number of angles is not the same as the row "
                         "number of the sinogram!!!")
    out = torch.empty((ncol, nslice, ncol), dtype=torch.float32, device=dev)
    lib = nat.load()
    args = (nat.ptr(filt_t), nat.dtype_code(filt_t.dtype), nrow, nslice, ncol,
            nat.ptr(ang), float(np.float32(center)), int(bool(edge_pad)),
            int(acc), float(mask_radius), int(bool(apply_scale)), nat.ptr(out))
    wide = filt_t.dtype in (torch.float16, torch.bfloat16)
    key = autotune.make_key("stack", acc, str(filt_t.dtype).split(".")[-1],
                            ncol, nslice)
    cfg = autotune.lookup(key, autotune.compiled_configs("stack", wide))
    if cfg is not None:
        st = lib.ath_back_projection_tiled(*args, int(cfg[0]), int(cfg[1]),
                                           nat.stream_ptr(dev))
        if st != nat.ATH_ERR_UNSUPPORTED:
            nat.check(st, "ath_back_projection_tiled")
            return out
        # a tuned configuration this library does not compile: forget it
        # and use the built-in default
        autotune.forget(key)
    nat.check(lib.ath_back_projection(*args, nat.stream_ptr(dev)),
              "ath_back_projection")
    return out


_to_host = _host.to_host

# Host (numpy) stacks of _STREAM_MIN bytes or more run in chunks of slices
# through a pipeline that overlaps the uploads and downloads of neighbouring
# chunks with the computation; each chunk carries about _CHUNK_BYTES of
# input plus output.
_STREAM_MIN = 256 << 20
_CHUNK_BYTES = 384 << 20


def _streams(sinogram, out_width, out_itemsize=4):
    nrow, nslice, ncol = sinogram.shape
    nbytes = nslice * (4 * nrow * ncol + out_itemsize * out_width * out_width)
    return nslice > 1 and nbytes >= _STREAM_MIN


def _chunk_ranges(nslice, slice_bytes):
    per = max(1, min(nslice, _CHUNK_BYTES // max(1, slice_bytes)))
    if per >= 8:
        per -= per % 4
    return [(j, min(nslice, j + per)) for j in range(0, nslice, per)]


def _stream_numpy(sinogram, out_shape, run, apply_log, dev, allow_f64,
                  out_dtype=np.float32):
    """``run`` over slice chunks of a host (nrow, nslice, ncol) stack with
    overlapped transfers; returns the new (a, nslice, b) array.  With
    ``apply_log`` the values <= 0 are set to 1 on the device, and, as
    reconstruction.py:431-435 does, a warning is given and the caller's
    array is fixed in place when there were any."""
    import torch
    a = np.asarray(sinogram)
    in_dtype = (torch.float64 if allow_f64 and a.dtype == np.float64
                else torch.float32)
    out = np.empty(out_shape, dtype=out_dtype)
    nrow, nslice, ncol = a.shape
    ranges = _chunk_ranges(nslice, 4 * nrow * ncol + out.itemsize
                           * out_shape[0] * out_shape[2])
    flags = []

    def step(part):
        if apply_log:
            bad = part <= 0.0
            flags.append(bad.any())
            part.masked_fill_(bad, 1.0)
        return run(part)

    _host.stream_chunks(a, out, ranges, step, in_dtype, dev)
    if flags and bool(torch.stack(flags).any()):
        warnings.warn("!!!Applying logarithm is enabled but "
                      "there are values <= 0.0 in the data!!!")
        sinogram[sinogram <= 0.0] = np.float32(1.0)
    return out


def _finish(out_t, as_numpy):
    if as_numpy:
        return _to_host(out_t)
    return out_t


_PRECISIONS = {None: None, "fp32": "fp32", "fp16": "fp16", "bf16": "bf16"}


def _storage_dtype(precision):
    import torch
    return {"fp16": torch.float16, "bf16": torch.bfloat16}.get(
        precision, torch.float32)


# --------------------------------------------------------------------------
# public API
# --------------------------------------------------------------------------
def apply_ramp_filter(sinogram, ramp_win=None, filter_name=None, pad=None,
                      pad_mode="edge"):
    """Apply the ramp filter to a sinogram (reconstruction.py:118-153).

    Parameters
    ----------
    sinogram : array_like
        2D array (nrow, ncol). A 3D (nrow, nslice, ncol) stack is filtered
        row by row as well.
    ramp_win : complex ndarray or None
        Ramp window in the Fourier space, shape (nrow, ncol + 2*pad);
        regenerated if None or of another shape.
    filter_name : {None, "hann", "bartlett", "blackman", "hamming",
                   "nuttall", "parzen", "triang"}
        Name of a smoothing window used.
    pad : int or None
        Padding before the FFT; min(int(0.15*ncol), 150) if None.
    pad_mode : str
        numpy.pad mode: "edge", "constant", "reflect", "symmetric", "wrap",
        "linear_ramp", "maximum", "mean", "median", "minimum" (numpy's
        defaults for the last five).

    Returns
    -------
    array_like
        Filtered sinogram, float64, same shape as the input.
    """
    import torch
    as_numpy = not _is_tensor(sinogram)
    dev = nat.device()
    shape = tuple(sinogram.shape)
    nrow, ncol = shape[0], shape[-1]
    if pad is None:
        pad = min(int(0.15 * ncol), 150)
    width = ncol + 2 * pad
    taps = _window_taps(ramp_win, nrow, width, filter_name)
    sino_t = _to_device(sinogram, dev)
    stack = sino_t.reshape(nrow, -1, ncol)
    out = _filter_stack(stack, pad, pad_mode, False, taps, torch.float64)
    return _finish(out.reshape(shape), as_numpy)


def back_projection_gpu(sinogram, angles, center, block=(16, 16),
                        edge_pad=False):
    """Back-projection of one sinogram with float64 accumulation
    (reconstruction.py:156-226).  Returns (ncol, ncol) float32."""
    dev = nat.device()
    as_numpy = not _is_tensor(sinogram)
    nrow, ncol = tuple(sinogram.shape)
    t = _to_device(sinogram, dev, allow_f64=False).reshape(nrow, 1, ncol)
    out = _backproject_stack(t, angles, center, nat.ACC_F64, edge_pad)
    return _finish(out.reshape(ncol, ncol), as_numpy)


def back_projection_gpu_chunk(sinograms, angles, center, block=(16, 16),
                              edge_pad=False):
    """Back-projection of a (nrow, nslice, ncol) stack with float32 storage
    after every angle (reconstruction.py:229-304).  Returns
    (ncol, nslice, ncol) float32."""
    dev = nat.device()
    as_numpy = not _is_tensor(sinograms)
    t = _to_device(sinograms, dev, allow_f64=False)
    out = _backproject_stack(t, angles, center, nat.ACC_F32SEQ, edge_pad)
    return _finish(out, as_numpy)


def back_projection_cpu(sinogram, angles, center, edge_pad=False):
    """The arithmetic of the reference's CPU back-projector
    (reconstruction.py:307-359: float32 storage after every angle), computed
    on the GPU without its prange race.  Returns (ncol, ncol) float32."""
    dev = nat.device()
    as_numpy = not _is_tensor(sinogram)
    nrow, ncol = tuple(sinogram.shape)
    t = _to_device(sinogram, dev, allow_f64=False).reshape(nrow, 1, ncol)
    out = _backproject_stack(t, angles, center, nat.ACC_F32SEQ, edge_pad)
    return _finish(out.reshape(ncol, ncol), as_numpy)


def fbp_reconstruction(sinogram, center, angles=None, ratio=1.0,
                       ramp_win=None, filter_name="hann", pad=None,
                       pad_mode="edge", apply_log=True, gpu=True,
                       block=(16, 16), ncore=None, *, precision=None):
    """Apply the FBP (filtered back-projection) reconstruction method to a
    sinogram-image or a chunk of sinogram-images (reconstruction.py:362-493).
    Angular axis is 0; for 3D input the slicing axis is 1.

    Parameters
    ----------
    sinogram : array_like
        2D/3D array. Sinogram image(s); numpy or a torch tensor.
    center : float
        Center of rotation.
    angles : array_like, optional
        1D array. List of angles (in radian) corresponding to the sinogram.
    ratio : float, optional
        To apply a circle mask to the reconstructed image.
    ramp_win : complex ndarray, optional
        Ramp window in the Fourier space. Generated if None.
    filter_name : {None, "hann", "bartlett", "blackman", "hamming",
                  "nuttall", "parzen", "triang"}
        Apply a smoothing filter.
    pad : int, optional
        To apply padding before the FFT; min(int(0.15*ncol), 150) if None.
    pad_mode : str, optional
        numpy.pad mode: "edge", "constant", "reflect", "symmetric", "wrap",
        "linear_ramp", "maximum", "mean", "median", "minimum" (numpy's
        defaults for the last five).
    apply_log : bool, optional
        Apply the logarithm function to the sinogram before reconstruction.
    gpu : bool, optional
        Which reference back-projector's arithmetic to reproduce (see the
        module docstring); the computation always runs on the GPU.
    block : tuple of two integer-values, optional
        Accepted for compatibility; ignored.
    ncore : int or None
        Accepted for compatibility; ignored.
    precision : {None, "fp32", "fp16", "bf16"}, keyword only
        None keeps the reference arithmetic.

    Returns
    -------
    array_like
        Square array(s): (ncol, ncol) for 2D input, (ncol, nslice, ncol)
        for 3D input, float32.
    """
    if precision not in _PRECISIONS:
        raise ValueError("precision must be None, 'fp32', 'fp16' or 'bf16'")
    shape = tuple(sinogram.shape)
    input_3d = len(shape) == 3
    if input_3d:
        (nrow, num_sino, ncol) = shape
    else:
        num_sino = 1
        (nrow, ncol) = shape
    if center < 0 or center >= ncol:
        raise ValueError("Center is out of the range [0, {}]".format(ncol - 1))
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, nrow))
    elif len(angles) != nrow:
        raise ValueError("!!!Number of angles is not the same as the row "
                         "number of the sinogram!!!")
    dev = nat.device()
    as_numpy = not _is_tensor(sinogram)
    mask_radius = -1.0
    if ratio is not None:
        if ratio == 0.0:
            ratio = min(center, ncol - center) / (0.5 * ncol)
        mask_radius = float(ratio * (ncol // 2))
    if pad is None:
        pad = min(int(0.15 * ncol), 150)
    taps = _window_taps(ramp_win, nrow, ncol + 2 * pad, filter_name)
    if precision is None:
        acc = nat.ACC_F64 if (gpu and num_sino == 1) else nat.ACC_F32SEQ
    else:
        acc = nat.ACC_F32
    storage = _storage_dtype(precision)
    ang = _angles_device(angles, dev)

    def run(sino_t):
        filt = _filter_stack(sino_t, pad, pad_mode, apply_log, taps, storage)
        return _backproject_stack(filt, ang, center, acc,
                                  mask_radius=mask_radius, apply_scale=True)

    if as_numpy and input_3d and _streams(sinogram, ncol):
        return _stream_numpy(sinogram, (ncol, num_sino, ncol), run,
                             apply_log, dev, allow_f64=True)
    sino_t = _to_device(sinogram, dev).reshape(nrow, num_sino, ncol)
    if apply_log:
        _clamp_nonpositive_device(sinogram, sino_t)
    out = run(sino_t)
    if not input_3d:
        out = out.reshape(ncol, ncol)
    return _finish(out, as_numpy)


# --------------------------------------------------------------------------
# BPF: back-projection filtering (reconstruction.py:494-684)
# --------------------------------------------------------------------------
def make_circular_ramp_window(width, filter_name=None):
    """Make a circular ramp window (2d) with the option of adding a smoothing
    window (reconstruction.py:494-521).

    Parameters
    ----------
    width : int
        Width of the window.
    filter_name : {None, "hann", "bartlett", "blackman", "hamming",
                   "nuttall", "parzen", "triang"}
        Name of a smoothing window used.

    Returns
    -------
    array_like
        Square float64 array, size (width, width): |windowed ramp spectrum|
        spun around the centre.
    """
    return np.array(_circular_window(int(width), filter_name))


@functools.lru_cache(maxsize=4)
def _circular_window(width, filter_name):
    from algotom_amd.util import utility
    ramp_1d = np.abs(np.fft.fftshift(np.fft.fft(_ramp_kernel(width))))
    ramp_1d = ramp_1d * make_smoothing_window(filter_name, width)
    win = utility.transform_1d_window_to_2d(ramp_1d, order=1, mode="nearest")
    win.setflags(write=False)
    return win


_circ_cache = _cache.DeviceCache(1)


def _circular_plan(ramp_win, dev, name_key=None):
    """Host kernel of a W x W circular window, with the device plans built
    from it (one per FFT length): the filter is the circular convolution
    with g2 = Re(ifft2(ifftshift(R2))) (reconstruction.py:555-559),
    evaluated by ath_circular_ramp_filter.  One window is kept resident
    (a plan holds the kernel's 2D spectrum: 537 MB at FFT length 8192,
    2.1 GB at 16384, 8.6 GB at 32768).
    Default windows are keyed by (width, filter_name); a caller's window by
    a hash of its values."""
    import hashlib
    win = np.ascontiguousarray(ramp_win, dtype=np.float64)
    width = win.shape[0]
    if name_key is not None:
        key = (dev.index, width) + tuple(name_key)
    else:
        key = (dev.index, width, hashlib.sha1(win.tobytes()).hexdigest())

    def build():
        g2 = np.ascontiguousarray(np.real(np.fft.ifft2(
            np.fft.ifftshift(win))))
        return {"width": width, "g2": g2, "plans": {},
                "lock": threading.Lock()}

    return _circ_cache.get(key, build)


def _circular_plan_for(entry, fft_len, dev):
    import torch
    with entry["lock"]:
        plan = entry["plans"].get(fft_len)
        if plan is not None:
            home = entry["homes"][fft_len]
            cur = torch.cuda.current_stream(dev)
            if cur.cuda_stream != home:
                plan.record_stream(cur)
            return plan
        lib = nat.load()
        width = entry["width"]
        nbytes = lib.ath_circular_filter_plan_bytes(fft_len)
        nwork = lib.ath_circular_filter_plan_work_bytes(width, fft_len)
        if nbytes == 0 or nwork == 0:
            raise NotImplementedError(
                "the circular ramp filter supports FFT lengths up to 32768 "
                "(width %d needs %d)" % (width, fft_len))
        plan = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        work = torch.empty(nwork, dtype=torch.uint8, device=dev)
        g2_d = torch.from_numpy(entry["g2"]).to(dev)
        nat.check(lib.ath_circular_filter_plan_build(
            nat.ptr(g2_d), width, fft_len, nat.ptr(plan), nbytes,
            nat.ptr(work), nwork, nat.stream_ptr(dev)),
            "ath_circular_filter_plan_build")
        entry["plans"][fft_len] = plan
        entry.setdefault("homes", {})[fft_len] = \
            torch.cuda.current_stream(dev).cuda_stream
        return plan


_CIRC_WORK_BYTES = 4 << 30


def _circular_filter(images, ncol, nimg, row_stride, img_stride, pad,
                     pad_mode, ramp_win, out, out_row_stride, out_img_stride,
                     mask_radius=-1.0, scale_nrow=0, name_key=None):
    """Run ath_circular_ramp_filter over nimg images of a device tensor,
    in groups that keep the workspace below 4 GiB."""
    import torch
    if pad_mode not in nat.PAD_MODES:
        raise NotImplementedError(
            "pad_mode %r is not supported by the GPU filter (supported: %s)"
            % (pad_mode, ", ".join(nat.PAD_MODES)))
    dev = images.device
    lib = nat.load()
    fft_len = nat.check(lib.ath_filter_fft_length(ncol, pad),
                        "ath_filter_fft_length")
    entry = _circular_plan(ramp_win, dev, name_key)
    plan = _circular_plan_for(entry, fft_len, dev)
    per_img = lib.ath_circular_ramp_filter_workspace_bytes(ncol, pad,
                                                           fft_len, 1)
    group = int(max(1, min(nimg, _CIRC_WORK_BYTES // max(per_img, 1))))
    work = torch.empty(per_img * group, dtype=torch.uint8, device=dev)
    esz_in = images.element_size()
    esz_out = out.element_size()
    for j0 in range(0, nimg, group):
        n = min(group, nimg - j0)
        nat.check(lib.ath_circular_ramp_filter(
            images.data_ptr() + j0 * img_stride * esz_in, n, ncol,
            row_stride, img_stride, int(pad), nat.PAD_MODES[pad_mode],
            nat.ptr(plan), fft_len, nat.ptr(work), per_img * n,
            out.data_ptr() + j0 * out_img_stride * esz_out,
            nat.dtype_code(out.dtype), out_row_stride, out_img_stride,
            float(mask_radius), int(scale_nrow), nat.stream_ptr(dev)),
            "ath_circular_ramp_filter")
    return out


def apply_circular_ramp_filter(rec_img, ramp_win=None, filter_name=None,
                               pad=None, pad_mode="edge"):
    """Apply the circular ramp filter to a back-projected image
    (reconstruction.py:524-560).

    Parameters
    ----------
    rec_img : array_like
        Square array. back-projected image (numpy or a GPU tensor).
    ramp_win : array_like
        2d circular ramp window, generated if None given or if its shape is
        not the padded image's.
    filter_name : {None, "hann", "bartlett", "blackman", "hamming",
                   "nuttall", "parzen", "triang"}
        Name of a smoothing window used.
    pad : int or None
        Padding before the FFT; min(int(0.15 * width), 150) if None.
    pad_mode : str
        numpy.pad mode: "edge", "constant", "reflect", "symmetric", "wrap".

    Returns
    -------
    array_like
        Square float64 array.
    """
    import torch
    ncol = rec_img.shape[0]
    if pad is None:
        pad = min(int(0.15 * ncol), 150)
    width = ncol + 2 * pad
    name_key = None
    if ramp_win is None or np.shape(ramp_win) != (width, width):
        ramp_win = _circular_window(width, filter_name)
        name_key = ("window", filter_name)
    dev = nat.device()
    as_numpy = not _is_tensor(rec_img)
    img = _to_device(rec_img, dev, allow_f64=False)
    out = torch.empty((ncol, ncol), dtype=torch.float64, device=dev)
    _circular_filter(img, ncol, 1, ncol, ncol * ncol, pad, pad_mode,
                     ramp_win, out, ncol, ncol * ncol, name_key=name_key)
    return _finish(out, as_numpy)


def _minus_log(sino_t):
    import torch
    out = torch.empty_like(sino_t)
    nat.check(nat.load().ath_minus_log(nat.ptr(sino_t), nat.ptr(out),
                                       sino_t.numel(),
                                       nat.stream_ptr(sino_t.device)),
              "ath_minus_log")
    return out


def bpf_reconstruction(sinogram, center, angles=None, ratio=1.0,
                       ramp_win=None, filter_name="hann", pad=None,
                       pad_mode="edge", apply_log=True, gpu=True,
                       block=(16, 16), ncore=None):
    """Apply the BPF (back-projection filtering) reconstruction method to a
    sinogram-image or a chunk of sinogram-images (reconstruction.py:563-684).
    Angular axis is 0; for 3D input the slicing axis is 1.

    The sinogram (logged if apply_log) is back-projected with edge padding,
    then every slice goes through the 2D circular ramp filter.  ``gpu``
    selects the reference back-projector's arithmetic as in
    fbp_reconstruction; ``block`` and ``ncore`` are ignored.

    Returns
    -------
    array_like
        (ncol, ncol) float64 for one sinogram (the reference's 2D branch,
        also for a (nrow, 1, ncol) stack, returned as (ncol, 1, ncol)), or
        (ncol, nslice, ncol) float32 for a stack.
    """
    import torch
    shape = tuple(sinogram.shape)
    input_3d = len(shape) == 3
    if input_3d:
        (nrow, num_sino, ncol) = shape
    else:
        num_sino = 1
        (nrow, ncol) = shape
    if center < 0 or center >= ncol:
        raise ValueError("Center is out of the range [0, {}]".format(ncol - 1))
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, nrow))
    elif len(angles) != nrow:
        raise ValueError("!!!Number of angles is not the same as the row "
                         "number of the sinogram!!!")
    dev = nat.device()
    as_numpy = not _is_tensor(sinogram)
    mask_radius = -1.0
    if ratio is not None:
        if ratio == 0.0:
            ratio = min(center, ncol - center) / (0.5 * ncol)
        mask_radius = float(ratio * (ncol // 2))
    if pad is None:
        pad = min(int(0.15 * ncol), 150)
    width = ncol + 2 * pad
    name_key = None
    if ramp_win is None:
        ramp_win = _circular_window(width, filter_name)
        name_key = ("window", filter_name)
    elif np.shape(ramp_win) != (width, width):
        # apply_circular_ramp_filter regenerates the window; the 2D branch
        # calls it without filter_name (:662), the 3D branch with it (:675)
        regen = filter_name if num_sino > 1 else None
        ramp_win = _circular_window(width, regen)
        name_key = ("window", regen)
    acc = nat.ACC_F64 if (gpu and num_sino == 1) else nat.ACC_F32SEQ
    ang = _angles_device(angles, dev)

    def run(sino_t):
        if apply_log:
            sino_t = _minus_log(sino_t)
        bp = _backproject_stack(sino_t, ang, center, acc, edge_pad=True)
        if num_sino == 1:
            out = torch.empty((ncol, ncol), dtype=torch.float64, device=dev)
            _circular_filter(bp, ncol, 1, ncol, ncol * ncol, pad, pad_mode,
                             ramp_win, out, ncol, ncol * ncol, mask_radius,
                             nrow, name_key)
            return out.reshape(ncol, 1, ncol) if input_3d else out
        ns = sino_t.shape[1]
        out = torch.empty((ncol, ns, ncol), dtype=torch.float32, device=dev)
        _circular_filter(bp, ncol, ns, ns * ncol, ncol, pad, pad_mode,
                         ramp_win, out, ns * ncol, ncol, mask_radius, nrow,
                         name_key)
        return out

    if as_numpy and input_3d and _streams(sinogram, ncol):
        return _stream_numpy(sinogram, (ncol, num_sino, ncol), run,
                             apply_log, dev, allow_f64=False)
    sino_t = _to_device(sinogram, dev, allow_f64=False).reshape(
        nrow, num_sino, ncol)
    if apply_log:
        _clamp_nonpositive_device(sinogram, sino_t)
    return _finish(run(sino_t), as_numpy)


# --------------------------------------------------------------------------
# DFI (direct Fourier inversion), reconstruction.py:687-907
# --------------------------------------------------------------------------
_dfi_plans = _cache.DeviceCache(8)


def _dfi_plan(width, filter_name, dev):
    """Device plan of ath_dfi_reconstruction for padded width ``width``:
    Bluestein tables, the smoothing window (reconstruction.py:875-878) and
    the B-spline prefilter taps; a few are kept per device (thread-safe
    LRU)."""
    import torch
    key = (dev.index, width, filter_name)

    def build():
        lib = nat.load()
        nbytes = lib.ath_dfi_plan_bytes(width)
        if nbytes == 0:
            raise NotImplementedError(
                "dfi_reconstruction on the GPU supports padded widths "
                "ncol + 2*int(pad_rate*ncol) up to 16383 (got %d)" % width)
        plan = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        win = None
        if filter_name is not None:
            win = torch.from_numpy(np.ascontiguousarray(
                make_smoothing_window(filter_name, width),
                dtype=np.float64)).to(dev)
        nat.check(lib.ath_dfi_plan_build(
            nat.ptr(win) if win is not None else None, width, nat.ptr(plan),
            nbytes, nat.stream_ptr(dev)), "ath_dfi_plan_build")
        return plan

    return _dfi_plans.get(key, build, lambda p: (p,))


def _dfi_angles(angles, nrow):
    """Angle handling of __dfi_handle_angles (reconstruction.py:731-746):
    (mode, rows, row shift, flip), the scalars in numpy's arithmetic."""
    if len(angles) != nrow:
        raise ValueError("!!! Number of angles is not the same as the row "
                         "number of the sinogram !!!")
    angles = np.asarray(angles)
    t_ang = np.sum(np.abs(np.diff(angles * 180.0 / np.pi)))
    fold = abs(t_ang - 360) < 10
    rows = nrow // 2 + 1 if fold else nrow
    step = np.mean(np.abs(np.diff(angles)))
    b_ang = angles[0] - (angles[0] // (2 * np.pi)) * (2 * np.pi)
    mode = nat.DFI_FOLD_360 if fold else nat.DFI_ANGLES
    return mode, rows, float(b_ang / step), int(bool(angles[-1] < angles[0]))


def dfi_reconstruction(sinogram, center, angles=None, ratio=1.0,
                       filter_name="hann", pad_rate=0.25, pad_mode="edge",
                       apply_log=True, ncore=None):
    """Apply the DFI (direct Fourier inversion) reconstruction method to a
    sinogram-image or a chunk of sinogram-images (reconstruction.py:809-907).
    Angular axis is 0; for 3D input the slicing axis is 1.

    Every slice runs on the GPU through ath_dfi_reconstruction: the odd-width
    sinogram shifted to the centre (cubic B-spline), optional angle handling
    (360-degree fold and first-angle offset), padding by int(pad_rate *
    width), -log, row FFTs, the smoothing window, order-5 B-spline gridding
    of the polar spectrum onto the Cartesian grid, and the inverse 2D FFT,
    all float64 after the shift.  ``ncore`` is ignored.

    Returns
    -------
    array_like
        (ncol, ncol) float64, or (ncol, nslice, ncol) float64 for 3D input
        ((ncol, 1, ncol) for a one-slice stack).
    """
    import torch
    shape = tuple(sinogram.shape)
    input_3d = len(shape) == 3
    nrow, ncol = shape[0], shape[-1]
    num_sino = shape[1] if input_3d else 1
    if center < 0 or center >= ncol:
        raise ValueError("Center is out of the range [0, {}]".format(ncol - 1))
    if ncol / nrow > 5.0:
        warnings.warn("!!!Sinogram is significantly undersampled!!! "
                      "Considering to use the 'upsample_sinogram' method "
                      "before reconstruction!")
    if pad_mode not in nat.PAD_MODES:
        raise NotImplementedError(
            "pad_mode %r is not supported by the GPU path (supported: %s)"
            % (pad_mode, ", ".join(nat.PAD_MODES)))
    ncol1 = ncol + 1 if ncol % 2 == 0 else ncol
    xshift = (ncol1 - 1) / 2.0 - center
    pad = int(pad_rate * ncol1)
    mode, rows, row_shift, flip = nat.DFI_NO_ANGLES, nrow, 0.0, 0
    if angles is not None:
        mode, rows, row_shift, flip = _dfi_angles(angles, nrow)
    width = ncol1 + 2 * pad
    mask_r2 = -1.0
    if ratio is not None:
        if ratio == 0.0:
            ratio = min(center, ncol - center) / (0.5 * ncol)
        radius = ratio * (ncol // 2)
        mask_r2 = float(radius * radius)
    dev = nat.device()
    as_numpy = not _is_tensor(sinogram)
    plan = _dfi_plan(width, filter_name, dev)
    lib = nat.load()
    nws = lib.ath_dfi_workspace_bytes(nrow, ncol, rows, width)
    work = torch.empty(nws, dtype=torch.uint8, device=dev)
    flags = []

    def run(sino_t):
        ns = sino_t.shape[1]
        out = torch.empty((ncol, ns, ncol), dtype=torch.float64, device=dev)
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        nat.check(lib.ath_dfi_reconstruction(
            nat.ptr(sino_t), nat.dtype_code(sino_t.dtype), nrow, ns, ncol,
            ns * ncol, ncol, float(xshift), pad, nat.PAD_MODES[pad_mode],
            int(bool(apply_log)), mode, row_shift, flip, rows, nat.ptr(plan),
            width, nat.ptr(work), nws, nat.ptr(out), nat.ATH_F64, ns * ncol,
            ncol, mask_r2, nat.ptr(flag), nat.stream_ptr(dev)),
            "ath_dfi_reconstruction")
        flags.append(flag)
        return out

    if as_numpy and input_3d and _streams(sinogram, ncol, 8):
        out = _stream_numpy(sinogram, (ncol, num_sino, ncol), run, False,
                            dev, allow_f64=True, out_dtype=np.float64)
    else:
        sino_t = _to_device(sinogram, dev).reshape(nrow, num_sino, ncol)
        out = run(sino_t)
        if not input_3d:
            out = out.reshape(ncol, ncol)
        out = _finish(out, as_numpy)
    if apply_log and bool(torch.cat(flags).any()):
        warnings.warn("!!!Applying logarithm is enabled but there are "
                      "values <= 0.0 in the data!!!")
    return out


# Centre search by slice metrics (reconstruction.py:1254-1458), implemented
# in algotom_amd.rec.center and reachable here, where the reference has it.
def _find_center_based_slice_metric(*args, **kwargs):
    """algotom_amd.rec.center._find_center_based_slice_metric
    (reconstruction.py:1254-1362)."""
    from algotom_amd.rec import center
    return center._find_center_based_slice_metric(*args, **kwargs)


def find_center_based_slice_metric(*args, **kwargs):
    """algotom_amd.rec.center.find_center_based_slice_metric
    (reconstruction.py:1365-1458)."""
    from algotom_amd.rec import center
    return center.find_center_based_slice_metric(*args, **kwargs)
