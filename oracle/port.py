"""ctypes binding of the oracle's C port -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
The C source (oracle/fbp_port.c) restates algotom/rec/reconstruction.py
:118-153 and :156-359 with numba's typing; see its header.
"""

import ctypes
import os

import numpy as np

from oracle import fbp_np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfbp_port.so")
_lib = None

PAD_MODES = {"edge": 0, "constant": 1, "reflect": 2, "symmetric": 3,
             "wrap": 4}


def build():
    """Compile oracle/fbp_port.c with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle port not built: run make -C oracle")
        _lib = ctypes.CDLL(_LIB_PATH)
        p = ctypes.c_void_p
        i = ctypes.c_int
        _lib.orc_trig_tables.argtypes = [p, i, p, p]
        _lib.orc_backproject.argtypes = [p, i, i, i, p, p, ctypes.c_float,
                                         i, i, p, i]
        _lib.orc_ramp_filter.argtypes = [p, i, i, i, i, i, i, p, p, i]
        _lib.orc_finalize.argtypes = [p, i, i, i, ctypes.c_double, i]
        _lib.orc_backproject_f64.argtypes = [p, i, i, i, p, p,
                                             ctypes.c_double, p, i]
        _lib.orc_forward_project.argtypes = [p, i, i, i, p, p,
                                             ctypes.c_float, ctypes.c_double,
                                             ctypes.c_double, p, i]
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def default_threads():
    return len(os.sched_getaffinity(0))


def trig_tables(angles):
    a = np.ascontiguousarray(angles, dtype=np.float32)
    c = np.empty_like(a)
    s = np.empty_like(a)
    lib().orc_trig_tables(_ptr(a), a.size, _ptr(c), _ptr(s))
    return c, s


def backproject(sino, angles, center, edge_pad=False, acc=fbp_np.ACC_F32SEQ,
                nthreads=None):
    """sino: (nrow, ncol) or (nrow, nslice, ncol) float32 -> float32 image(s)
    of shape (ncol, ncol) or (ncol, nslice, ncol)."""
    s = np.ascontiguousarray(sino, dtype=np.float32)
    two_d = s.ndim == 2
    if two_d:
        s = s[:, None, :]
    nrow, nslice, ncol = s.shape
    c, sn = trig_tables(angles)
    out = np.empty((ncol, nslice, ncol), dtype=np.float32)
    lib().orc_backproject(_ptr(s), nrow, nslice, ncol, _ptr(c), _ptr(sn),
                          float(np.float32(center)), int(bool(edge_pad)),
                          int(acc), _ptr(out), nthreads or default_threads())
    return out[:, 0, :] if two_d else out


def forward_project(image, angles, center, mask_radius=-1.0, scale=1.0,
                    nthreads=None):
    """Transpose of backproject's float64-sum form: image (ncol, ncol) or
    (ncol, nslice, ncol) float32 -> sinogram (nrow, ncol) or (nrow, nslice,
    ncol) float32; mask_radius < 0: no mask."""
    im = np.ascontiguousarray(image, dtype=np.float32)
    two_d = im.ndim == 2
    if two_d:
        im = im[:, None, :]
    ncol, nslice, _ = im.shape
    c, sn = trig_tables(angles)
    nrow = c.size
    out = np.empty((nrow, nslice, ncol), dtype=np.float32)
    lib().orc_forward_project(_ptr(im), nrow, nslice, ncol, _ptr(c), _ptr(sn),
                              float(np.float32(center)), float(mask_radius),
                              float(scale), _ptr(out),
                              nthreads or default_threads())
    return out[:, 0, :] if two_d else out


def backproject_f64(sino, angles, center, nthreads=None, round_center=True):
    """All-float64 back-projection: sino (nrow, ncol) or (nrow, nslice, ncol)
    float64; angles used as float32 (the reference's rounding), trig in
    float64.  Returns float64 (ncol, ncol) or (ncol, nslice, ncol)."""
    s = np.ascontiguousarray(sino, dtype=np.float64)
    two_d = s.ndim == 2
    if two_d:
        s = s[:, None, :]
    nrow, nslice, ncol = s.shape
    th = -np.asarray(angles, dtype=np.float32).astype(np.float64)
    c = np.ascontiguousarray(np.cos(th))
    sn = np.ascontiguousarray(np.sin(th))
    out = np.empty((ncol, nslice, ncol), dtype=np.float64)
    cen = float(np.float32(center)) if round_center else float(center)
    lib().orc_backproject_f64(_ptr(s), nrow, nslice, ncol, _ptr(c), _ptr(sn),
                              cen, _ptr(out),
                              nthreads or default_threads())
    return out[:, 0, :] if two_d else out


def fbp_f64(sinogram, center, angles, ratio=1.0, filter_name="hann",
            pad=None, apply_log=True, nthreads=None, round_center=True):
    """All-float64 FBP of float32 input (log, filter, BP, mask, scale in
    float64; only the reference's float32 rounding of angles and centre)."""
    x = np.asarray(sinogram)
    input_3d = x.ndim == 3
    if not input_3d:
        x = x[:, None, :]
    nrow, nslice, ncol = x.shape
    if pad is None:
        pad = min(int(0.15 * ncol), 150)
    xd = x.astype(np.float64)
    if apply_log:
        xd = -np.log(np.where(xd <= 0.0, 1.0, xd))
    g = fbp_np.filter_kernel(fbp_np.ramp_window_row(ncol + 2 * pad,
                                                    filter_name))
    filt = ramp_filter(xd, g, pad, nthreads=nthreads)
    out = backproject_f64(filt, angles, center, nthreads, round_center)
    if ratio is not None:
        if ratio == 0.0:
            ratio = min(center, ncol - center) / (0.5 * ncol)
        out = out * fbp_np.circle_mask(ncol, ratio)[:, None, :]
    out = out * np.pi / (nrow - 1)
    return out if input_3d else out[:, 0, :]


def ramp_filter(rows, g, pad, pad_mode="edge", apply_log=False,
                nthreads=None):
    """rows: (..., ncol) float32/float64; g: real kernel of length
    ncol + 2*pad.  Returns float64 (..., ncol)."""
    r = np.asarray(rows)
    in_f64 = r.dtype == np.float64
    r = np.ascontiguousarray(r, dtype=np.float64 if in_f64 else np.float32)
    if pad_mode in fbp_np.VALUE_PAD_MODES and pad > 0:
        # the pad values depend on the logged row: log (float64, rounded to
        # the row's type), np.pad, then the same filter with nothing to pad
        if apply_log:
            r = (-np.log(r.astype(np.float64))).astype(r.dtype)
        ncol = r.shape[-1]
        full = ramp_filter(fbp_np.pad_rows(r, pad, pad_mode), g, 0, "edge",
                           False, nthreads)
        return np.ascontiguousarray(full[..., pad:pad + ncol])
    ncol = r.shape[-1]
    nrows = r.size // ncol
    g = np.ascontiguousarray(g, dtype=np.float64)
    assert g.size == ncol + 2 * pad
    out = np.empty(r.shape, dtype=np.float64)
    lib().orc_ramp_filter(_ptr(r), int(in_f64), nrows, ncol, int(pad),
                          PAD_MODES[pad_mode], int(bool(apply_log)), _ptr(g),
                          _ptr(out), nthreads or default_threads())
    return out


def fbp(sinogram, center, angles=None, ratio=1.0, ramp_row=None,
        filter_name="hann", pad=None, pad_mode="edge", apply_log=True,
        gpu=True, nthreads=None):
    """Same contract as fbp_np.fbp, evaluated by the C port."""
    sino = np.array(sinogram, copy=True)
    input_3d = sino.ndim == 3
    if not input_3d:
        sino = sino[:, None, :]
    nrow, nslice, ncol = sino.shape
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, nrow))
    if apply_log:
        sino[sino <= 0.0] = 1.0
    if pad is None:
        pad = min(int(0.15 * ncol), 150)
    width = ncol + 2 * pad
    if ramp_row is None or np.shape(ramp_row)[-1] != width:
        ramp_row = fbp_np.ramp_window_row(width, filter_name)
    g = fbp_np.filter_kernel(ramp_row)
    filt = ramp_filter(sino, g, pad, pad_mode, apply_log,
                       nthreads).astype(np.float32)
    acc = fbp_np.ACC_F64 if (gpu and nslice == 1) else fbp_np.ACC_F32SEQ
    out = backproject(filt, angles, center, acc=acc, nthreads=nthreads)
    use_mask = ratio is not None
    if use_mask and ratio == 0.0:
        ratio = min(center, ncol - center) / (0.5 * ncol)
    lib().orc_finalize(_ptr(out), ncol, nslice, nrow,
                       float(ratio if use_mask else 0.0), int(use_mask))
    return out if input_3d else out[:, 0, :]
