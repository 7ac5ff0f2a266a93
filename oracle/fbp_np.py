"""CPU oracle for the FBP path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only to check results against.  The product path
(``algotom_amd``) never imports it and fails loudly when its HIP library is
missing.

This is a NumPy restatement of algotom's filtered back-projection
(algotom 1.6.0, ``algotom/rec/reconstruction.py``), written from the
reference's documented behaviour:

* ``smoothing_window``       follows reconstruction.py:50-83 (scipy.signal
  window definitions, symmetric form), restated from the closed forms;
* ``ramp_window_row``        follows reconstruction.py:86-115 (discrete
  Ram-Lak kernel, FFT, fftshift, times the window);
* ``ramp_filter``            follows reconstruction.py:118-153 (pad, shifted
  spectral multiply, real part, crop) -- evaluated as the equivalent
  length-W circular convolution with ``filter_kernel``;
* ``backproject``            follows reconstruction.py:156-190 (2D GPU kernel,
  FP64 register accumulation), :229-264 (chunk kernel) and :307-359 (CPU
  kernel), both FP32 read-modify-write per angle;
* ``circle_mask``            follows algotom/util/utility.py:160-182;
* ``fbp``                    follows reconstruction.py:362-493.

Numba typing of the reference kernels is emulated exactly: float32 storage,
float32 angles/centre, float32 cos/sin (``float(cos(double(x)))``), float64
detector position ``f``, float64 interpolated value, and either a float64
accumulator (2D GPU kernel) or float32 storage after every angle (CPU and
chunk kernels).

Pinning (each check lives in ``tests/test_oracle_*.py``):
windows -> closed forms and scipy.signal.windows; ramp -> closed-form Ram-Lak
values and DC/Nyquist identities; filter -> O(W^2) brute-force circular
convolution, linearity, the reference's own fftshift chain; back-projection ->
pure-Python loops on tiny inputs, constant-sinogram closed form, single-angle
shift, dense-matrix adjointness, integer centre shift, 180-degree mirror
symmetry; end to end -> the reference's test case thresholds
(tests/test_rec/test_reconstruction.py:88-144).  Parity unpinned at the ULP
level: the float32 log of numpy (SIMD, up to 4 ulp from correctly rounded) and
libm cosf/sinf are modelled as correctly rounded.
"""

import math
import numpy as np

ACC_F64 = 0      # reconstruction.py:168,182 -- ``num`` float64 register
ACC_F32SEQ = 1   # reconstruction.py:256 and :351 -- float32 += per angle

_PAD_MODES = ("edge", "constant", "reflect", "symmetric", "wrap")


# --------------------------------------------------------------------------
# windows and ramp (reconstruction.py:50-115)
# --------------------------------------------------------------------------
def smoothing_window(filter_name, width):
    """Symmetric smoothing windows from their closed forms.

    Follows reconstruction.py:50-83, which calls scipy.signal.windows.<name>
    (sym=True).  Unknown names and None give ones (reconstruction.py:81-82).
    """
    m = int(width)
    if filter_name not in ("hann", "bartlett", "blackman", "hamming",
                           "nuttall", "parzen", "triang"):
        return np.ones(m)
    if m <= 1:
        return np.ones(m)
    n = np.arange(m, dtype=np.float64)
    phase = 2.0 * np.pi * n / (m - 1)
    if filter_name == "hann":
        return 0.5 - 0.5 * np.cos(phase)
    if filter_name == "hamming":
        return 0.54 - 0.46 * np.cos(phase)
    if filter_name == "blackman":
        return 0.42 - 0.5 * np.cos(phase) + 0.08 * np.cos(2.0 * phase)
    if filter_name == "nuttall":
        a = (0.3635819, 0.4891775, 0.1365995, 0.0106411)
        return (a[0] - a[1] * np.cos(phase) + a[2] * np.cos(2.0 * phase)
                - a[3] * np.cos(3.0 * phase))
    if filter_name == "bartlett":
        half = 0.5 * (m - 1)
        return 1.0 - np.abs(n - half) / half
    if filter_name == "triang":
        k = np.arange(1, (m + 1) // 2 + 1, dtype=np.float64)
        if m % 2 == 0:
            w = (2.0 * k - 1.0) / m
            return np.concatenate([w, w[::-1]])
        w = 2.0 * k / (m + 1.0)
        return np.concatenate([w, w[-2::-1]])
    # parzen: piecewise cubic B-spline window
    t = np.abs(n - 0.5 * (m - 1)) / (0.5 * m)
    inner = 1.0 - 6.0 * t ** 2 + 6.0 * t ** 3
    outer = 2.0 * (1.0 - t) ** 3
    return np.where(np.abs(n - 0.5 * (m - 1)) <= 0.25 * (m - 1), inner, outer)


def ramlak_kernel(width):
    """Discrete Ram-Lak impulse response on n = k - ceil((W-1)/2).

    reconstruction.py:106-111: h(0) = 1/4, h(even) = 0, h(odd) = -1/(n pi)^2.
    """
    n = np.arange(width, dtype=np.float64) - math.ceil((width - 1.0) / 2)
    h = np.zeros(width)
    h[n == 0] = 0.25
    odd = (np.abs(n) % 2) == 1
    h[odd] = -1.0 / (n[odd] * np.pi) ** 2
    return h


def ramp_window_row(width, filter_name=None):
    """One row of make_2d_ramp_window (reconstruction.py:86-115)."""
    spec = np.fft.fftshift(np.fft.fft(ramlak_kernel(width)))
    return spec * smoothing_window(filter_name, width)


def filter_kernel(ramp_row):
    """Real circular-convolution kernel g equivalent to the spectral filter.

    The reference (reconstruction.py:150-152) computes
    real(ifftshift(ifft(ifftshift(fftshift(fft(p)) * R)))).  Since
    ifftshift(fftshift(P) * R) = P * ifftshift(R), this is
    real(p (*) g) with g = ifftshift(ifft(ifftshift(R))), and for real p the
    real part only involves Re(g).
    """
    r = np.asarray(ramp_row)
    g = np.fft.ifftshift(np.fft.ifft(np.fft.ifftshift(r)))
    return np.real(g)


def pad_index(width_in, pad, width_out, mode):
    """Source index of each padded sample for numpy.pad's index-map modes."""
    i = np.arange(width_out) - pad
    n = width_in
    if mode == "edge":
        return np.clip(i, 0, n - 1)
    if mode == "wrap":
        return np.mod(i, n)
    if mode == "reflect":
        if n == 1:
            return np.zeros_like(i)
        p = 2 * (n - 1)
        j = np.mod(i, p)
        return np.where(j < n, j, p - j)
    if mode == "symmetric":
        p = 2 * n
        j = np.mod(i, p)
        return np.where(j < n, j, p - 1 - j)
    raise ValueError(mode)


# numpy.pad's value modes (reconstruction.py:147 passes any mode to np.pad):
# with their defaults the pad values are a statistic of the whole row or a
# linspace ramp, which np.pad itself evaluates here
VALUE_PAD_MODES = ("linear_ramp", "maximum", "mean", "median", "minimum")


def pad_rows(rows, pad, mode):
    """numpy.pad((0,0),(pad,pad)) restated with explicit index maps (the
    value modes through np.pad itself, in the rows' own type)."""
    rows = np.asarray(rows)
    n = rows.shape[-1]
    w = n + 2 * pad
    if mode in VALUE_PAD_MODES:
        return np.pad(rows, [(0, 0)] * (rows.ndim - 1) + [(pad, pad)],
                      mode=mode)
    if mode == "constant":
        out = np.zeros(rows.shape[:-1] + (w,), dtype=rows.dtype)
        out[..., pad:pad + n] = rows
        return out
    return rows[..., pad_index(n, pad, w, mode)]


def ramp_filter(sino, ramp_row=None, filter_name=None, pad=None,
                pad_mode="edge"):
    """Row-wise windowed ramp filter, float64 (reconstruction.py:118-153).

    ``sino`` is (..., ncol); returns float64 (..., ncol).
    """
    sino = np.asarray(sino)
    ncol = sino.shape[-1]
    if pad is None:
        pad = min(int(0.15 * ncol), 150)
    width = ncol + 2 * pad
    if ramp_row is None or np.shape(ramp_row)[-1] != width:
        ramp_row = ramp_window_row(width, filter_name)
    g = filter_kernel(ramp_row)
    p = pad_rows(sino, pad, pad_mode).astype(np.float64)
    q = np.fft.ifft(np.fft.fft(p, axis=-1) * np.fft.fft(g), axis=-1)
    return np.ascontiguousarray(np.real(q)[..., pad:pad + ncol])


def ramp_filter_spectral(sino, ramp_row, pad, pad_mode="edge"):
    """The reference's own fftshift chain (reconstruction.py:147-153), FP64.

    Kept beside ``ramp_filter`` so the circular-convolution form is checked
    against the literal spectral form in the tests.
    """
    p = pad_rows(np.asarray(sino), pad, pad_mode).astype(np.float64)
    ncol = p.shape[-1] - 2 * pad
    spec = np.fft.fftshift(np.fft.fft(p, axis=-1), axes=-1) * ramp_row
    q = np.fft.ifftshift(np.fft.ifft(np.fft.ifftshift(spec, axes=-1),
                                     axis=-1), axes=-1)
    return np.real(q)[..., pad:pad + ncol]


# --------------------------------------------------------------------------
# back-projection (reconstruction.py:156-190, 229-264, 307-359)
# --------------------------------------------------------------------------
def trig_tables(angles):
    """float32 cos/sin of the float32 negated angles, as numba types them.

    reconstruction.py:169-171 / :334-336: theta = -angles[i] (float32), then
    math.cos/sin on float32 give float32 (modelled as correctly rounded).
    """
    th = -np.asarray(angles, dtype=np.float32).astype(np.float64)
    return (np.cos(th).astype(np.float32), np.sin(th).astype(np.float32))


def backproject(sino, angles, center, edge_pad=False, acc=ACC_F32SEQ):
    """Numba-typed back-projection of one (nrow, ncol) float32 sinogram.

    Returns a (ncol, ncol) float32 image indexed [y, x].
    """
    sino = np.asarray(sino, dtype=np.float32)
    nrow, ncol = sino.shape
    cos_t, sin_t = trig_tables(angles)
    center = np.float64(np.float32(center))
    w1 = ncol - 1
    icenter = 0.5 * w1
    xc = (np.arange(ncol, dtype=np.float64) - icenter)[None, :]
    yc = (np.arange(ncol, dtype=np.float64) - icenter)[:, None]
    if acc == ACC_F64:
        accum = np.zeros((ncol, ncol), dtype=np.float64)
    else:
        accum = np.zeros((ncol, ncol), dtype=np.float32)
    for i in range(nrow):
        c = np.float64(cos_t[i])
        s = np.float64(sin_t[i])
        f = (xc * c + yc * s) + center
        inside = (f >= 0.0) & (f <= w1)
        fd = np.floor(np.where(inside, f, 0.0))
        d = fd.astype(np.int64)
        u = np.minimum(d + 1, w1)
        row = sino[i]
        yd = row[d]
        yu = row[u]
        delta = (yu - yd).astype(np.float64)            # float32 subtraction
        val = yd.astype(np.float64) + delta * (f - fd)  # float64 mul, add
        val = np.where(f == fd, yd.astype(np.float64), val)
        if edge_pad:
            val = np.where(inside, val,
                           np.where(f < 0, np.float64(row[0]),
                                    np.float64(row[w1])))
            take = np.ones_like(inside)
        else:
            take = inside
        if acc == ACC_F64:
            accum = np.where(take, accum + val, accum)
        else:
            accum = np.where(take, (accum.astype(np.float64) + val)
                             .astype(np.float32), accum)
    return accum.astype(np.float32)


def backproject_loops(sino, angles, center, edge_pad=False, acc=ACC_F32SEQ):
    """Pure-Python scalar loops of the same arithmetic, for tiny inputs."""
    sino = np.asarray(sino, dtype=np.float32)
    nrow, ncol = sino.shape
    cos_t, sin_t = trig_tables(angles)
    center = float(np.float32(center))
    w1 = ncol - 1
    icenter = 0.5 * w1
    out = np.zeros((ncol, ncol), dtype=np.float32)
    for y in range(ncol):
        for x in range(ncol):
            total = 0.0 if acc == ACC_F64 else np.float32(0.0)
            for i in range(nrow):
                f = ((x - icenter) * float(cos_t[i])
                     + (y - icenter) * float(sin_t[i])) + center
                if 0.0 <= f <= w1:
                    d = math.floor(f)
                    u = math.ceil(f)
                    if u != d:
                        yd = sino[i, d]
                        val = float(yd) + float(sino[i, u] - yd) * (f - d)
                    else:
                        val = float(sino[i, d])
                elif edge_pad:
                    val = float(sino[i, 0] if f < 0 else sino[i, w1])
                else:
                    continue
                if acc == ACC_F64:
                    total += val
                else:
                    total = np.float32(float(total) + val)
            out[y, x] = np.float32(total)
    return out


def backproject_f64(sino, angles, center):
    """All-FP64 back-projection (no float32 rounding of trig or data)."""
    sino = np.asarray(sino, dtype=np.float64)
    nrow, ncol = sino.shape
    w1 = ncol - 1
    icenter = 0.5 * w1
    xc = (np.arange(ncol, dtype=np.float64) - icenter)[None, :]
    yc = (np.arange(ncol, dtype=np.float64) - icenter)[:, None]
    out = np.zeros((ncol, ncol))
    for i in range(nrow):
        th = -float(angles[i])
        f = xc * math.cos(th) + yc * math.sin(th) + float(center)
        inside = (f >= 0.0) & (f <= w1)
        fd = np.floor(np.where(inside, f, 0.0))
        d = fd.astype(np.int64)
        u = np.minimum(d + 1, w1)
        val = sino[i][d] + (sino[i][u] - sino[i][d]) * (f - fd)
        out += np.where(inside, val, 0.0)
    return out


def forward_project_matrix(ncol, angles, center, float32_trig=True):
    """Dense (nrow*ncol, ncol*ncol) matrix of the back-projector's transpose.

    Built pixel by pixel from the interpolation weights so that
    <BP s, x> = <s, A x> holds; used for the adjointness checks.  With
    ``float32_trig`` the float32 cos/sin of ``backproject`` are used, else the
    float64 ones of ``backproject_f64``.
    """
    if float32_trig:
        cos_t, sin_t = trig_tables(angles)
    else:
        th = -np.asarray(angles, dtype=np.float64)
        cos_t, sin_t = np.cos(th), np.sin(th)
    center = np.float64(np.float32(center))
    nrow = len(angles)
    w1 = ncol - 1
    icenter = 0.5 * w1
    a = np.zeros((nrow * ncol, ncol * ncol))
    for i in range(nrow):
        for y in range(ncol):
            for x in range(ncol):
                f = ((x - icenter) * np.float64(cos_t[i])
                     + (y - icenter) * np.float64(sin_t[i])) + center
                if 0.0 <= f <= w1:
                    d = int(math.floor(f))
                    t = f - d
                    a[i * ncol + d, y * ncol + x] += 1.0 - t
                    if t > 0.0:
                        a[i * ncol + d + 1, y * ncol + x] += t
    return a


# --------------------------------------------------------------------------
# mask and orchestration (utility.py:160-182, reconstruction.py:362-493)
# --------------------------------------------------------------------------
def circle_mask(width, ratio):
    """utility.py:160-182: ones where x^2 + y^2 <= (ratio * (width//2))^2."""
    c = width // 2
    radius = ratio * c
    y, x = np.ogrid[-c:width - c, -c:width - c]
    return (x * x + y * y <= radius * radius).astype(np.float32)


def fbp(sinogram, center, angles=None, ratio=1.0, ramp_row=None,
        filter_name="hann", pad=None, pad_mode="edge", apply_log=True,
        gpu=True):
    """The reference's fbp_reconstruction pipeline (reconstruction.py:362-493).

    ``gpu`` selects which reference kernel's accumulation is emulated: the 2D
    GPU kernel accumulates in float64 (reconstruction.py:168), the chunk
    kernel and the CPU kernel store float32 after each angle (:256, :351).
    Returns float32, (ncol, ncol) for 2D input or (ncol, nslice, ncol) for 3D.
    """
    sino = np.array(sinogram, copy=True)
    input_3d = sino.ndim == 3
    if input_3d:
        nrow, nslice, ncol = sino.shape
    else:
        nrow, ncol = sino.shape
        nslice = 1
        sino = sino[:, None, :]
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, nrow))
    if apply_log:
        sino[sino <= 0.0] = 1.0
        sino = (-np.log(sino.astype(np.float64))).astype(sino.dtype)
    if pad is None:
        pad = min(int(0.15 * ncol), 150)
    if ratio is not None and ratio == 0.0:
        ratio = min(center, ncol - center) / (0.5 * ncol)
    filt = ramp_filter(np.moveaxis(sino, 1, 0), ramp_row, filter_name, pad,
                       pad_mode).astype(np.float32)       # (S, N, C)
    acc = ACC_F64 if (gpu and nslice == 1) else ACC_F32SEQ
    out = np.zeros((ncol, nslice, ncol), dtype=np.float32)
    for j in range(nslice):
        out[:, j, :] = backproject(filt[j], angles, center, acc=acc)
    if ratio is not None:
        out = out * circle_mask(ncol, ratio)[:, None, :]
    out = out * np.float32(np.pi) / np.float32(nrow - 1)
    if not input_3d:
        return out[:, 0, :]
    return out
