"""Oracle for vertical-slice reconstruction -- TEST INFRASTRUCTURE ONLY.

Restates algotom/rec/vertrec.py:53-335 (the four vertical back-projectors,
with numba's typing), :338-448 (the points of a slice) and :450-706
(vertical_reconstruction's pipeline: crop, flat field, log, chunked
back-projection, ramp filter before or after, mask, scale) in NumPy, one
projection at a time, vectorised over rows and points.

Typing (from the kernels' argument types):
  "gpu"  float32 angles, points, centre, trig (cos of the float32 angle,
         rounded to float32), position r = x c + y s and f = r + centre in
         float32; float64 sum over a chunk of projections;
  "cpu"  float64 trig and positions; the chunk image is float32, stored
         after every projection.
Chunks are added in float32 (the reference's ver_slice += chunk result).
The log is taken in float64 and rounded, the oracle's convention for the
FBP path.

Pinned in tests/test_oracle_vertrec.py: the points by their geometry (unit
spacing, distance of the line from the centre, direction), the
back-projector by scalar pure-Python loops and by the standard
back-projector (an alpha = 0 slice of an odd-width image is one row of the
horizontal slices), the pipeline by the reference test's thresholds
(tests/test_rec/test_vertrec.py:146-320).
"""

import numpy as np

from oracle import fbp_np


def line_points(slice_index, alpha, width):
    """Points of the line at slice_index, angle alpha (degree), float64
    (vertrec.py:338-384, whose radian-vs-45/90 comparisons are always
    true for alpha in [0, 180])."""
    mid = np.ceil(0.5 * (width - 1))
    t = np.arange(-mid, width - mid)
    a = np.deg2rad(alpha)
    off = slice_index - mid
    if alpha <= 45 or alpha >= 135:
        k = np.tan(a)
        b = -off / np.cos(a)
        x = -k * b / (1 + k ** 2) + t * np.cos(a)
        return x, k * x + b
    k = np.cos(a) / np.sin(a)
    b = off / np.sin(a)
    y = -k * b / (1 + k ** 2) + t * np.sin(a)
    return k * y + b, y


def vertical_backproject(proj, angles, xs, ys, center, edge_pad=False,
                         mode="gpu", chunk=None):
    """proj (nproj, height, width) float32; xs, ys (nslice, width) ->
    (nslice, height, width) float32, chunks of `chunk` projections."""
    proj = np.asarray(proj, dtype=np.float32)
    nproj, height, width = proj.shape
    xs = np.atleast_2d(np.asarray(xs, dtype=np.float64))
    ys = np.atleast_2d(np.asarray(ys, dtype=np.float64))
    chunk = nproj if chunk is None else chunk
    w1 = width - 1
    out = np.zeros((xs.shape[0], height, width), dtype=np.float32)
    for s in range(xs.shape[0]):
        total = np.zeros((height, width), dtype=np.float32)
        acc64 = np.zeros((height, width))
        acc32 = np.zeros((height, width), dtype=np.float32)
        for n in range(nproj):
            if mode == "gpu":
                th = np.float32(angles[n])
                c = np.float32(np.cos(np.float64(th)))
                sn = np.float32(np.sin(np.float64(th)))
                x32 = xs[s].astype(np.float32)
                y32 = ys[s].astype(np.float32)
                f = (x32 * c + y32 * sn) + np.float32(center)
            else:
                th = np.float64(angles[n])
                f = (xs[s] * np.cos(th) + ys[s] * np.sin(th)) + \
                    np.float64(center)
            inr = (f >= 0) & (f.astype(np.float64) <= w1)
            d = np.floor(f).astype(np.int64)
            u = np.ceil(f).astype(np.int64)
            frac = f.astype(np.float64) - d
            dc, uc = np.clip(d, 0, w1), np.clip(u, 0, w1)
            row = proj[n]
            yd, yu = row[:, dc], row[:, uc]
            val = yd.astype(np.float64) + (yu - yd).astype(np.float64) * frac
            edge = np.where(f < 0, row[:, :1], row[:, w1:]).astype(np.float64)
            add = np.where(inr, val, edge if edge_pad else 0.0)
            if mode == "gpu":
                acc64 = acc64 + add
            else:
                acc32 = (acc32.astype(np.float64) + add).astype(np.float32)
            if (n + 1) % chunk == 0 or n + 1 == nproj:
                if mode == "gpu":
                    total = total + acc64.astype(np.float32)
                    acc64[:] = 0.0
                else:
                    total = total + acc32
                    acc32[:] = 0.0
        out[s] = total
    return out


def vertical_backproject_loops(proj, angles, x, y, center, edge_pad, mode):
    """Scalar restatement of one slice, one chunk (tiny inputs only)."""
    nproj, height, width = proj.shape
    w1 = width - 1
    out = np.zeros((height, width), dtype=np.float32)
    for i in range(height):
        for j in range(width):
            acc = 0.0 if mode == "gpu" else np.float32(0.0)
            for n in range(nproj):
                if mode == "gpu":
                    th = np.float32(angles[n])
                    c = np.float32(np.cos(float(th)))
                    s = np.float32(np.sin(float(th)))
                    f = np.float32(np.float32(np.float32(x[j]) * c)
                                   + np.float32(np.float32(y[j]) * s))
                    f = np.float32(f + np.float32(center))
                else:
                    f = float(x[j]) * np.cos(float(angles[n])) + \
                        float(y[j]) * np.sin(float(angles[n])) + center
                row = proj[n, i]
                if 0 <= f <= w1:
                    d, u = int(np.floor(f)), int(np.ceil(f))
                    if u != d:
                        diff = np.float32(row[u] - row[d])
                        v = float(row[d]) + float(diff) * (float(f) - d)
                    else:
                        v = float(row[d])
                elif edge_pad:
                    v = float(row[0] if f < 0 else row[w1])
                else:
                    continue
                if mode == "gpu":
                    acc += v
                else:
                    acc = np.float32(float(acc) + v)
            out[i, j] = np.float32(acc)
    return out


def after_filter(ver, ramp_row_abs, pad, pad_mode="edge"):
    """The ramp filter along the rows of a back-projected vertical slice
    (vertrec.py:621-627): the shifted spectrum times |R|, inverse FFT, real
    part -- without the final ifftshift of apply_ramp_filter (|R| carries
    no centring phase).  complex128, float64 out."""
    width = ver.shape[-1]
    p = np.pad(np.asarray(ver, np.float64), ((0, 0), (pad, pad)),
               mode=pad_mode)
    spec = np.fft.fftshift(np.fft.fft(p), axes=1) * ramp_row_abs
    return np.real(np.fft.ifft(np.fft.ifftshift(spec, axes=1)))[
        :, pad:pad + width]


def vertical_reconstruction(proj, slice_index, center, alpha=0.0,
                            flat_field=None, dark_field=None, angles=None,
                            crop=(0, 0, 0, 0), proj_start=0, proj_stop=-1,
                            chunk_size=30, ramp_filter="after",
                            filter_name="hann", pad=None, pad_mode="edge",
                            apply_log=True, gpu=True, masking=False,
                            points=None):
    """vertical_reconstruction (vertrec.py:450-695).  `points` = (xs, ys)
    replaces the single line for the multiple-slice forms."""
    num0, height0, width0 = proj.shape
    if proj_stop == -1:
        proj_stop = num0
    num_proj = proj_stop - proj_start
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, num_proj))
    top, bottom, left, right = crop
    sub = np.asarray(proj[proj_start:proj_stop, top:height0 - bottom,
                          left:width0 - right])
    nproj, height, width = sub.shape
    img = sub.astype(np.float32)
    if flat_field is not None or dark_field is not None:
        flat = np.ones((height0, width0), np.float32) if flat_field is None \
            else np.asarray(flat_field)
        dark = np.zeros((height0, width0), np.float32) if dark_field is None \
            else np.asarray(dark_field)
        flat = flat[top:height0 - bottom, left:width0 - right]
        dark = dark[top:height0 - bottom, left:width0 - right]
        fd = flat - dark
        fd[fd == 0.0] = 1.0
        img = (img - np.float32(dark)) / np.float32(fd)
    if apply_log:
        img = img.copy()
        img[img <= 0.0] = 1.0
        img = (-np.log(img.astype(np.float64))).astype(np.float32)
    if chunk_size is None:
        chunk_size = min(30, num_proj // 4 + 1)
    if chunk_size > num_proj or chunk_size < 1:
        chunk_size = num_proj
    if pad is None:
        pad = min(int(0.15 * width), 150)
    if points is None:
        xs, ys = line_points(slice_index, alpha, width)
        xs, ys = xs[None], ys[None]
        single = True
    else:
        xs, ys = points
        single = False
    edge_pad = ramp_filter != "before"
    if ramp_filter == "before":
        rrow = fbp_np.ramp_window_row(width + 2 * pad, filter_name)
        img = np.stack([fbp_np.ramp_filter(p, rrow, None, pad, pad_mode)
                        for p in img]).astype(np.float32)
    ver = vertical_backproject(img, angles, xs, ys, center, edge_pad,
                               "gpu" if gpu else "cpu", chunk_size)
    out = ver
    if ramp_filter == "after":
        rrow = np.abs(fbp_np.ramp_window_row(width + 2 * pad, filter_name))
        out = np.stack([after_filter(v, rrow, pad, pad_mode) for v in ver])
        if not single:
            out = out.astype(np.float32)
    if masking:
        rad = int(min(center, width - center))
        lo = (width - 2 * rad) // 2
        m = np.zeros(width, np.float32)
        m[lo:lo + 2 * rad] = 1.0
        out = out * m
    if out.dtype == np.float32:
        out = out * np.float32(np.pi) / np.float32(num_proj - 1)
    else:
        out = out * np.pi / (num_proj - 1)
    return out[0] if single else out


def autocorrelation_metric(image):
    """vertrec.py:1255-1261: mean over rows of max |ifft(|fft(row)|^2)|,
    complex128 (numpy 1.x semantics for the float32 image)."""
    rows = np.fft.fft(np.float32(image).astype(np.float64))
    ac = np.abs(np.fft.ifft(np.abs(rows) ** 2))
    return np.mean(np.max(ac, axis=1))


def find_center_vertical(proj, slice_index, start, stop, step=1.0,
                         metric="entropy", alpha=0.0, angles=None,
                         chunk_size=30, ramp_filter="after", sigma=3,
                         apply_log=True, gpu=True, masking=True,
                         invert_metric=False, metric_function=None, **kw):
    """find_center_vertical_slice (vertrec.py:1264-1428) with
    return_metric=True: (list_center, list_metric)."""
    import scipy.ndimage as ndi
    from oracle import center_np
    if ramp_filter is None:
        ramp_filter = "after"
    centers = np.arange(start, stop + step, step)
    num_proj, height, width = proj.shape
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, num_proj))
    p = np.asarray(proj, dtype=np.float32)
    if sigma > 0:
        p = ndi.gaussian_filter(p, (0, sigma, sigma))
    if apply_log:
        p = p.copy()
        p[p <= 0.0] = 1.0
        p = (-np.log(p.astype(np.float64))).astype(np.float32)
    mask = None
    if masking:
        cmin = np.min(centers)
        rad = int(min(cmin, width - cmin))
        lo = (width - 2 * rad) // 2
        mask = np.zeros(width, np.float32)
        mask[lo:lo + 2 * rad] = 1.0
    pad = min(int(0.15 * width), 150)
    kwr = dict(alpha=alpha, angles=angles, chunk_size=chunk_size,
               ramp_filter=ramp_filter, pad=pad, pad_mode="edge",
               filter_name="hann", apply_log=False, gpu=gpu, masking=False)
    nmin, nmax = 0.0, 1.0
    if metric_function is None and metric == "entropy":
        ref = vertical_reconstruction(p, slice_index, np.mean(centers), **kwr)
        nmin, nmax = np.min(ref), np.max(ref)
    values = np.zeros_like(centers)
    for i, c in enumerate(centers):
        rec = vertical_reconstruction(p, slice_index, c, **kwr)
        if mask is not None:
            rec = rec * mask
        if metric_function is not None:
            values[i] = metric_function(rec, **kw)
        elif metric == "entropy":
            values[i] = center_np.slice_metric(rec, "entropy", nmin, nmax)
        elif metric == "autocorrelation":
            values[i] = autocorrelation_metric(rec)
        else:
            values[i] = center_np.edge_sharpness(rec)
    if invert_metric:
        values = np.max(values) - values
    return centers, values
