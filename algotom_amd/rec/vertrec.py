"""Vertical-slice reconstruction on AMD Instinct MI355X -- drop-in for
algotom's ``algotom.rec.vertrec`` (algotom/rec/vertrec.py).

A vertical slice is reconstructed directly from the projections
(num_projection, height, width) along a line of points in the horizontal
plane, for every detector row, without reconstructing and reslicing the
volume.  Flat-field correction and log (ath_flat_field_log), the optional
ramp filter before (every projection row) or after (every slice row)
back-projection (ath_ramp_filter), the vertical back-projection
(ath_vertical_back_projection) and the mask/scale (ath_vertical_finish) all
run on the GPU; nothing falls back to the CPU.

* ``gpu`` selects which reference kernel's arithmetic is reproduced:
  ``True`` = __vertical_back_projection_gpu(_chunk)_kernel (float32
  positions, float64 sums per projection chunk), ``False`` =
  vertical_back_projection_cpu(_chunk) (float64 positions, float32 chunk
  images).  Both run on the GPU; projection chunks are summed in float32 as
  the reference's chunk loop does, so ``chunk_size`` keeps its meaning.
* ``block``, ``ncore``, ``prefer`` are accepted and ignored;
  ``show_progress`` prints one line per pass.
* Projections: numpy arrays (float32 or uint16; other types are converted
  to float32) or torch tensors on the GPU.  Flat and dark fields are used as
  float32.
* Centre search: ``find_center_vertical_slice`` (metrics on the GPU) and
  ``find_center_visual_vertical_slices`` (one TIFF per centre, written with
  ``algotom_amd.io.loadersaver``).
"""

import sys
import time
import warnings

import numpy as np

from algotom_amd import _host
from algotom_amd import _native as nat
from algotom_amd.rec import reconstruction as rec

_LOG_WARNING = ("!!! Applying logarithm is enabled but there are values "
                "<= 0.0 in the data !!!")
_GROUP_BYTES = 2 << 30


# --------------------------------------------------------------------------
# points of a slice (vertrec.py:338-448), host metadata
# --------------------------------------------------------------------------
def _get_points_single_line(slice_index, alpha, width):
    """x and y coordinates (float64, relative to the image centre) of the
    width points of the line at slice_index, angle alpha (degree)
    (vertrec.py:338-384).  The points are one pixel apart; the line is at
    distance |slice_index - centre| from the centre, at angle alpha."""
    if alpha < 0.0 or alpha > 180.0:
        raise ValueError("Angle is out of range [0, 180] (degree)")
    if slice_index < 0 or slice_index > width - 1:
        raise ValueError(f"Slice index is out of range [0, {width - 1}]")
    mid = np.ceil(0.5 * (width - 1))
    steps = np.arange(-mid, width - mid)
    theta = np.deg2rad(alpha)
    offset = slice_index - mid
    # the reference compares the angle in radians with 45 and 90 (:367,
    # :376), which holds for every alpha in [0, 180]; kept as it is
    if alpha <= 45 or alpha >= 135:
        slope = np.tan(theta)
        icept = (-offset if theta <= 45 else offset) / np.cos(theta)
        foot = - slope * icept / (1 + slope ** 2)
        xs = foot + steps * np.cos(theta)
        ys = slope * xs + icept
    else:
        slope = np.cos(theta) / np.sin(theta)
        icept = (offset if theta <= 90 else -offset) / np.sin(theta)
        foot = - slope * icept / (1 + slope ** 2)
        ys = foot + steps * np.sin(theta)
        xs = slope * ys + icept
    return xs, ys


def _get_points_multiple_lines(start_index, stop_index, alpha, width,
                               step_index=1):
    """float32 (nslice, width) coordinates of parallel lines
    (vertrec.py:387-418)."""
    pts = [_get_points_single_line(i, alpha, width)
           for i in np.arange(start_index, stop_index + 1, step_index)]
    return (np.float32([p[0] for p in pts]), np.float32([p[1] for p in pts]))


def _get_points_multiple_lines_different_angles(list_index, list_alpha,
                                                width):
    """float32 (nslice, width) coordinates of lines at their own angles
    (vertrec.py:421-447)."""
    pts = [_get_points_single_line(i, list_alpha[k], width)
           for k, i in enumerate(list_index)]
    return (np.float32([p[0] for p in pts]), np.float32([p[1] for p in pts]))


# --------------------------------------------------------------------------
# device helpers
# --------------------------------------------------------------------------
def _points_device(xs, ys, dev):
    import torch
    x = np.ascontiguousarray(np.atleast_2d(np.asarray(xs, np.float64)))
    y = np.ascontiguousarray(np.atleast_2d(np.asarray(ys, np.float64)))
    if x.shape != y.shape:
        raise ValueError("x and y coordinates differ in shape")
    return torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)


def _angles64(angles, dev):
    import torch
    if rec._is_tensor(angles):
        return angles.detach().to(dev, torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(
        np.asarray(angles, dtype=np.float64))).to(dev)


def _vbp(proj_t, ang_t, x_t, y_t, center, edge_pad, mode, chunk, out,
         accumulate, centers_t=None):
    nproj, height, width = proj_t.shape
    if x_t.shape[1] != width:
        raise ValueError("the lines need one point per projection column")
    nat.check(nat.load().ath_vertical_back_projection(
        nat.ptr(proj_t), nproj, height, width, nat.ptr(ang_t), nat.ptr(x_t),
        nat.ptr(y_t), x_t.shape[0], float(center),
        nat.ptr(centers_t) if centers_t is not None else None,
        int(bool(edge_pad)), int(mode), int(chunk), int(bool(accumulate)),
        nat.ptr(out), nat.stream_ptr(proj_t.device)),
        "ath_vertical_back_projection")
    return out


def _bp_public(projections, angles, xs, ys, center, edge_pad, mode, single):
    import torch
    dev = nat.device()
    as_numpy = not rec._is_tensor(projections)
    p = rec._to_device(projections, dev, allow_f64=False)
    nproj, height, width = p.shape
    if len(angles) != nproj:
        raise ValueError("!!! Number of angles is not the same as the "
                         "number of projections !!!")
    x_t, y_t = _points_device(xs, ys, dev)
    out = torch.empty((x_t.shape[0], height, width), dtype=torch.float32,
                      device=dev)
    _vbp(p, _angles64(angles, dev), x_t, y_t, center, edge_pad, mode, nproj,
         out, False)
    if single:
        out = out[0]
    return rec._finish(out, as_numpy)


def vertical_back_projection_cpu(projections, angles, xlist, ylist, center,
                                 edge_pad=False):
    """Vertical back-projection with the arithmetic of the reference's CPU
    kernel (vertrec.py:53-110): float64 positions, float32 storage after
    every projection.  Returns (height, width) float32."""
    return _bp_public(projections, angles, xlist, ylist, center, edge_pad,
                      nat.VERT_CPU, True)


def vertical_back_projection_cpu_chunk(projections, angles, x_mat, y_mat,
                                       center, edge_pad=False):
    """Multiple-slice form of vertical_back_projection_cpu
    (vertrec.py:113-173).  Returns (num_slice, height, width) float32."""
    return _bp_public(projections, angles, x_mat, y_mat, center, edge_pad,
                      nat.VERT_CPU, False)


def vertical_back_projection_gpu(projections, angles, xlist, ylist, center,
                                 edge_pad=False, block=(16, 16)):
    """Vertical back-projection with the arithmetic of the reference's GPU
    kernel (vertrec.py:176-253): float32 positions, float64 sum.  Returns
    (height, width) float32."""
    return _bp_public(projections, angles, xlist, ylist, center, edge_pad,
                      nat.VERT_GPU, True)


def vertical_back_projection_gpu_chunk(projections, angles, x_mat, y_mat,
                                       center, edge_pad=False,
                                       block=(16, 16)):
    """Multiple-slice form of vertical_back_projection_gpu
    (vertrec.py:256-335).  Returns (num_slice, height, width) float32."""
    return _bp_public(projections, angles, x_mat, y_mat, center, edge_pad,
                      nat.VERT_GPU, False)


# --------------------------------------------------------------------------
# the reconstruction pipeline (vertrec.py:450-706, :709-977, :980-1252)
# --------------------------------------------------------------------------
def _check_common(projections, angles, crop, proj_start, proj_stop, center,
                  ramp_filter):
    if not (ramp_filter == "before" or ramp_filter == "after"
            or ramp_filter is None):
        raise ValueError("Must use one of these options: "
                         "'before', 'after', None")
    (num_proj0, height0, width0) = projections.shape
    if proj_stop == -1:
        proj_stop = num_proj0
    num_proj = proj_stop - proj_start
    if num_proj < 1:
        raise ValueError("Wrong value of proj_start or proj_stop !!! Given "
                         "the number of projections {}".format(num_proj0))
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, num_proj))
    elif len(angles) != num_proj:
        raise ValueError("!!! Number of angles is not the same as the "
                         "number of projections !!!")
    (cr_top, cr_bottom, cr_left, cr_right) = crop
    top, bot = cr_top, height0 - cr_bottom
    left, right = cr_left, width0 - cr_right
    width, height = right - left, bot - top
    if height < 1 or width < 1:
        raise ValueError("Can't crop images with the given parameters !!!")
    if center < 0 or center > (width - 1):
        raise ValueError("Center (relative to the cropped image) is out of "
                         "range {})".format((0, width - 1)))
    return (proj_stop, num_proj, angles, (top, bot, left, right), height,
            width, height0, width0)


def _flat_dark(flat_field, dark_field, height0, width0, box):
    """(dark, flat - dark with zeros set to 1) cropped, float32, or None
    (vertrec.py:548-570)."""
    if flat_field is None and dark_field is None:
        return None
    if flat_field is None:
        flat_field = np.ones((height0, width0), dtype=np.float32)
    elif flat_field.shape != (height0, width0):
        raise ValueError("!!! Shape of flat-field and projection-image"
                         " is not the same !!!")
    if dark_field is None:
        dark_field = np.zeros((height0, width0), dtype=np.float32)
    elif dark_field.shape != (height0, width0):
        raise ValueError("!!! Shape of dark-field and projection-image"
                         " is not the same !!!")
    top, bot, left, right = box
    flat = np.asarray(flat_field)[top:bot, left:right]
    dark = np.asarray(dark_field)[top:bot, left:right]
    flat_dark = flat - dark
    flat_dark[flat_dark == 0.0] = np.float32(1.0)
    return (np.ascontiguousarray(dark, dtype=np.float32),
            np.ascontiguousarray(flat_dark, dtype=np.float32))


def _chunk(chunk_size, num_proj):
    if chunk_size is None:
        chunk_size = min(30, num_proj // 4 + 1)
    if chunk_size > num_proj or chunk_size < 1:
        chunk_size = num_proj
    return int(chunk_size)


def _upload(projections, n0, n1, box, dev):
    """Cropped projections [n0, n1) on the device: a contiguous float32 or
    uint16 tensor, or a strided view of a device tensor (+ its strides)."""
    import torch
    top, bot, left, right = box
    if rec._is_tensor(projections):
        t = projections[n0:n1, top:bot, left:right]
        if t.dtype not in (torch.float32, torch.uint16):
            t = t.to(torch.float32)
        if not t.is_cuda:
            t = t.contiguous().to(dev)
        return t
    a = np.asarray(projections[n0:n1, top:bot, left:right])
    if a.dtype not in (np.float32, np.uint16):
        a = a.astype(np.float32)
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        return torch.from_numpy(a.view(np.int16)).to(dev).view(torch.uint16)
    return torch.from_numpy(a).to(dev)


def _groups(projections, proj_start, ranges, box, dev):
    """(g0, g1, cropped projections [g0, g1) on the device) per group.  A
    numpy array (a memory map included) streams through the prefetch
    pipeline: group k+1 uploads on a copy stream while group k is
    processed; other inputs go group by group through _upload."""
    import torch
    if isinstance(projections, np.ndarray):
        top, bot, left, right = box
        n0, n1 = proj_start + ranges[0][0], proj_start + ranges[-1][1]
        src = projections[n0:n1, top:bot, left:right]
        if src.dtype == np.uint16:
            src, dtype = src.view(np.int16), torch.int16
        else:
            dtype = torch.float32
        shifted = [(g0 - ranges[0][0], g1 - ranges[0][0]) for g0, g1 in ranges]
        for k, t in _host.prefetch_chunks(src, shifted, dtype, dev):
            g0, g1 = ranges[k]
            yield g0, g1, (t.view(torch.uint16) if dtype == torch.int16
                           else t)
        return
    for g0, g1 in ranges:
        yield g0, g1, _upload(projections, proj_start + g0,
                              proj_start + g1, box, dev)


def _prepare(raw, fd, apply_log, flag, dev):
    """Flat-field + log into a contiguous float32 (n, height, width)."""
    import torch
    n, height, width = raw.shape
    if fd is None and not apply_log and raw.dtype == torch.float32 and \
            raw.is_contiguous():
        return raw
    out = torch.empty((n, height, width), dtype=torch.float32, device=dev)
    dark_t = flat_t = None
    if fd is not None:
        dark_t, flat_t = fd
    code = nat.ATH_U16 if raw.dtype == torch.uint16 else nat.ATH_F32
    st0, st1, _ = raw.stride()
    nat.check(nat.load().ath_flat_field_log(
        nat.ptr(raw), code, n, height, width, st0, st1,
        nat.ptr(dark_t) if dark_t is not None else None,
        nat.ptr(flat_t) if flat_t is not None else None,
        int(bool(apply_log)), nat.ptr(out),
        nat.ptr(flag) if flag is not None else None,
        nat.stream_ptr(dev)), "ath_flat_field_log")
    return out


def _vertical_pipeline(projections, xs, ys, single, center, flat_field,
                       dark_field, angles, crop, proj_start, proj_stop,
                       chunk_size, ramp_filter, filter_name, pad, pad_mode,
                       apply_log, gpu, show_progress, masking, centers=None,
                       mask_lohi=None, keep_device=False):
    """The reconstruction of vertical_reconstruction (vertrec.py:450-695)
    for one or more lines.  centers: candidate rotation centres of one line
    (find_center_vertical_slice), each reconstructed as its own slice with
    the single-slice output dtype; mask_lohi: explicit mask columns."""
    import torch
    (proj_stop, num_proj, angles, box, height, width, height0,
     width0) = _check_common(projections, angles, crop, proj_start,
                             proj_stop, center, ramp_filter)
    dev = nat.device()
    as_numpy = not rec._is_tensor(projections)
    fd = _flat_dark(flat_field, dark_field, height0, width0, box)
    if fd is not None:
        fd = tuple(torch.from_numpy(a).to(dev) for a in fd)
    chunk = _chunk(chunk_size, num_proj)
    if pad is None:
        pad = min(int(0.15 * width), 150)
    edge_pad = ramp_filter != "before"
    taps = None
    if ramp_filter == "before":
        taps = rec._window_taps(None, height, width + 2 * pad, filter_name)
    x_t, y_t = _points_device(xs, ys, dev)
    centers_t = None
    if centers is not None:
        centers_t = torch.from_numpy(np.ascontiguousarray(
            np.asarray(centers, dtype=np.float64))).to(dev)
        x_t = x_t.expand(len(centers), -1).contiguous()
        y_t = y_t.expand(len(centers), -1).contiguous()
    nslice = x_t.shape[0]
    ang_t = _angles64(angles, dev)
    mode = nat.VERT_GPU if gpu else nat.VERT_CPU
    ver = torch.empty((nslice, height, width), dtype=torch.float32,
                      device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev) if apply_log \
        else None
    # whole chunks per pass, so chunk boundaries (and the float32 chunk sums)
    # are those of the reference
    per_proj = height * width * 4 * (2 if ramp_filter == "before" else 1)
    group = max(1, _GROUP_BYTES // max(per_proj * chunk, 1)) * chunk
    t0 = time.time()
    ranges = [(g0, min(num_proj, g0 + group))
              for g0 in range(0, num_proj, group)]
    for g0, g1, raw in _groups(projections, proj_start, ranges, box, dev):
        img = _prepare(raw, fd, apply_log, flag, dev)
        if ramp_filter == "before":
            img = rec._filter_stack(img, pad, pad_mode, False, taps,
                                    torch.float32)
        _vbp(img, ang_t[g0:g1], x_t, y_t, center, edge_pad, mode, chunk, ver,
             g0 > 0, centers_t)
        if show_progress:
            sys.stdout.write("\rProcessed {}/{} images ({:.0f}%) - Time "
                             "elapsed: {:.2f} seconds".format(
                                 g1, num_proj, 100.0 * g1 / num_proj,
                                 time.time() - t0))
            sys.stdout.flush()
    if flag is not None and int(flag.item()) != 0:
        warnings.warn(_LOG_WARNING)
    result = ver
    if ramp_filter == "after":
        # |ramp window| along every slice row (vertrec.py:621-627, 698-706)
        row = np.abs(rec._ramp_row(width + 2 * pad, filter_name))
        taps_a = rec._taps(row, outer_shift=False)[None, :]
        out_dtype = torch.float64 if (single or centers is not None) \
            else torch.float32
        result = rec._filter_stack(ver.reshape(nslice * height, 1, width),
                                   pad, pad_mode, False, taps_a,
                                   out_dtype).reshape(nslice, height, width)
    if show_progress:
        print("\nDone! Total time elapsed: {0:.2f}".format(time.time() - t0))
    lo = hi = -1
    if mask_lohi is not None:
        lo, hi = mask_lohi
    elif masking:
        rad = int(min(center, width - center))
        lo = (width - 2 * rad) // 2
        hi = lo + 2 * rad
    nat.check(nat.load().ath_vertical_finish(
        nat.ptr(result), nat.dtype_code(result.dtype), result.numel(), width,
        lo, hi, num_proj, nat.stream_ptr(dev)), "ath_vertical_finish")
    if single and centers is None:
        result = result[0]
    if keep_device:
        return result
    return rec._finish(result, as_numpy)


def vertical_reconstruction(projections, slice_index, center, alpha=0.0,
                            flat_field=None, dark_field=None, angles=None,
                            crop=(0, 0, 0, 0), proj_start=0, proj_stop=-1,
                            chunk_size=30, ramp_filter="after",
                            filter_name="hann", pad=None, pad_mode="edge",
                            apply_log=True, gpu=True, block=(16, 16),
                            ncore=None, prefer="threads", show_progress=True,
                            masking=False):
    """Reconstruct a vertical slice from a stack of projection-images
    (num_projection, height, width) with optional use of the ramp-filter
    (vertrec.py:450-695).  Returns (height, width): float64 with
    ramp_filter="after", float32 otherwise."""
    width = _check_common(projections, angles, crop, proj_start, proj_stop,
                          center, ramp_filter)[5]
    if slice_index < 0 or slice_index > (width - 1):
        raise ValueError("Index (relative to the cropped image) is out of "
                         "range {})".format((0, width - 1)))
    xs, ys = _get_points_single_line(slice_index, alpha, width)
    return _vertical_pipeline(projections, xs, ys, True, center, flat_field,
                              dark_field, angles, crop, proj_start,
                              proj_stop, chunk_size, ramp_filter,
                              filter_name, pad, pad_mode, apply_log, gpu,
                              show_progress, masking)


def vertical_reconstruction_multiple(projections, start_index, stop_index,
                                     center, alpha=0.0, step_index=1,
                                     flat_field=None, dark_field=None,
                                     angles=None, crop=(0, 0, 0, 0),
                                     proj_start=0, proj_stop=-1,
                                     chunk_size=30, ramp_filter="after",
                                     filter_name="hann", pad=None,
                                     pad_mode="edge", apply_log=True,
                                     gpu=True, block=(16, 16), ncore=None,
                                     prefer="threads", show_progress=True,
                                     masking=False):
    """Reconstruct parallel vertical slices start_index..stop_index
    (vertrec.py:709-977).  Returns (num_slice, height, width) float32."""
    width = _check_common(projections, angles, crop, proj_start, proj_stop,
                          center, ramp_filter)[5]
    if start_index < 0 or start_index > (width - 1):
        raise ValueError("Start index (relative to the cropped image) is out "
                         "of range {})".format((0, width - 1)))
    if stop_index == -1:
        stop_index = width - 1
    if stop_index < 0 or stop_index > (width - 1):
        raise ValueError("Start index (relative to the cropped image) is out "
                         "of range {})".format((0, width - 1)))
    xs, ys = _get_points_multiple_lines(start_index, stop_index, alpha,
                                        width, step_index)
    return _vertical_pipeline(projections, xs, ys, False, center, flat_field,
                              dark_field, angles, crop, proj_start,
                              proj_stop, chunk_size, ramp_filter,
                              filter_name, pad, pad_mode, apply_log, gpu,
                              show_progress, masking)


def vertical_reconstruction_different_angles(projections, slice_indices,
                                             alphas, center, flat_field=None,
                                             dark_field=None, angles=None,
                                             crop=(0, 0, 0, 0), proj_start=0,
                                             proj_stop=-1, chunk_size=30,
                                             ramp_filter="after",
                                             filter_name="hann", pad=None,
                                             pad_mode="edge", apply_log=True,
                                             gpu=True, block=(16, 16),
                                             ncore=None, prefer="threads",
                                             show_progress=True,
                                             masking=False):
    """Reconstruct vertical slices each at its own index and angle
    (vertrec.py:980-1252).  Returns (num_slice, height, width) float32."""
    if len(slice_indices) != len(alphas):
        raise ValueError("!!! Number of slice indices is not the same as the "
                         "number of angles !!!")
    width = _check_common(projections, angles, crop, proj_start, proj_stop,
                          center, ramp_filter)[5]
    if any(idx < 0 or idx > width - 1 for idx in slice_indices):
        raise ValueError("Slice index (relative to the cropped image) is out "
                         "of range {})".format((0, width - 1)))
    xs, ys = _get_points_multiple_lines_different_angles(slice_indices,
                                                         alphas, width)
    return _vertical_pipeline(projections, xs, ys, False, center, flat_field,
                              dark_field, angles, crop, proj_start,
                              proj_stop, chunk_size, ramp_filter,
                              filter_name, pad, pad_mode, apply_log, gpu,
                              show_progress, masking)


# --------------------------------------------------------------------------
# centre search on a vertical slice (vertrec.py:1255-1428)
# --------------------------------------------------------------------------
def _gaussian_2d(p, sigma):
    """ndi.gaussian_filter(p, (0, sigma, sigma)) of a (n, height, width)
    float32 device tensor: axis 1 then axis 2, float32 between the passes
    (vertrec.py:1331-1332)."""
    import torch
    from algotom_amd.rec import center as ctr
    w, radius = ctr._gaussian_weights(sigma)
    w_d = torch.from_numpy(np.ascontiguousarray(w)).to(p.device)
    n, height, width = p.shape
    lib = nat.load()
    mid = torch.empty_like(p)
    nat.check(lib.ath_gaussian_filter_axis(
        nat.ptr(p), nat.ptr(mid), n, height, width, nat.ptr(w_d), radius,
        nat.stream_ptr(p.device)), "ath_gaussian_filter_axis")
    out = torch.empty_like(p)
    nat.check(lib.ath_gaussian_filter_axis(
        nat.ptr(mid), nat.ptr(out), n * height, width, 1, nat.ptr(w_d),
        radius, nat.stream_ptr(p.device)), "ath_gaussian_filter_axis")
    return out


_METRICS = {"entropy": 0, "sharpness": 1, "autocorrelation": 2}
_SEARCH_BYTES = 2 << 30


def find_center_vertical_slice(projections, slice_index, start, stop,
                               step=1.0, metric="entropy", alpha=0.0,
                               angles=None, chunk_size=30,
                               ramp_filter="after", sigma=3, apply_log=True,
                               gpu=True, block=(16, 16), ncore=None,
                               prefer="threads", show_progress=True,
                               masking=True, return_metric=False,
                               invert_metric=False, metric_function=None,
                               **kwargs):
    """Find the centre of rotation from metrics of a vertical slice
    reconstructed at centres arange(start, stop + step, step)
    (vertrec.py:1264-1428); the minimum metric gives the centre.

    The projections are smoothed (Gaussian over rows and columns) and
    logged once; every candidate centre is then reconstructed as its own
    line in one vertical back-projection launch (per pass of whole chunks)
    and all metrics -- "entropy", "sharpness", "autocorrelation" -- are
    computed on the GPU.  A custom metric_function is called on each slice
    as a numpy array.  Returns the centre, or (list_center, list_metric)
    if return_metric.
    """
    import torch
    if (metric not in _METRICS) and metric_function is None:
        raise ValueError("Please select one of three options: 'entropy', "
                         "'sharpness', 'autocorrelation'; or provide a "
                         "custom method.")
    if ramp_filter is None:
        ramp_filter = "after"
    list_center = np.arange(start, stop + step, step)
    (num_proj, height, width) = projections.shape
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, num_proj))
    dev = nat.device()
    p = rec._to_device(projections, dev, allow_f64=False)
    if sigma > 0:
        p = _gaussian_2d(p, sigma)
    if apply_log:
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        p = _prepare(p, None, True, flag, dev)
        if int(flag.item()) != 0:
            warnings.warn(_LOG_WARNING)
    lohi = None
    if masking:
        cmin = np.min(list_center)
        rad = int(min(cmin, width - cmin))
        lo = (width - 2 * rad) // 2
        lohi = (lo, lo + 2 * rad)
    pad = min(int(0.15 * width), 150)
    xs, ys = _get_points_single_line(slice_index, alpha, width)
    if np.any(list_center < 0) or np.any(list_center > width - 1):
        raise ValueError("Center (relative to the cropped image) is out of "
                         "range {})".format((0, width - 1)))

    def recon(cs, mask):
        return _vertical_pipeline(
            p, xs, ys, True, float(cs[0]), None, None, angles, (0, 0, 0, 0),
            0, -1, chunk_size, ramp_filter, "hann", pad, "edge", False, gpu,
            False, False, centers=cs, mask_lohi=mask, keep_device=True)

    nmin, nmax = 0.0, 1.0
    if metric_function is None and metric == "entropy":
        ref = recon([np.mean(list_center)], None)
        nmin = float(torch.min(ref).item())
        nmax = float(torch.max(ref).item())
    esize = 8 if ramp_filter == "after" else 4
    per_batch = max(1, _SEARCH_BYTES // (esize * height * width))
    lib = nat.load()
    values = []
    t0 = time.time()
    for b0 in range(0, len(list_center), per_batch):
        cs = list_center[b0:b0 + per_batch]
        imgs = recon(cs, lohi)
        if metric_function is not None:
            for k in range(len(cs)):
                values.append(metric_function(imgs[k].cpu().numpy(),
                                              **kwargs))
        else:
            code = _METRICS[metric]
            out = torch.empty(len(cs), dtype=torch.float64, device=dev)
            nws = lib.ath_slice_metrics_workspace_bytes(len(cs), height,
                                                        width, code)
            ws = torch.empty(max(nws, 1), dtype=torch.uint8, device=dev)
            nat.check(lib.ath_slice_metrics(
                nat.ptr(imgs), nat.dtype_code(imgs.dtype), len(cs), height,
                width, height * width, nmin, nmax, code, nat.ptr(out),
                nat.ptr(ws), nws, nat.stream_ptr(dev)), "ath_slice_metrics")
            values.extend(out.cpu().numpy().tolist())
        if show_progress:
            sys.stdout.write("\rReconstructed {}/{} centres - Time elapsed: "
                             "{:.2f} seconds".format(
                                 min(b0 + per_batch, len(list_center)),
                                 len(list_center), time.time() - t0))
            sys.stdout.flush()
    list_metric = np.zeros_like(list_center)
    list_metric[:] = values
    if invert_metric:
        list_metric = np.max(list_metric) - list_metric
    center = list_center[np.argmin(list_metric)]
    if show_progress:
        print("\nDone! Optimal center: {0}.".format(center))
    if return_metric is True:
        return list_center, list_metric
    return center


def find_center_visual_vertical_slices(projections, output_folder,
                                       slice_index, start, stop, step=1.0,
                                       alpha=0.0, angles=None, chunk_size=30,
                                       ramp_filter="after", apply_log=True,
                                       gpu=True, block=(16, 16), ncore=None,
                                       prefer="processes", display=True,
                                       masking=True):
    """For finding the centre of rotation visually: a vertical slice
    reconstructed at every centre of arange(start, stop + step, step),
    each saved as ``center_<c:.2f>.tif`` in a new ``Find_center_NNN``
    folder under ``output_folder`` (vertrec.py:1431-1536).  Returns that
    folder's path (with a trailing "/").

    The projections are logged once (values <= 0 are set to 1 in place,
    with the reference's warning); all centres of a pass are then
    reconstructed as separate lines of one vertical back-projection launch.
    The mask keeps the columns of radius int(min(c_min, width - c_min))
    about the middle, as the reference does.
    """
    import torch
    from algotom_amd.io import loadersaver as losa
    output_name = losa.make_folder_name(output_folder,
                                        name_prefix="Find_center",
                                        zero_prefix=3)
    output_base = output_folder + "/" + output_name + "/"
    list_center = np.arange(start, stop + step, step)
    (num_proj, height, width) = projections.shape
    if angles is None:
        angles = np.deg2rad(np.linspace(0.0, 180.0, num_proj))
    if np.any(list_center < 0) or np.any(list_center > width - 1):
        raise ValueError("Center (relative to the cropped image) is out of "
                         "range {})".format((0, width - 1)))
    dev = nat.device()
    p = rec._to_device(projections, dev, allow_f64=False)
    if apply_log:
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        p = _prepare(p, None, True, flag, dev)
        if int(flag.item()) != 0:
            # vertrec.py:1493-1502: warn and set the values <= 0 of the
            # caller's array to 1 before the log
            warnings.warn(_LOG_WARNING)
            projections[projections <= 0.0] = 1.0
    if ramp_filter is None:
        ramp_filter = "after"
    lohi = None
    if masking:
        cmin = np.min(list_center)
        rad = int(min(cmin, width - cmin))
        lo = (width - 2 * rad) // 2
        lohi = (lo, lo + 2 * rad)
    pad = min(int(0.15 * width), 150)
    xs, ys = _get_points_single_line(slice_index, alpha, width)
    esize = 8 if ramp_filter == "after" else 4
    per_batch = max(1, _SEARCH_BYTES // (esize * height * width))
    t0 = time.time()
    for b0 in range(0, len(list_center), per_batch):
        cs = list_center[b0:b0 + per_batch]
        imgs = _vertical_pipeline(
            p, xs, ys, True, float(cs[0]), None, None, angles, (0, 0, 0, 0),
            0, -1, chunk_size, ramp_filter, "hann", pad, "edge", False, gpu,
            False, False, centers=cs, mask_lohi=lohi, keep_device=True)
        imgs = rec._to_host(imgs)
        for k, center in enumerate(cs):
            file_name = "center_{0:.2f}".format(center) + ".tif"
            losa.save_image(output_base + file_name, imgs[k])
            if display:
                print("Done: {0}. Time elapsed {1:.2f}".format(
                    output_base + file_name, time.time() - t0))
    return output_base
