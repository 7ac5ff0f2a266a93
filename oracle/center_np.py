"""Oracle for the centre search by slice metrics -- TEST INFRASTRUCTURE ONLY.

Restates algotom/rec/reconstruction.py:1183-1208 (_reconstruct_slice, the
methods "fbp", "bpf" and "dfi"), :1211-1228 (the two slice metrics),
:1231-1251 (metric of one slice), :1254-1362
(_find_center_based_slice_metric) and :1365-1458
(find_center_based_slice_metric).  The metrics are compositions of numpy /
scipy calls in the reference, so they are evaluated with the same calls
here; the reconstructions come from the oracle's FBP port (oracle/port.py),
BPF (oracle/bpf_np.py) and DFI (oracle/dfi_np.py) restatements.

Pinned in tests/test_oracle_center.py by closed forms (the entropy of a
constant image, the sharpness of a linear ramp, the histogram level of
every uint8 value counted by brute force) and by the search finding the
known rotation centre of an analytic off-centre phantom.
"""

import numpy as np
import scipy.ndimage as ndi
from scipy import signal

from oracle import bpf_np, dfi_np, port


def histogram_entropy(img):
    """reconstruction.py:1211-1220 on an image already scaled to [0, 1]."""
    levels = np.uint8(img * 255)
    counts = ndi.histogram(levels, 0, 255, 256)
    smooth = signal.convolve(1.0 + counts, signal.windows.boxcar(7),
                             mode="valid")
    return -np.sum(smooth * np.log2(smooth))


def edge_sharpness(img):
    """reconstruction.py:1223-1228: mean gradient magnitude, float32."""
    g0, g1 = np.gradient(np.float32(img))
    return np.mean(np.sqrt(g0 ** 2 + g1 ** 2))


def slice_metric(recon, metric="entropy", nmin=0.0, nmax=1.0):
    """reconstruction.py:1240-1247."""
    if metric == "entropy":
        scaled = np.clip((recon - nmin) / (nmax - nmin), 0, 1)
        return histogram_entropy(scaled)
    return edge_sharpness(recon)


def preprocess(sinogram, sigma=2, apply_log=True):
    """Smoothing and log of reconstruction.py:1315-1322 (float32 input).
    The log is evaluated in float64 and rounded, the convention of the
    oracle's fbp."""
    sino = np.asarray(sinogram, dtype=np.float32)
    if sigma > 0:
        sino = ndi.gaussian_filter1d(sino, sigma, axis=1)
    else:
        sino = sino.copy()
    if apply_log:
        sino[sino <= 0.0] = np.float32(1.0)
        sino = (-np.log(sino.astype(np.float64))).astype(np.float32)
    return sino


def reconstruct_slice(sino, center, method="fbp", angles=None, ratio=1.0,
                      filter_name="hann", gpu=True):
    """_reconstruct_slice (reconstruction.py:1183-1208) of a logged
    sinogram: every name but "fbp" and "bpf" goes to DFI, as there
    ("gridrec" and "astra" need external libraries and are not restated)."""
    if method == "fbp":
        return port.fbp(sino, center, angles=angles, ratio=ratio,
                        filter_name=filter_name, apply_log=False, gpu=gpu)
    if method == "bpf":
        return bpf_np.bpf(sino, center, angles=angles, ratio=ratio,
                          filter_name=filter_name, apply_log=False, gpu=gpu)
    return dfi_np.dfi(sino, center, angles=angles, ratio=ratio,
                      filter_name=filter_name, apply_log=False)


def center_metrics(sinogram, start, stop, step=1, metric="entropy",
                   gpu=True, angles=None, ratio=1.0, filter_name="hann",
                   apply_log=True, sigma=2, invert_metric=False,
                   method="fbp"):
    """_find_center_based_slice_metric (reconstruction.py:1254-1362) with
    return_metric=True: (list_center, list_metric)."""
    centers = np.arange(start, stop, step)
    if len(centers) == 0:
        raise ValueError("empty centre list")
    sino = preprocess(sinogram, sigma, apply_log)

    def recon(c):
        return reconstruct_slice(sino, c, method, angles, ratio, filter_name,
                                 gpu)

    nmin, nmax = 0.0, 1.0
    if metric == "entropy":
        ref = recon(np.mean(centers))
        nmin, nmax = np.min(ref), np.max(ref)
        if nmin == nmax:
            raise ValueError("Empty image!!!")
    values = np.ones(len(centers))
    for i, c in enumerate(centers):
        values[i] = slice_metric(recon(c), metric, nmin, nmax)
    if invert_metric:
        values = np.max(values) - values
    return centers, values


def find_center(sinogram, start, stop, step=0.5, metric="entropy",
                radius=2, zoom=1.0, gpu=True, angles=None, ratio=1.0,
                filter_name="hann", apply_log=True, sigma=2, method="fbp"):
    """find_center_based_slice_metric (reconstruction.py:1365-1458):
    integer-step coarse search (on the zoomed sinogram), then a sub-pixel
    search within +-radius of the coarse centre."""
    zoom = float(np.clip(zoom, 0.01, 1.0))
    sino_z, angles_z = sinogram, angles
    if zoom < 1.0:
        sino_z = ndi.zoom(np.asarray(sinogram, np.float32), zoom, order=1,
                          mode="nearest")
        start, stop = start * zoom, stop * zoom
        if angles is not None:
            angles_z = ndi.zoom(np.tile(angles, (1, 1)), (1.0, zoom))[0]
    kw = dict(metric=metric, gpu=gpu, ratio=ratio, filter_name=filter_name,
              apply_log=apply_log, sigma=sigma, method=method)
    if stop > start:
        cs, vs = center_metrics(sino_z, start, stop + 1, 1.0,
                                angles=angles_z, **kw)
        coarse = cs[np.argmin(vs)] / zoom
    else:
        coarse = start / zoom
    if radius == 0.0:
        return coarse
    radius = max(radius, step + 1.0 / zoom)
    cs, vs = center_metrics(sinogram, coarse - radius, coarse + radius + step,
                            step, angles=angles, **kw)
    return cs[np.argmin(vs)]
