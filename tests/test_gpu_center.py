"""Centre search on the GPU against the oracle (oracle/center_np.py).

* Gaussian row smoothing == scipy.ndimage.gaussian_filter1d, bit for bit;
* order-1 zoom == scipy.ndimage.zoom(order=1, mode="nearest") to float32
  rounding, same output shape;
* sharpness metric == the oracle's numpy evaluation bit for bit (the kernel
  follows numpy's summation order); entropy metric within 1e-12 relative
  (the two libraries' log2 may differ in the last bit);
* the whole search (entropy / sharpness, both accumulation modes, with and
  without zoom): the same centres and metrics as the oracle, and the known
  rotation centre of an analytic phantom.
"""

import numpy as np
import pytest
import scipy.ndimage as ndi

pytestmark = pytest.mark.gpu


def _gauss(x, sigma):
    import torch
    from algotom_amd.rec import center
    t = torch.from_numpy(x).cuda()
    return center._smooth_rows(t, sigma).cpu().numpy()


@pytest.mark.parametrize("shape,sigma", [((181, 300), 2), ((37, 5), 2),
                                         ((64, 129), 1), ((16, 1000), 3.5),
                                         ((8, 1), 2)])
def test_gaussian_rows_bitwise(shape, sigma):
    rng = np.random.default_rng(0)
    x = rng.random(shape, dtype=np.float32) * np.float32(2.0)
    got = _gauss(x, sigma)
    ref = ndi.gaussian_filter1d(x, sigma, axis=1)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("zoom", [0.5, 0.37, 0.8, 0.25])
def test_zoom_linear_vs_scipy(zoom):
    import torch
    from algotom_amd.rec import center
    rng = np.random.default_rng(1)
    x = rng.random((181, 300), dtype=np.float32)
    got = center._zoom_linear(torch.from_numpy(x).cuda(), zoom).cpu().numpy()
    ref = ndi.zoom(x, zoom, order=1, mode="nearest")
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 4 * np.finfo(np.float32).eps


def _metrics_gpu(imgs, metric, nmin=0.0, nmax=1.0):
    import torch
    from algotom_amd import _native as nat
    t = torch.from_numpy(np.ascontiguousarray(imgs)).cuda()
    n, c, _ = imgs.shape
    code = 0 if metric == "entropy" else 1
    lib = nat.load()
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    nws = lib.ath_slice_metrics_workspace_bytes(n, c, c, code)
    ws = torch.full((nws,), 0xAB, dtype=torch.uint8, device="cuda")
    nat.check(lib.ath_slice_metrics(
        nat.ptr(t), nat.ATH_F32, n, c, c, c * c, float(np.float32(nmin)),
        float(np.float32(nmax)), code, nat.ptr(out), nat.ptr(ws), nws,
        nat.stream_ptr(t.device)), "ath_slice_metrics")
    return out.cpu().numpy()


@pytest.mark.parametrize("size", [2, 64, 97, 1000, 2560])
def test_sharpness_bitwise(size):
    from oracle import center_np
    rng = np.random.default_rng(size)
    imgs = rng.standard_normal((3, size, size)).astype(np.float32)
    imgs[1] = np.cumsum(imgs[1], axis=0) * np.float32(0.1)
    got = _metrics_gpu(imgs, "sharpness")
    ref = [np.float64(center_np.edge_sharpness(im)) for im in imgs]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("size", [1, 64, 300, 1000])
def test_entropy_vs_oracle(size):
    from oracle import center_np
    rng = np.random.default_rng(size + 7)
    imgs = rng.standard_normal((4, size, size)).astype(np.float32)
    imgs[2] = imgs[2] * np.float32(0.01)     # most pixels in a few bins
    nmin, nmax = np.float32(-1.5), np.float32(2.0)  # clipping on both ends
    got = _metrics_gpu(imgs, "entropy", nmin, nmax)
    ref = [center_np.slice_metric(im, "entropy", nmin, nmax) for im in imgs]
    assert np.allclose(got, ref, rtol=1e-12, atol=0)


def _phantom(nrow=181, ncol=96, center=51.0, seed=2):
    import torch
    from algotom_amd.util import phantom
    params = phantom.random_ellipses(64, count=10, seed=seed,
                                     axis_range=(3.0, 20.0))
    angles = np.deg2rad(np.linspace(0.0, 180.0, nrow))
    proj = phantom.ellipse_sinogram(params, torch.from_numpy(angles), ncol,
                                    center)
    return phantom.transmission(proj, photons=1e6, seed=seed).numpy(), angles


@pytest.mark.parametrize("metric", ["entropy", "sharpness"])
@pytest.mark.parametrize("gpu", [True, False])
def test_center_metrics_vs_oracle(metric, gpu):
    from oracle import center_np
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom()
    cs, vs = rec._find_center_based_slice_metric(
        sino.copy(), 45, 57.5, 0.5, metric=metric, gpu=gpu, angles=angles)
    rcs, rvs = center_np.center_metrics(sino, 45, 57.5, 0.5, metric=metric,
                                        gpu=gpu, angles=angles)
    assert np.array_equal(cs, rcs)
    if metric == "sharpness":
        assert np.array_equal(vs, rvs)
    else:
        assert np.allclose(vs, rvs, rtol=1e-12, atol=0)
    assert cs[np.argmin(vs)] == rcs[np.argmin(rvs)] == 51.0


@pytest.mark.parametrize("method,gpu", [("bpf", True), ("bpf", False),
                                        ("dfi", True)])
@pytest.mark.parametrize("metric", ["entropy", "sharpness"])
def test_center_metrics_bpf_dfi_vs_oracle(method, gpu, metric):
    """method="bpf" / "dfi" (reconstruction.py:1188-1207): float64 slices
    whose values match the oracle's to FFT rounding (~1e-12), so a few
    pixels may land in the neighbouring histogram level or float32 value;
    bar: metrics within 1e-6 relative and the same chosen centre."""
    from oracle import center_np
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom()
    cs, vs = rec._find_center_based_slice_metric(
        sino.copy(), 47, 55.5, 0.5, metric=metric, gpu=gpu, angles=angles,
        method=method)
    rcs, rvs = center_np.center_metrics(sino, 47, 55.5, 0.5, metric=metric,
                                        gpu=gpu, angles=angles, method=method)
    rel = np.abs(vs - rvs).max() / np.abs(rvs).max()
    print("%s gpu=%s %s: max rel metric diff %.3g, centre %.1f" %
          (method, gpu, metric, rel, cs[np.argmin(vs)]))
    assert np.array_equal(cs, rcs)
    assert rel <= 1e-6
    assert cs[np.argmin(vs)] == rcs[np.argmin(rvs)]
    assert abs(cs[np.argmin(vs)] - 51.0) <= 0.5


def test_find_center_bpf_ratio_zero_vs_oracle():
    """ratio=0: a mask radius per centre (one filter call per radius)."""
    from oracle import center_np
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom(nrow=241, ncol=128, center=66.5, seed=6)
    got = rec.find_center_based_slice_metric(sino.copy(), 60, 72, 0.5,
                                             method="bpf", ratio=0.0,
                                             angles=angles)
    ref = center_np.find_center(sino, 60, 72, 0.5, method="bpf", ratio=0.0,
                                angles=angles)
    assert got == ref
    assert abs(got - 66.5) <= 0.5


@pytest.mark.parametrize("zoom", [1.0, 0.5])
@pytest.mark.parametrize("metric", ["entropy", "sharpness"])
def test_find_center_vs_oracle(zoom, metric):
    from oracle import center_np
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom(nrow=361, ncol=160, center=83.5, seed=4)
    got = rec.find_center_based_slice_metric(sino.copy(), 75, 92, 0.5,
                                             metric=metric, zoom=zoom,
                                             angles=angles)
    ref = center_np.find_center(sino, 75, 92, 0.5, metric=metric, zoom=zoom,
                                angles=angles)
    print("zoom %.2f %s: gpu %.2f oracle %.2f" % (zoom, metric, got, ref))
    assert got == ref
    assert abs(got - 83.5) <= 0.5


def test_center_search_on_c1_geometry():
    """721 angles x 512 columns, entropy, fine step 0.25."""
    from oracle import center_np
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom(nrow=721, ncol=512, center=261.25, seed=9)
    cs, vs = rec._find_center_based_slice_metric(sino.copy(), 256, 266,
                                                 0.25, angles=angles)
    rcs, rvs = center_np.center_metrics(sino, 256, 266, 0.25, angles=angles)
    assert np.array_equal(cs, rcs)
    assert np.allclose(vs, rvs, rtol=1e-12, atol=0)
    assert abs(cs[np.argmin(vs)] - 261.25) <= 0.25


def test_errors():
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom()
    with pytest.raises(ValueError):
        rec._find_center_based_slice_metric(sino.copy(), 50, 40, 1.0)
    with pytest.raises(ValueError):
        rec._find_center_based_slice_metric(sino.copy(), 90, 100, 1.0)
    for method in ("gridrec", "astra"):
        with pytest.raises(NotImplementedError):
            rec._find_center_based_slice_metric(sino.copy(), 45, 50, 1.0,
                                                method=method)
