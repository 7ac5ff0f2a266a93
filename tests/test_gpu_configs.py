"""GPU parity at the workloads' own sizes, against the oracle (no
self-comparison): the batched uneven workload (every width x angle-count
pair it draws), the differentiable workload's 16 x 901 x 1024 projector
pair over the 120-degree arc, and a slice of the large half-acquisition
volume (3601 x 4096 over 360 degrees, stitched to 7192 columns).

Bars:
* reference arithmetic vs the oracle's numba-typed C port: max|gpu - port|
  <= 1e-4 * max|port| (the two filters' float32 roundings may differ by an
  ulp on rare samples, so not every pixel need be bit-identical; the count
  of differing pixels is printed);
* fp32 / fp16 / bf16 storage vs the all-float64 oracle: <= 1e-3 relative;
* the forward projection (adjoint) vs the port's scatter transpose:
  <= 1e-6 * max (float64 sums in different orders, rounded to float32);
* <B s, x> = <s, B^T x> at full size to 2e-6 relative.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WIDTHS = (1024, 2048, 2560, 4096)
NANGLES = (901, 1801, 3601)


def _flush(msg):
    print(msg, flush=True)


def _uneven_batch(dev):
    """One exact-projection sinogram per (width, angle count) pair, centres
    (w - 1)/2 + U[-20, 20] (seed 3000), angles over [0, pi)."""
    import torch
    from algotom_amd.util import phantom
    rng = np.random.default_rng(3000)
    sinos, centers, angles = [], [], []
    for w in WIDTHS:
        for n in NANGLES:
            c = float((w - 1) / 2 + rng.uniform(-20.0, 20.0))
            ang = np.linspace(0.0, np.pi, n, endpoint=False)
            prm = phantom.random_ellipses(w, 20, seed=len(sinos) + 3000)
            s = phantom.ellipse_sinogram(prm, ang, w, np.float32(c),
                                         device=dev).float()
            sinos.append(s)
            centers.append(c)
            angles.append(ang)
    torch.cuda.synchronize()
    return sinos, centers, angles


@pytest.fixture(scope="module")
def uneven():
    from algotom_amd import _native as nat
    from oracle import port
    dev = nat.device()
    sinos, centers, angles = _uneven_batch(dev)
    host = [s.cpu().numpy() for s in sinos]
    _flush("uneven: 12 sinograms ready; oracle port (reference arithmetic)")
    ref = [port.fbp(h, c, a, apply_log=False, gpu=True)
           for h, c, a in zip(host, centers, angles)]
    _flush("uneven: oracle float64")
    ref64 = [port.fbp_f64(h, c, a, apply_log=False)
             for h, c, a in zip(host, centers, angles)]
    return sinos, centers, angles, ref, ref64


@pytest.mark.parametrize("precision", [None, "fp32", "fp16", "bf16"])
def test_uneven_batch_every_shape_vs_oracle(uneven, precision):
    from algotom_amd.rec import batch
    sinos, centers, angles, ref, ref64 = uneven
    plan = batch.BatchFBP([tuple(s.shape) for s in sinos], centers, angles,
                          apply_log=False, precision=precision)
    plan.load(sinos)
    plan.autotune(persist=False)
    plan.capture()
    plan.run()
    plan.run()  # a second replay: the graph's counter reset works
    got = plan.results()
    worst = 0.0
    for i, g in enumerate(got):
        n, w = sinos[i].shape
        if precision is None:
            r = ref[i]
            err = np.abs(g - r).max() / np.abs(r).max()
            ndiff = int((g != r).sum())
            _flush("%dx%d c=%.2f reference: max rel %.3g, %d px differ"
                   % (n, w, centers[i], err, ndiff))
            assert err <= 1e-4
        else:
            r = ref64[i]
            err = np.abs(g.astype(np.float64) - r).max() / np.abs(r).max()
            _flush("%dx%d %s vs fp64: max rel %.3g" % (n, w, precision, err))
            assert err <= 1e-3
        worst = max(worst, err)
    _flush("precision %s: worst %.3g" % (precision, worst))
    del plan


@pytest.fixture(scope="module")
def refine_case():
    """The differentiable workload's geometry: 16 x 901 x 1024 over
    linspace(0, 2 pi / 3, 901), centre near mid-width."""
    import torch
    from algotom_amd import _native as nat
    dev = nat.device()
    rng = np.random.default_rng(4000)
    nrow, nsl, ncol = 901, 16, 1024
    angles = np.linspace(0.0, 2 * np.pi / 3, nrow)
    center = 511.87
    img = torch.from_numpy(rng.standard_normal((ncol, nsl, ncol))
                           .astype(np.float32)).to(dev)
    sino = torch.from_numpy(rng.standard_normal((nrow, nsl, ncol))
                            .astype(np.float32)).to(dev)
    return angles, center, img, sino


def test_forward_projection_full_size_vs_port(refine_case):
    from oracle import port
    from algotom_amd.rec import autograd
    angles, center, img, _ = refine_case
    geom = autograd.FBPGeometry(901, 1024, angles, ratio=1.0)
    got = geom.forward_project(img, center).cpu().numpy()
    _flush("refine: GPU forward projection done; oracle scatter")
    ref = port.forward_project(img.cpu().numpy(), angles, center,
                               mask_radius=geom.mask_radius,
                               scale=geom.sigma)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    _flush("forward 16x901x1024 (120 deg): max rel %.3g" % err)
    assert err <= 1e-6


def test_adjoint_identity_full_size(refine_case):
    import torch
    from algotom_amd.rec import autograd
    angles, center, img, sino = refine_case
    geom = autograd.FBPGeometry(901, 1024, angles, ratio=1.0)
    bs = geom.backproject(sino, center)      # sigma * mask * B s
    fx = geom.forward_project(img, center)   # sigma * B^T (mask * x)
    lhs = torch.sum(bs.double() * img.double()).item()
    rhs = torch.sum(sino.double() * fx.double()).item()
    _flush("<Bs, x> %.12g  <s, B^T x> %.12g" % (lhs, rhs))
    assert abs(lhs - rhs) <= 2e-6 * (abs(lhs) + abs(rhs))


def test_volume_slice_half_acquisition_vs_oracle():
    """One slice of the large volume workload: 3601 x 4096 over [0, 2 pi)
    with the axis at column 499.5 (overlap 1000), Poisson transmission;
    reconstruct_volume_360 against the oracle's conversion and port FBP."""
    import torch
    from algotom_amd import _native as nat
    from algotom_amd.rec import volume
    from algotom_amd.util import phantom
    from oracle import conversion_np, port
    dev = nat.device()
    nrow, ncol, cor = 3601, 4096, 499.5
    step = 2.0 * np.pi / nrow
    ang360 = np.arange(nrow) * step
    prm = phantom.random_ellipses(7192, 20, seed=5000)
    p = phantom.ellipse_sinogram(prm, ang360, ncol, cor, device=dev)
    s = phantom.transmission(p, seed=5000)
    n180 = nrow // 2 + 1
    angles = np.arange(n180) * step
    got, new_cor = volume.reconstruct_volume_360(
        s.reshape(nrow, 1, ncol), cor, chunk=1, angles=angles)
    got = got[:, 0, :].cpu().numpy()
    _flush("volume slice: GPU done (%d x %d); oracle" % got.shape)
    host = s.cpu().numpy()
    stitched, rc = conversion_np.convert_360_to_180(host, cor)
    assert rc == new_cor and stitched.shape == (n180, 7192)
    ref = port.fbp(stitched, new_cor, angles, apply_log=True, gpu=True)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    ndiff = int((got != ref).sum())
    _flush("volume slice 1801x7192: max rel %.3g, %d px differ" % (err, ndiff))
    assert err <= 1e-4
