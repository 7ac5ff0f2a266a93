"""End-to-end GPU parity of the fbp_reconstruction drop-in.

Bar: max|gpu - oracle| <= 1e-4 * max|oracle| (the float32 contract), where
the oracle is the numba-typed C port with the same accumulation as the
reference kernel the call selects; the error against the all-FP64 oracle is
printed for every case.  Also re-asserts the reference's own test
(tests/test_rec/test_reconstruction.py:88-144) against our drop-in.
"""

import os
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ref_case():
    from __graft_entry__ import _reference_test_case
    return _reference_test_case()


def _fp64_oracle(sino, center, angles, apply_log, ratio=1.0,
                 filter_name="hann"):
    """All-float64 FBP (same f32 rounding of centre/angles as the reference,
    nothing else rounded) for the reported error."""
    from oracle import fbp_np
    x = np.asarray(sino, dtype=np.float64)
    two_d = x.ndim == 2
    if two_d:
        x = x[:, None, :]
    if apply_log:
        x = -np.log(x)
    nrow, nslice, ncol = x.shape
    filt = fbp_np.ramp_filter(np.moveaxis(x, 1, 0), None, filter_name, None)
    ang = np.asarray(angles, dtype=np.float32).astype(np.float64)
    cen = float(np.float32(center))
    out = np.stack([fbp_np.backproject_f64(filt[j], ang, cen)
                    for j in range(nslice)], axis=1)
    if ratio is not None:
        if ratio == 0.0:
            ratio = min(center, ncol - center) / (0.5 * ncol)
        out = out * fbp_np.circle_mask(ncol, ratio)[:, None, :]
    out = out * np.pi / (nrow - 1)
    return out[:, 0, :] if two_d else out


def test_reference_thresholds():
    """test_reconstruction.py:88-144, against our drop-in."""
    from algotom_amd.rec import reconstruction as rec
    mat, sino_180, angles2, center = _ref_case()
    import scipy.ndimage as ndi
    sino_360 = np.zeros((73, 65), dtype=np.float32)
    ang360 = np.linspace(0.0, 360.0, 73, dtype=np.float32)
    for i, a in enumerate(ang360):
        sino_360[i] = np.sum(ndi.rotate(mat, -a, reshape=False), axis=0)
    r1 = rec.fbp_reconstruction(sino_180, center, apply_log=False, gpu=False)
    assert np.median(np.abs(mat - r1)) <= 0.2
    r1 = rec.fbp_reconstruction(sino_180, center, apply_log=False, gpu=True)
    assert np.max(np.abs(mat - r1)) <= 0.1
    r2 = rec.fbp_reconstruction(sino_360, center, angles=np.deg2rad(ang360),
                                apply_log=False, gpu=True)
    assert np.max(np.abs(mat - r2)) <= 0.1
    stack = np.pad(np.expand_dims(sino_180, 1), ((0, 0), (2, 2), (0, 0)),
                   mode="edge")
    rs = rec.fbp_reconstruction(stack, center, apply_log=False, gpu=False,
                                angles=angles2, ratio=0.0, ncore=1)
    assert rs.shape[1] == 5 and np.median(np.abs(mat - rs[:, 0, :])) <= 0.2
    rs = rec.fbp_reconstruction(stack, center, apply_log=False, gpu=True,
                                ratio=0.0, ncore=1)
    assert rs.shape[1] == 5 and np.max(np.abs(mat - rs[:, 0, :])) <= 0.1
    rs = rec.fbp_reconstruction(stack, center, apply_log=False, gpu=True,
                                ratio=1.0, ncore=None)
    assert np.max(np.abs(mat - rs[:, 0, :])) <= 0.1
    with pytest.warns(Warning):
        rec.fbp_reconstruction(sino_180.copy(), center, apply_log=True)
    with pytest.raises(ValueError):
        rec.fbp_reconstruction(sino_180, center, np.deg2rad(ang360),
                               apply_log=False)
    with pytest.raises(ValueError):
        rec.fbp_reconstruction(sino_180, -1, apply_log=False)


@pytest.mark.parametrize("gpu", [True, False])
@pytest.mark.parametrize("ratio", [1.0, None, 0.0])
def test_reference_case_vs_oracle(gpu, ratio):
    from algotom_amd.rec import reconstruction as rec
    from oracle import port
    mat, sino, angles, center = _ref_case()
    got = rec.fbp_reconstruction(sino, center, angles, ratio=ratio,
                                 apply_log=False, gpu=gpu)
    ref = port.fbp(sino, center, angles, ratio=ratio, apply_log=False,
                   gpu=gpu)
    err = np.abs(got.astype(np.float64) - ref).max()
    assert err <= 1e-4 * np.abs(ref).max()
    e64 = np.abs(got - _fp64_oracle(sino, center, angles, False, ratio)).max()
    print("ref case gpu=%s ratio=%s: max|d| vs port %.3g, vs fp64 %.3g, "
          "bit-identical %s" % (gpu, ratio, err, e64,
                                np.array_equal(got, ref)))


_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(name):
    with np.load(os.path.join(_GOLDEN, name), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def _c1_input():
    """The committed CPU workload (tests/golden/c1_fourier_slice.npz): 721
    angles x 512 columns over [0, pi), centre 256.0, a 20-ellipse phantom
    (seed 1) projected by the Fourier-slice method, and its float64 FBP."""
    d = _golden("c1_fourier_slice.npz")
    return d["sinogram"], d["angles"]


def test_reference_case_fixture():
    """tests/golden/reference_case.npz: the reference test's input with the
    oracle's outputs for both accumulations, reproduced bit for bit."""
    from algotom_amd.rec import reconstruction as rec
    d = _golden("reference_case.npz")
    for gpu, key in ((True, "fbp_acc_f64"), (False, "fbp_acc_f32seq")):
        got = rec.fbp_reconstruction(d["sinogram"], float(d["center"]),
                                     d["angles"], apply_log=False, gpu=gpu)
        assert np.array_equal(got, d[key]), key


@pytest.mark.parametrize("gpu", [True, False])
def test_c1_full_size(gpu):
    from algotom_amd.rec import reconstruction as rec
    from oracle import port
    sino, angles = _c1_input()
    got = rec.fbp_reconstruction(sino, 256.0, angles, apply_log=False,
                                 gpu=gpu)
    ref = port.fbp(sino, 256.0, angles, apply_log=False, gpu=gpu)
    err = np.abs(got.astype(np.float64) - ref).max()
    scale = np.abs(ref).max()
    e64 = np.abs(got - _golden("c1_fourier_slice.npz")["fbp_f64"]).max()
    print("C1 gpu=%s: max|d| vs port %.3g (rel %.3g), vs fp64 rel %.3g, "
          "mismatching pixels %d" % (gpu, err, err / scale, e64 / scale,
                                     int((got != ref).sum())))
    assert err <= 1e-4 * scale


def test_stack_with_log_and_noise():
    """4 slices of the 2560-column, 1801-angle stack workload, Poisson
    transmission at 1e4 photons, centre 1280.37, apply_log=True."""
    import torch
    from algotom_amd.rec import reconstruction as rec
    from algotom_amd.util import phantom
    from oracle import port
    ncol, nrow, nslice, center = 2560, 1801, 4, 1280.37
    angles = np.linspace(0.0, np.pi, nrow, endpoint=False)
    slices = []
    for j in range(nslice):
        prm = phantom.random_ellipses(ncol, 20, seed=1000 + j)
        p = phantom.ellipse_sinogram(prm, angles, ncol, np.float32(center),
                                     device="cuda")
        slices.append(phantom.transmission(p, seed=1000 + j).cpu().numpy())
    sino = np.stack(slices, axis=1)
    got = rec.fbp_reconstruction(sino.copy(), center, angles)
    ref = port.fbp(sino, center, angles)
    err = np.abs(got.astype(np.float64) - ref).max()
    scale = np.abs(ref).max()
    print("stack 4x1801x2560: max|d| rel %.3g, mismatching pixels %d of %d"
          % (err / scale, int((got != ref).sum()), got.size))
    assert err <= 1e-4 * scale
    torch.cuda.synchronize()


def test_fbp_width_past_16384_point_filter():
    """A phantom 8100 columns wide (padded row 8400: FFT length 32768, the
    two-half filter path) through fbp_reconstruction against the C port's
    reference arithmetic; bar as for the stack slices, 1e-4 * max|ref|."""
    from algotom_amd.rec import reconstruction as rec
    from algotom_amd.util import phantom
    from oracle import port
    ncol, nrow, center = 8100, 121, 4049.63
    angles = np.linspace(0.0, np.pi, nrow, endpoint=False)
    prm = phantom.random_ellipses(ncol, 20, seed=2024)
    p = phantom.ellipse_sinogram(prm, angles, ncol, np.float32(center),
                                 device="cuda")
    sino = phantom.transmission(p, seed=2024).cpu().numpy()
    got = rec.fbp_reconstruction(sino.copy(), center, angles)
    print("8100 wide: GPU done, oracle port", flush=True)
    ref = port.fbp(sino, center, angles)
    err = np.abs(got.astype(np.float64) - ref).max()
    scale = np.abs(ref).max()
    print("121x8100: max|d| rel %.3g, mismatching pixels %d of %d"
          % (err / scale, int((got != ref).sum()), got.size))
    assert err <= 1e-4 * scale


def test_device_tensor_in_tensor_out():
    import torch
    from algotom_amd.rec import reconstruction as rec
    mat, sino, angles, center = _ref_case()
    t = torch.from_numpy(sino).cuda()
    out = rec.fbp_reconstruction(t, center, angles, apply_log=False)
    assert isinstance(out, torch.Tensor) and out.is_cuda
    ref = rec.fbp_reconstruction(sino, center, angles, apply_log=False)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_fused_c_entry_matches_python_pipeline():
    import torch
    from algotom_amd import _native as nat
    from algotom_amd.rec import reconstruction as rec
    rng = np.random.default_rng(3)
    nrow, nslice, ncol, pad = 91, 3, 150, 22
    sino = rng.uniform(0.1, 1.0, (nrow, nslice, ncol)).astype(np.float32)
    angles = np.linspace(0, np.pi, nrow, endpoint=False)
    center = 74.6
    ref = rec.fbp_reconstruction(sino.copy(), center, angles, pad=pad,
                                 apply_log=True)
    dev = nat.device()
    lib = nat.load()
    taps = torch.from_numpy(rec._taps(rec._ramp_row(ncol + 2 * pad, "hann"))
                            [None, :].copy()).to(dev)
    ws_bytes = lib.ath_fbp_workspace_bytes(nrow, nslice, ncol, pad, 1,
                                           nat.ATH_F32)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    out = torch.empty((ncol, nslice, ncol), dtype=torch.float32, device=dev)
    s = torch.from_numpy(sino).to(dev)
    ang = rec._angles_device(angles, dev)
    nat.check(lib.ath_fbp_reconstruction(
        nat.ptr(s), nat.ATH_F32, nrow, nslice, ncol, nat.ptr(ang),
        float(np.float32(center)), float(1.0 * (ncol // 2)), nat.ptr(taps), 1,
        pad, 0, 1, nat.ACC_F32SEQ, nat.ATH_F32, nat.ptr(ws), ws_bytes,
        nat.ptr(out), nat.stream_ptr(dev)), "fused")
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("mode", ["linear_ramp", "maximum", "mean",
                                  "median", "minimum"])
@pytest.mark.parametrize("gpu", [True, False])
def test_fbp_value_pad_modes_vs_port(mode, gpu):
    """fbp_reconstruction with every numpy.pad value mode (reference
    reconstruction.py:147 hands pad_mode to np.pad) against the C port, on
    a 2D sinogram and a 3-slice stack."""
    from algotom_amd.rec import reconstruction as rec
    from oracle import port
    rng = np.random.default_rng(17)
    for shape, center in (((121, 150), 74.3), ((91, 3, 128), 63.0)):
        sino = rng.uniform(0.1, 1.0, shape).astype(np.float32)
        angles = np.linspace(0.0, np.pi, shape[0], endpoint=False)
        got = rec.fbp_reconstruction(sino.copy(), center, angles,
                                     pad_mode=mode, gpu=gpu)
        ref = port.fbp(sino, center, angles, pad_mode=mode, gpu=gpu)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print("%s %s: rel err vs port %.3g" % (mode, shape, err))
        assert err <= 1e-5


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16"])
def test_reduced_precision_error_vs_fp64(precision):
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _c1_input()
    got = rec.fbp_reconstruction(sino, 256.0, angles, apply_log=False,
                                 precision=precision)
    ref = _golden("c1_fourier_slice.npz")["fbp_f64"]
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    print("precision %s: max rel err vs fp64 %.3g" % (precision, rel))
    bound = {"fp32": 1e-5, "fp16": 1e-4, "bf16": 1e-3}[precision]
    assert rel <= bound


def test_apply_log_mutates_caller_array():
    from algotom_amd.rec import reconstruction as rec
    rng = np.random.default_rng(4)
    sino = rng.uniform(0.2, 1.0, (30, 64)).astype(np.float32)
    sino[3, 5] = 0.0
    sino[7, 9] = -1.0
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        rec.fbp_reconstruction(sino, 32.0, apply_log=True)
    assert any("values <= 0.0" in str(x.message) for x in w)
    assert sino[3, 5] == 1.0 and sino[7, 9] == 1.0


@pytest.mark.parametrize("dtype,shape", [
    ("float32", (1801, 7, 5003)),     # ~252 MB: one full and one partial buffer
    ("float64", (2049, 67, 331)),     # ~363 MB: both buffers reused
    ("float32", (7, 5, 3))])          # below the staging threshold
def test_staged_host_copy(dtype, shape):
    """The staged device->host copy returns exactly what .cpu() does,
    including for a non-contiguous view."""
    import torch
    from algotom_amd import _host
    g = torch.Generator(device="cuda").manual_seed(5)
    t = torch.randn(shape, dtype=getattr(torch, dtype), device="cuda",
                    generator=g)
    got = _host.to_host(t)
    assert got.dtype == np.dtype(dtype) and got.shape == shape
    assert np.array_equal(got, t.cpu().numpy())
    view = t.transpose(0, 1)
    assert np.array_equal(_host.to_host(view), view.cpu().numpy())
