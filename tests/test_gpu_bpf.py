"""BPF on the GPU against the oracle (oracle/bpf_np.py).

* apply_circular_ramp_filter: float64 output within 1e-11 of the oracle's
  complex128 spectral product (FFT lengths 64 .. 32768, every pad mode);
* bpf_reconstruction, one sinogram: float64 output within 1e-10 (the
  back-projection is bit-identical; the filters differ by float64 rounding);
* stacks: float32 output within 2e-6 of max (float32 rounding flips);
* the reference test's thresholds (tests/test_rec/test_reconstruction.py
  :146-202), warnings and errors.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ncol,pad,mode", [
    (5, 2, "edge"), (64, None, "edge"), (65, 7, "constant"),
    (65, 7, "reflect"), (65, 7, "symmetric"), (65, 7, "wrap"),
    (257, None, "edge"), (512, None, "edge"), (1024, None, "edge"),
    (2560, None, "edge"),
    (4096, None, "edge"),   # W = 4396: FFT length 16384, two-level columns
    # FFT length 32768: two-half row transforms, radix-4 two-level columns
    (7968, 150, "reflect"), (8000, None, "edge")])
def test_circular_filter_vs_oracle(ncol, pad, mode):
    from oracle import bpf_np
    from algotom_amd.rec import reconstruction as rec
    rng = np.random.default_rng(ncol)
    img = rng.standard_normal((ncol, ncol)).astype(np.float32)
    p = min(int(0.15 * ncol), 150) if pad is None else pad
    win = rec.make_circular_ramp_window(ncol + 2 * p, "hann")
    got = rec.apply_circular_ramp_filter(img, win, pad=p, pad_mode=mode)
    ref = bpf_np.circular_filter(img, win, p, mode)
    assert got.dtype == np.float64 and got.shape == (ncol, ncol)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("ncol %d pad %d %s: max rel %.2e" % (ncol, p, mode, err))
    assert err <= 1e-11


def _phantom_sino(nrow=181, ncol=96, center=47.5, nslice=1, seed=3):
    import torch
    from algotom_amd.util import phantom
    angles = np.deg2rad(np.linspace(0.0, 180.0, nrow))
    out = []
    for j in range(nslice):
        prm = phantom.random_ellipses(ncol, 12, seed=seed + j)
        p = phantom.ellipse_sinogram(prm, torch.from_numpy(angles), ncol,
                                     center)
        out.append(phantom.transmission(p, photons=1e5, seed=seed + j)
                   .numpy())
    s = np.stack(out, axis=1)
    return (s[:, 0, :] if nslice == 1 else s), angles


@pytest.mark.parametrize("gpu", [True, False])
@pytest.mark.parametrize("ratio", [1.0, None, 0.0])
def test_bpf_2d_vs_oracle(gpu, ratio):
    from oracle import bpf_np
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom_sino()
    got = rec.bpf_reconstruction(sino.copy(), 47.5, angles, ratio=ratio,
                                 gpu=gpu)
    ref = bpf_np.bpf(sino, 47.5, angles, ratio=ratio, gpu=gpu)
    assert got.dtype == np.float64 and got.shape == (96, 96)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("gpu %s ratio %s: max rel %.2e" % (gpu, ratio, err))
    assert err <= 1e-10


@pytest.mark.parametrize("gpu", [True, False])
def test_bpf_stack_vs_oracle(gpu):
    from oracle import bpf_np
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom_sino(nrow=361, ncol=200, center=101.25,
                                 nslice=5)
    got = rec.bpf_reconstruction(sino.copy(), 101.25, angles, gpu=gpu)
    ref = bpf_np.bpf(sino, 101.25, angles, gpu=gpu)
    assert got.dtype == np.float32 and got.shape == (200, 5, 200)
    err = np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max()
    print("stack gpu %s: max rel %.2e, equal %.4f" % (
        gpu, err, np.mean(got == ref)))
    assert err <= 2e-6


def test_bpf_single_slice_stack_is_2d_branch():
    from oracle import bpf_np
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom_sino()
    got = rec.bpf_reconstruction(sino[:, None, :].copy(), 47.5, angles)
    ref = bpf_np.bpf(sino[:, None, :], 47.5, angles)
    assert got.shape == (96, 1, 96) and got.dtype == np.float64
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def test_bpf_c1_geometry_and_tensor_io():
    import torch
    from oracle import bpf_np
    from algotom_amd.rec import reconstruction as rec
    sino, angles = _phantom_sino(nrow=721, ncol=512, center=256.0, seed=1)
    t = torch.from_numpy(sino.copy()).cuda()
    got = rec.bpf_reconstruction(t, 256.0, angles)
    assert isinstance(got, torch.Tensor) and got.is_cuda
    ref = bpf_np.bpf(sino, 256.0, angles)
    err = (np.abs(got.cpu().numpy() - ref).max() / np.abs(ref).max())
    print("721x512: max rel %.2e" % err)
    assert err <= 1e-10


def test_reference_thresholds_and_errors():
    from __graft_entry__ import _reference_test_case
    from algotom_amd.rec import reconstruction as rec
    mat, sino, ang, center = _reference_test_case()
    f = rec.bpf_reconstruction
    r = f(sino.copy(), center, apply_log=False, gpu=False)
    assert np.median(np.abs(mat - r)) <= 0.2
    r = f(sino.copy(), center, apply_log=False, gpu=True)
    assert np.max(np.abs(mat - r)) <= 0.1
    stack = np.pad(np.expand_dims(sino, 1), ((0, 0), (2, 2), (0, 0)),
                   mode="edge")
    rs = f(stack.copy(), center, apply_log=False, gpu=True, ratio=0.0)
    assert rs.shape[1] == 5
    assert np.max(np.abs(mat - rs[:, 0, :])) <= 0.1
    rs = f(stack.copy(), center, np.deg2rad(ang), apply_log=False, gpu=False,
           ratio=0.0)
    assert np.median(np.abs(mat - rs[:, 0, :])) <= 0.2
    with pytest.warns(Warning):
        f(sino.copy(), center, apply_log=True)
    with pytest.raises(ValueError):
        f(sino.copy(), center, np.deg2rad(np.linspace(0, 360, 73)),
          apply_log=False)
    with pytest.raises(ValueError):
        f(sino.copy(), -1, apply_log=False)
