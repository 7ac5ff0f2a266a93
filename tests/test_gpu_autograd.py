"""Differentiable FBP: the hand-written adjoint against FP64 references.

* forward projection == dense transpose of the back-projection matrix
  (oracle.fbp_np.forward_project_matrix, float64), and <B s, x> = <s, B^T x>;
* filter adjoint == dense transpose of the filter matrix (built column by
  column with the oracle's float64 filter), every pad mode;
* dL/dsinogram == the dense-matrix gradient F^T B^T (sigma mask dL/dy);
* dL/dcenter == the FP64 central finite difference of the FP64 oracle loss
  on a 128-angle x 64-column case (h = 1e-6), and the derivative image
  matches FP64 finite differences of the back-projection away from kinks;
* both checkpoint modes give identical gradients;
* every check runs on a full 180-degree scan and on the limited 120-degree
  arc of the refinement workload (angles linspace(0, 2 pi / 3, n)).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


ARCS = [np.pi, 2 * np.pi / 3]


def _setup(nrow=128, ncol=64, nslice=1, center=None, seed=0, arc=np.pi):
    rng = np.random.default_rng(seed)
    # a full scan over [0, pi); the limited arc includes its end angle
    angles = np.linspace(0.0, arc, nrow, endpoint=arc != np.pi)
    center = (ncol - 1) / 2 + 0.37 if center is None else center
    sino = rng.standard_normal((nrow, nslice, ncol)).astype(np.float32)
    return rng, angles, center, sino


@pytest.mark.parametrize("arc", ARCS)
def test_forward_projection_is_dense_transpose(arc):
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import autograd
    rng, angles, center, _ = _setup(nrow=40, ncol=33, arc=arc)
    geom = autograd.FBPGeometry(40, 33, angles, ratio=None)
    img = rng.standard_normal((33, 2, 33)).astype(np.float32)
    got = geom.forward_project(torch.from_numpy(img).cuda(), center)
    got = got.cpu().numpy().astype(np.float64) / geom.sigma
    a = fbp_np.forward_project_matrix(33, angles, center)
    for j in range(2):
        ref = (a @ img[:, j, :].astype(np.float64).ravel()).reshape(40, 33)
        assert np.abs(got[:, j, :] - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("nrow,nslice,ncol,ratio", [
    (37, 1, 65, None), (90, 3, 120, 1.0), (181, 9, 200, 0.8),
    (64, 16, 257, 1.0), (45, 20, 130, None)])
def test_forward_projection_ws_matches_plain(nrow, nslice, ncol, ratio):
    # the transposed-copy entry stages the same values into the same LDS
    # records: bit-identical to the plain entry, for every slice batch (1,
    # 4, 8, 16) and both walking directions
    import torch
    from algotom_amd import _native as nat
    lib = nat.load()
    dev = nat.device()
    rng = np.random.default_rng(nrow * ncol)
    angles = np.sort(rng.uniform(0.0, 2 * np.pi, nrow))
    ang = torch.from_numpy(angles.astype(np.float32)).to(dev)
    img = torch.from_numpy(rng.standard_normal((ncol, nslice, ncol))
                           .astype(np.float32)).to(dev)
    mr = -1.0 if ratio is None else float(ratio * (ncol // 2))
    center = float(np.float32((ncol - 1) / 2 + 1.3))
    plain = torch.empty((nrow, nslice, ncol), dtype=torch.float32,
                        device=dev)
    ws = torch.empty_like(plain)
    nat.check(lib.ath_forward_projection(
        nat.ptr(img), nrow, nslice, ncol, nat.ptr(ang), center, mr, 0.5,
        nat.ptr(plain), nat.stream_ptr(dev)), "ath_forward_projection")
    nbytes = lib.ath_forward_projection_workspace_bytes(nslice, ncol)
    assert nbytes == ncol * nslice * ncol * 4
    work = torch.full((nbytes,), 255, dtype=torch.uint8, device=dev)
    nat.check(lib.ath_forward_projection_ws(
        nat.ptr(img), nrow, nslice, ncol, nat.ptr(ang), center, mr, 0.5,
        nat.ptr(work), nbytes, nat.ptr(ws), nat.stream_ptr(dev)),
        "ath_forward_projection_ws")
    assert torch.equal(plain, ws)
    # the workspace holds the transposed image
    t = work.view(torch.float32).view(ncol, nslice, ncol)
    assert torch.equal(t, img.permute(2, 1, 0))


@pytest.mark.parametrize("arc", ARCS)
def test_forward_projection_adjoint_identity_with_mask(arc):
    import torch
    from algotom_amd.rec import autograd
    rng, angles, center, sino = _setup(nrow=90, ncol=120, nslice=3, arc=arc)
    geom = autograd.FBPGeometry(90, 120, angles, ratio=1.0)
    s = torch.from_numpy(sino).cuda()
    x = torch.from_numpy(rng.standard_normal((120, 3, 120))
                         .astype(np.float32)).cuda()
    bs = geom.backproject(s, center)          # sigma * mask * B s
    fx = geom.forward_project(x, center)      # sigma * B^T(mask * x)
    lhs = torch.sum(bs.double() * x.double()).item()
    rhs = torch.sum(s.double() * fx.double()).item()
    assert abs(lhs - rhs) <= 2e-6 * (abs(lhs) + abs(rhs))


def _filter_matrix(ncol, pad, mode, name="hann"):
    from oracle import fbp_np
    r = fbp_np.ramp_window_row(ncol + 2 * pad, name)
    eye = np.eye(ncol)
    return fbp_np.ramp_filter(eye, r, None, pad, mode).T  # columns = F e_k


@pytest.mark.parametrize("mode", ["edge", "constant", "reflect", "symmetric",
                                  "wrap"])
def test_filter_adjoint_is_dense_transpose(mode):
    import torch
    from algotom_amd.rec import autograd
    rng = np.random.default_rng(1)
    nrow, ncol, pad = 7, 50, 12
    geom = autograd.FBPGeometry(nrow, ncol, pad=pad, pad_mode=mode,
                                ratio=None)
    u = rng.standard_normal((nrow, 2, ncol)).astype(np.float32)
    got = geom.filter_adjoint(torch.from_numpy(u).cuda()).cpu().numpy()
    fmat = _filter_matrix(ncol, pad, mode)
    ref = np.einsum("kc,rjk->rjc", fmat, u.astype(np.float64))
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("mode", ["edge", "constant", "reflect", "symmetric",
                                  "wrap"])
def test_filter_adjoint_fft_32768(mode):
    """8000 columns, pad 150: the adjoint needs FFT length 32768 (the
    two-half path with the folded outputs' values stashed in LDS).  Oracle:
    the transpose of F = crop . C_g . pad in float64 -- u embedded at
    [pad, pad + ncol) of a zero row, circular cross-correlation with the
    taps g (numpy FFT), then folded through numpy.pad's index map."""
    import torch
    from algotom_amd.rec import autograd
    rng = np.random.default_rng(8)
    nrow, ncol, pad = 3, 8000, 150
    geom = autograd.FBPGeometry(nrow, ncol, pad=pad, pad_mode=mode,
                                ratio=None)
    assert geom.fft_len_adj == 32768
    u = rng.standard_normal((nrow, 2, ncol)).astype(np.float32)
    got = geom.filter_adjoint(torch.from_numpy(u).cuda()).cpu().numpy()
    W = ncol + 2 * pad
    g = np.asarray(geom.taps, dtype=np.float64).reshape(-1, W)[0]
    if mode == "constant":
        src = np.pad(np.arange(ncol), pad, mode="constant",
                     constant_values=-1)
    else:
        src = np.pad(np.arange(ncol), pad, mode=mode)
    ghat = np.conj(np.fft.fft(g))
    for r in range(nrow):
        for j in range(2):
            e = np.zeros(W)
            e[pad:pad + ncol] = u[r, j].astype(np.float64)
            z = np.real(np.fft.ifft(np.fft.fft(e) * ghat))
            ref = np.zeros(ncol)
            keep = src >= 0
            np.add.at(ref, src[keep], z[keep])
            err = np.abs(got[r, j] - ref).max() / np.abs(ref).max()
            assert err <= 1e-6, (mode, r, j, err)


def _fp64_loss(sino, center, angles, target, ratio=1.0):
    """float64 loss sum((FBP(s; c) - target)^2) with the oracle (centre not
    rounded, so finite differences below float32 resolution work)."""
    from oracle import port
    y = port.fbp_f64(sino, center, angles, ratio=ratio, apply_log=False,
                     round_center=False)
    return float(np.sum((y - target) ** 2)), y


@pytest.mark.parametrize("arc", ARCS)
@pytest.mark.parametrize("checkpoint", ["filtered", "input", "derivative"])
def test_gradients_vs_dense_and_finite_differences(checkpoint, arc):
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import autograd
    rng, angles, center, sino = _setup(nrow=128, ncol=64, arc=arc)
    sino2 = sino[:, 0, :]
    target = rng.standard_normal((64, 64))
    geom = autograd.FBPGeometry(128, 64, angles, ratio=1.0,
                                checkpoint=checkpoint)
    s = torch.from_numpy(sino2.copy()).cuda().requires_grad_(True)
    c = torch.tensor(float(np.float32(center)), dtype=torch.float64,
                     requires_grad=True)
    y = autograd.fbp_differentiable(s, c, geom)
    tgt = torch.from_numpy(target).cuda()
    loss = torch.sum((y.double() - tgt) ** 2)
    loss.backward()
    g_s = s.grad.cpu().numpy().astype(np.float64)
    g_c = c.grad.item()

    # dense-matrix reference gradient in float64
    cf = float(np.float32(center))
    _, yref = _fp64_loss(sino2, cf, angles, target)
    resid = 2.0 * (yref - target)
    mask = fbp_np.circle_mask(64, 1.0).astype(np.float64)
    a = fbp_np.forward_project_matrix(64, angles, cf, float32_trig=True)
    fmat = _filter_matrix(64, min(int(0.15 * 64), 150), "edge")
    bt = (a @ (np.pi / 127 * mask * resid).ravel()).reshape(128, 64)
    g_ref = bt @ fmat  # F^T applied to every row
    assert np.abs(g_s - g_ref).max() <= 1e-4 * np.abs(g_ref).max()

    # FP64 central finite difference in the centre
    h = 1e-6
    lp, _ = _fp64_loss(sino2, cf + h, angles, target)
    lm, _ = _fp64_loss(sino2, cf - h, angles, target)
    fd = (lp - lm) / (2 * h)
    print("dL/dc analytic %.8g  fp64 FD %.8g" % (g_c, fd))
    assert abs(g_c - fd) <= 1e-3 * abs(fd) + 1e-6


@pytest.mark.parametrize("arc", ARCS)
def test_center_derivative_image_vs_fd(arc):
    import torch
    from oracle import fbp_np
    from algotom_amd.rec import autograd
    rng, angles, center, sino = _setup(nrow=60, ncol=48, seed=5, arc=arc)
    geom = autograd.FBPGeometry(60, 48, angles, ratio=None)
    filt = geom.filter(torch.from_numpy(sino).cuda())
    d_img = geom.backproject(filt, center, acc=3).cpu().numpy()[:, 0, :]
    f64 = filt.cpu().numpy()[:, 0, :].astype(np.float64)
    cf = float(np.float32(center))
    h = 1e-7
    ang64 = np.asarray(angles, np.float32).astype(np.float64)
    bp = lambda cc: fbp_np.backproject_f64(f64, ang64, cc)
    fd = (bp(cf + h) - bp(cf - h)) / (2 * h) * geom.sigma
    # pixels whose f crosses an integer within +-h have a one-sided slope;
    # all others must agree to float32 rounding
    close = np.abs(d_img - fd) <= 1e-4 * np.abs(fd).max()
    assert close.mean() > 0.999


@pytest.mark.parametrize("shape,center,mask", [
    ((91, 3, 64), 31.7, 32.0), ((181, 1, 97), 40.2, -1.0),
    ((60, 5, 130), 70.9, 65.0)])
def test_fused_derivative_pass_is_bitwise_the_two_passes(shape, center,
                                                          mask):
    """ath_back_projection_with_derivative: both outputs equal the separate
    ATH_ACC_F64 and ATH_ACC_DERIV calls bit for bit."""
    import torch
    from algotom_amd import _native as nat
    from algotom_amd.rec import reconstruction as rec
    nrow, nslice, ncol = shape
    rng = np.random.default_rng(nrow)
    filt = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)
                            ).cuda()
    angles = np.linspace(0.0, np.pi, nrow, endpoint=False)
    ang = rec._angles_device(angles, filt.device)
    out = torch.empty((ncol, nslice, ncol), dtype=torch.float32,
                      device=filt.device)
    der = torch.empty_like(out)
    nat.check(nat.load().ath_back_projection_with_derivative(
        nat.ptr(filt), nrow, nslice, ncol, nat.ptr(ang),
        float(np.float32(center)), 0, mask, 1, nat.ptr(out), nat.ptr(der),
        nat.stream_ptr(filt.device)), "with_derivative")
    ref = rec._backproject_stack(filt, angles, center, nat.ACC_F64,
                                 mask_radius=mask, apply_scale=True)
    dref = rec._backproject_stack(filt, angles, center, nat.ACC_DERIV,
                                  mask_radius=mask, apply_scale=True)
    assert torch.equal(out, ref)
    assert torch.equal(der, dref)


def test_derivative_checkpoint_matches_filtered_gradients():
    """checkpoint='derivative' gives the same gradients as 'filtered',
    and runs without a centre gradient too."""
    import torch
    from algotom_amd.rec import autograd
    rng, angles, center, sino = _setup(nrow=128, ncol=64, arc=ARCS[0])
    grads = {}
    for ck in ("filtered", "derivative"):
        geom = autograd.FBPGeometry(128, 64, angles, ratio=1.0,
                                    checkpoint=ck)
        s = torch.from_numpy(sino.copy()).cuda().requires_grad_(True)
        c = torch.tensor(float(center), dtype=torch.float64,
                         requires_grad=True)
        y = autograd.fbp_differentiable(s, c, geom)
        torch.sum(y * y).backward()
        grads[ck] = (s.grad.clone(), c.grad.clone())
        s2 = torch.from_numpy(sino.copy()).cuda().requires_grad_(True)
        y2 = autograd.fbp_differentiable(s2, float(center), geom)
        torch.sum(y2 * y2).backward()
        assert torch.equal(s2.grad, grads[ck][0])
    assert torch.equal(grads["filtered"][0], grads["derivative"][0])
    assert torch.equal(grads["filtered"][1], grads["derivative"][1])
