"""The registered operators (algotom_amd/rec/torch_ops.py): schemas and
fake implementations on the CPU; on the GPU, the same results as the
drop-in functions, torch.library.opcheck and a torch.compile trace."""

import numpy as np
import pytest
import torch


def test_ops_are_registered_with_fake_shapes():
    from algotom_amd.rec import torch_ops  # noqa: F401
    ops = torch.ops.algotom_amd
    s = torch.empty((181, 3, 96), device="meta")
    a = torch.empty((181,), device="meta")
    assert ops.fbp(s, a, 47.5, 1.0, "hann", 14, "edge", True,
                   True).shape == (96, 3, 96)
    assert ops.ramp_filter(s, 14, "median", "hann",
                           False).dtype == torch.float32
    assert ops.back_projection(s, a, 47.5, 1, -1.0, True,
                               False).shape == (96, 3, 96)
    img = torch.empty((96, 3, 96), device="meta")
    assert ops.forward_projection(img, a, 47.5, -1.0, 1.0).shape == s.shape


def _case(seed=5, shape=(181, 3, 96)):
    rng = np.random.default_rng(seed)
    sino = rng.uniform(0.1, 1.0, shape).astype(np.float32)
    angles = np.linspace(0.0, np.pi, shape[0], endpoint=False)
    return sino, angles


@pytest.mark.gpu
def test_fbp_op_matches_dropin_and_leaves_input():
    from algotom_amd.rec import reconstruction as rec
    from algotom_amd.rec import torch_ops  # noqa: F401
    sino, angles = _case()
    sino[3, 1, 7] = 0.0                       # <= 0: fixed in a copy only
    s = torch.from_numpy(sino).cuda()
    a = torch.from_numpy(angles).cuda()
    with pytest.warns(UserWarning):
        got = torch.ops.algotom_amd.fbp(s, a, 47.3, 1.0, "hann", 14, "edge",
                                        True, False)
    assert float(s[3, 1, 7]) == 0.0
    ref = rec.fbp_reconstruction(sino.copy(), 47.3, angles, pad=14,
                                 gpu=False)
    assert np.array_equal(got.cpu().numpy(), ref)


@pytest.mark.gpu
def test_opcheck_and_compile():
    from algotom_amd.rec import torch_ops  # noqa: F401
    sino, angles = _case(seed=6, shape=(91, 2, 64))
    s = torch.from_numpy(sino).cuda()
    a = torch.from_numpy(angles).float().cuda()
    ops = torch.ops.algotom_amd
    torch.library.opcheck(ops.fbp.default,
                          (s, a, 31.5, 1.0, "hann", 9, "edge", True, True),
                          test_utils=("test_schema", "test_faketensor"))
    torch.library.opcheck(ops.back_projection.default,
                          (s, a, 31.5, 0, -1.0, True, False),
                          test_utils=("test_schema", "test_faketensor"))

    def pipeline(x, ang):
        filt = ops.ramp_filter(x, 9, "edge", "hann", True)
        return ops.back_projection(filt, ang, 31.5, 1, 32.0, True, False)

    eager = pipeline(s, a)
    compiled = torch.compile(pipeline, backend="aot_eager", fullgraph=True)
    assert torch.equal(compiled(s, a), eager)
    from algotom_amd.rec import reconstruction as rec
    ref = rec.fbp_reconstruction(sino, 31.5, angles, pad=9, gpu=False)
    assert np.array_equal(eager.cpu().numpy(), ref)
    # the adjoint op against the autograd geometry's projector
    from algotom_amd.rec import autograd
    img = torch.from_numpy(np.random.default_rng(7).standard_normal(
        (64, 2, 64)).astype(np.float32)).cuda()
    geom = autograd.FBPGeometry(91, 64, angles, ratio=1.0)
    fp = ops.forward_projection(img, a, 31.5, geom.mask_radius, geom.sigma)
    assert torch.equal(fp, geom.forward_project(img, 31.5))
