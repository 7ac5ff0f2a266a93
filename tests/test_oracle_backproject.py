"""Pins the oracle's back-projection and end-to-end FBP (CPU).

Checks: vectorized oracle == scalar pure-Python loops == C port bit for bit
(three independent codings of reconstruction.py:156-190/229-264/307-359);
closed forms (constant sinogram, single angle at 0 rad is an exact row
shift, integer centre shift moves the image); adjointness against a dense
forward-projection matrix; 180-degree mirror symmetry; and the reference's
own accuracy thresholds (tests/test_rec/test_reconstruction.py:88-144).
"""

import numpy as np
import pytest

from oracle import fbp_np, port


@pytest.mark.parametrize("acc", [0, 1])
@pytest.mark.parametrize("edge_pad", [False, True])
@pytest.mark.parametrize("ncol,nrow,center", [(9, 7, 4.0), (10, 5, 4.8),
                                              (13, 11, 2.0), (8, 3, 7.0)])
def test_three_codings_agree_bitwise(acc, edge_pad, ncol, nrow, center):
    rng = np.random.default_rng(ncol * nrow)
    s = rng.standard_normal((nrow, ncol)).astype(np.float32)
    ang = np.sort(rng.uniform(0, 2 * np.pi, nrow))
    a = fbp_np.backproject_loops(s, ang, center, edge_pad, acc)
    b = fbp_np.backproject(s, ang, center, edge_pad, acc)
    c = port.backproject(s, ang, center, edge_pad, acc)
    assert np.array_equal(a, b) and np.array_equal(b, c)


def test_port_stack_equals_per_slice():
    rng = np.random.default_rng(5)
    s = rng.standard_normal((20, 3, 31)).astype(np.float32)
    ang = np.linspace(0, np.pi, 20, endpoint=False)
    st = port.backproject(s, ang, 15.3, acc=1)
    for j in range(3):
        assert np.array_equal(st[:, j, :],
                              port.backproject(s[:, j, :], ang, 15.3, acc=1))


def test_accumulation_modes_differ_only_by_rounding():
    rng = np.random.default_rng(6)
    s = rng.standard_normal((400, 64)).astype(np.float32)
    ang = np.linspace(0, np.pi, 400, endpoint=False)
    a = port.backproject(s, ang, 31.5, acc=0)
    b = port.backproject(s, ang, 31.5, acc=1)
    assert not np.array_equal(a, b)
    assert np.abs(a - b).max() <= 1e-5 * np.abs(a).max()


def test_constant_sinogram():
    ncol, nrow, v = 40, 25, 0.75
    s = np.full((nrow, ncol), v, dtype=np.float32)
    ang = np.linspace(0, np.pi, nrow, endpoint=False)
    out = port.backproject(s, ang, (ncol - 1) / 2, acc=0)
    yy, xx = np.mgrid[:ncol, :ncol] - (ncol - 1) / 2
    inside = xx ** 2 + yy ** 2 <= ((ncol - 1) / 2 - 1) ** 2
    assert np.all(out[inside] == np.float32(nrow * v))


def test_single_angle_zero_is_row_shift():
    # theta = 0: f = x - ic + center, the same for every row y
    rng = np.random.default_rng(7)
    ncol = 30
    s = rng.standard_normal((1, ncol)).astype(np.float32)
    center = 14.5 + 0.25  # f = x + 0.25
    out = port.backproject(s, [0.0], center, acc=0)
    xs = np.arange(ncol) + 0.25
    ref = np.interp(xs, np.arange(ncol), s[0].astype(np.float64),
                    right=np.nan)
    ok = ~np.isnan(ref)
    assert np.abs(out[0][ok] - ref[ok]).max() < 1e-6
    assert np.all(out[0][~ok] == 0.0)
    assert np.all(out == out[0:1])


def test_integer_centre_shift_moves_image():
    # at theta = 0 the detector axis is x: centre + k shifts columns by -k
    rng = np.random.default_rng(8)
    ncol = 24
    s = rng.standard_normal((1, ncol)).astype(np.float32)
    a = port.backproject(s, [0.0], 10.0, acc=0)
    b = port.backproject(s, [0.0], 12.0, acc=0)
    assert np.array_equal(a[:, 2:], b[:, :-2])


def test_adjointness_dense_matrix():
    rng = np.random.default_rng(9)
    ncol, nrow = 11, 9
    ang = np.sort(rng.uniform(0, np.pi, nrow))
    center = 5.3
    s = rng.standard_normal((nrow, ncol)).astype(np.float32)
    x = rng.standard_normal((ncol, ncol))
    a64 = fbp_np.forward_project_matrix(ncol, ang, center, False)
    bp = fbp_np.backproject_f64(s, ang, np.float32(center))
    lhs = np.sum(bp * x)
    rhs = np.sum(s.astype(np.float64).ravel() * (a64 @ x.ravel()))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1)
    a = fbp_np.forward_project_matrix(ncol, ang, center)
    # the float32-typed oracle equals A^T s up to float32 rounding of trig
    bt = (a.T @ s.astype(np.float64).ravel()).reshape(ncol, ncol)
    b32 = fbp_np.backproject(s, ang, center, acc=0)
    assert np.abs(b32 - bt).max() <= 1e-5 * np.abs(bt).max()


@pytest.mark.parametrize("arc,center,mask", [(np.pi, 6.3, -1.0),
                                             (2 * np.pi / 3, 5.0, 6.0),
                                             (2 * np.pi, 7.9, 4.5)])
def test_port_forward_projection_is_dense_transpose(arc, center, mask):
    # the port's scatter forward projection against the transpose of the
    # back-projection matrix assembled column by column from the port's
    # float64-sum back-projection of unit sinograms
    rng = np.random.default_rng(12)
    ncol, nrow = 13, 7
    ang = np.linspace(0.0, arc, nrow)
    cols = []
    for i in range(nrow):
        for d in range(ncol):
            e = np.zeros((nrow, ncol), np.float32)
            e[i, d] = 1.0
            cols.append(port.backproject(e, ang, center,
                                         acc=fbp_np.ACC_F64).ravel())
    b = np.array(cols, dtype=np.float64).T          # (ncol^2, nrow*ncol)
    x = rng.standard_normal((ncol, ncol)).astype(np.float32)
    xm = x.astype(np.float64)
    if mask >= 0:
        yy, xx = np.mgrid[:ncol, :ncol] - ncol // 2
        xm = xm * ((xx * xx + yy * yy) <= mask * mask)
    exp = (b.T @ xm.ravel()).reshape(nrow, ncol) * 0.5
    got = port.forward_project(x, ang, center, mask_radius=mask, scale=0.5)
    assert np.abs(got - exp).max() <= 1e-6 * np.abs(exp).max()


def test_mirror_symmetry_180():
    # projecting at theta + pi reverses the detector axis about the centre
    rng = np.random.default_rng(10)
    ncol = 33
    c = (ncol - 1) / 2
    s = rng.standard_normal((1, ncol)).astype(np.float64)
    a = fbp_np.backproject_f64(s, [0.3], c)
    b = fbp_np.backproject_f64(s[:, ::-1], [0.3 + np.pi], c)
    assert np.abs(a - b).max() < 1e-9


def _reference_case():
    from __graft_entry__ import _reference_test_case
    return _reference_test_case()


def test_reference_thresholds_on_oracle():
    """tests/test_rec/test_reconstruction.py:88-144 on the oracle."""
    mat, sino, ang, center = _reference_case()
    r = port.fbp(sino, center, apply_log=False, gpu=False)
    assert np.median(np.abs(mat - r)) <= 0.2
    r = port.fbp(sino, center, apply_log=False, gpu=True)
    assert np.max(np.abs(mat - r)) <= 0.1
    stack = np.pad(np.expand_dims(sino, 1), ((0, 0), (2, 2), (0, 0)),
                   mode="edge")
    rs = port.fbp(stack, center, ang, apply_log=False, gpu=False, ratio=0.0)
    assert rs.shape[1] == 5
    assert np.median(np.abs(mat - rs[:, 0, :])) <= 0.2


def test_reference_case_fixture_is_current():
    """tests/golden/reference_case.npz holds the reference test's input
    and what the oracle makes of it now (read by the GPU tests)."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "golden", "reference_case.npz")
    with np.load(path, allow_pickle=False) as d:
        mat, sino, ang, center = _reference_case()
        assert np.array_equal(d["sinogram"], sino)
        assert np.array_equal(d["phantom"], mat)
        assert np.array_equal(d["angles"], ang)
        assert float(d["center"]) == center
        assert np.array_equal(d["fbp_acc_f64"],
                              port.fbp(sino, center, ang, apply_log=False,
                                       gpu=True))
        assert np.array_equal(d["fbp_acc_f32seq"],
                              port.fbp(sino, center, ang, apply_log=False,
                                       gpu=False))


def test_numpy_and_port_fbp_agree():
    mat, sino, ang, center = _reference_case()
    for gpu in (True, False):
        a = fbp_np.fbp(sino, center, ang, apply_log=False, gpu=gpu)
        b = port.fbp(sino, center, ang, apply_log=False, gpu=gpu)
        assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()


def test_circle_mask_closed_form():
    for width in (1, 2, 7, 64, 65):
        for ratio in (0.0, 0.5, 1.0, 1.3):
            m = fbp_np.circle_mask(width, ratio)
            c = width // 2
            r = ratio * c
            for y in range(width):
                for x in range(width):
                    want = (x - c) ** 2 + (y - c) ** 2 <= r * r
                    assert m[y, x] == (1.0 if want else 0.0)
