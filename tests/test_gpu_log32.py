"""Device build of the float32 -log (csrc/log32.h) through ath_minus_log:
for every positive finite float32 it must equal (float)(-log((double)x))
with the device's float64 library log (torch's float64 log), which is what
the kernels computed before the table-driven form."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_minus_log_exhaustive():
    import torch
    from algotom_amd import _native as nat
    lib = nat.load()
    dev = torch.device("cuda", 0)
    end = 0x7F800000
    step = 1 << 28
    out = torch.empty(step, dtype=torch.float32, device=dev)
    bad = 0
    for lo in range(1, end, step):
        hi = min(end, lo + step)
        bits = torch.arange(lo, hi, dtype=torch.int64, device=dev).to(
            torch.int32)
        x = bits.view(torch.float32)
        o = out[:hi - lo]
        nat.check(lib.ath_minus_log(nat.ptr(x), nat.ptr(o), hi - lo,
                                    nat.stream_ptr(dev)), "ath_minus_log")
        ref = (-torch.log(x.double())).float()
        bad += int((o.view(torch.int32) != ref.view(torch.int32)).sum())
        del bits, x, ref
    torch.cuda.synchronize()
    assert bad == 0


def test_minus_log_special_values():
    import torch
    from algotom_amd import _native as nat
    lib = nat.load()
    dev = torch.device("cuda", 0)
    x = torch.tensor([1.0, 0.0, -0.0, -1.0, float("inf"), float("nan"),
                      1e-45, 3e38, 1.0 - 2.0 ** -24, 1.0 + 2.0 ** -23],
                     dtype=torch.float32, device=dev)
    o = torch.empty_like(x)
    nat.check(lib.ath_minus_log(nat.ptr(x), nat.ptr(o), x.numel(),
                                nat.stream_ptr(dev)), "ath_minus_log")
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = (-np.log(x.cpu().numpy().astype(np.float64))).astype(np.float32)
    got = o.cpu().numpy()
    assert np.array_equal(got, ref, equal_nan=True)
    assert np.signbit(got[0])   # -log(1) = -0.0


@pytest.mark.parametrize("dtype,width,fd", [
    ("float32", 64, True), ("float32", 61, True), ("uint16", 64, True),
    ("uint16", 37, False), ("float32", 48, False)])
def test_flat_field_log_bitwise(dtype, width, fd):
    """ath_flat_field_log against numpy on a strided input of 13
    projections (runs of 8 with a partial last one): (raw - dark) / flat_dark
    in float32, values <= 0 set to 1 and flagged, -log float64-rounded."""
    import torch
    from algotom_amd import _native as nat
    lib = nat.load()
    dev = torch.device("cuda", 0)
    nproj, height = 13, 9
    rng = np.random.default_rng(width)
    big = (rng.random((nproj, height + 2, width)) * 3000 + 50).astype(dtype)
    big[4, 2, 5] = 0
    raw = big[:, 1:height + 1, :]                   # proj stride > plane
    dark = (rng.random((height, width)) * 40).astype(np.float32)
    flat_dark = (rng.random((height, width)) * 2000 + 500).astype(np.float32)
    v = raw.astype(np.float32)
    if fd:
        v = (v - dark) / flat_dark
    ref = v.copy()
    ref[ref <= 0] = 1.0
    ref = (-np.log(ref.astype(np.float64))).astype(np.float32)
    t_big = torch.from_numpy(big).to(dev)
    t_raw = t_big[:, 1:height + 1, :]
    code = nat.ATH_U16 if dtype == "uint16" else nat.ATH_F32
    out = torch.empty((nproj, height, width), dtype=torch.float32, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    d_t = torch.from_numpy(dark).to(dev)
    f_t = torch.from_numpy(flat_dark).to(dev)
    st0, st1, _ = t_raw.stride()
    nat.check(lib.ath_flat_field_log(
        nat.ptr(t_raw), code, nproj, height, width, st0, st1,
        nat.ptr(d_t) if fd else None, nat.ptr(f_t) if fd else None, 1,
        nat.ptr(out), nat.ptr(flag), nat.stream_ptr(dev)),
        "ath_flat_field_log")
    got = out.cpu().numpy()
    assert np.array_equal(got, ref)
    assert int(flag.item()) == int(bool((v <= 0).any()))
