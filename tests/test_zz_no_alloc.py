"""ath_fbp_reconstruction_noalloc (the allocation-free fused FBP) against
ath_fbp_reconstruction with a caller workspace, through the C ABI.

Two calls in a row share the library's device array; the second is smaller,
so stale bytes of the first would show up if the entry read any byte it did
not write.  Bar: bit-identical (same kernels, same arithmetic).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _taps(width):
    from algotom_amd.rec import reconstruction as rec
    return rec._taps(rec._ramp_row(width, "hann"))[None, :].copy()


def _fused(lib, nat, dev, sino, angles, center, pad, acc, filt_dtype,
           noalloc):
    import torch
    from algotom_amd.rec import reconstruction as rec
    nrow, nslice, ncol = sino.shape
    taps = torch.from_numpy(_taps(ncol + 2 * pad)).to(dev)
    out = torch.full((ncol, nslice, ncol), float("nan"), dtype=torch.float32,
                     device=dev)
    s = torch.from_numpy(sino).to(dev)
    ang = rec._angles_device(angles, dev)
    args = (nat.ptr(s), nat.ATH_F32, nrow, nslice, ncol, nat.ptr(ang),
            float(np.float32(center)), float(1.0 * (ncol // 2)),
            nat.ptr(taps), 1, pad, 0, 1, acc, filt_dtype)
    if noalloc:
        nat.check(lib.ath_fbp_reconstruction_noalloc(
            *args, nat.ptr(out), nat.stream_ptr(dev)), "noalloc")
    else:
        ws_bytes = lib.ath_fbp_workspace_bytes(nrow, nslice, ncol, pad, 1,
                                               filt_dtype)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        nat.check(lib.ath_fbp_reconstruction(
            *args, nat.ptr(ws), ws_bytes, nat.ptr(out),
            nat.stream_ptr(dev)), "fused")
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("acc,filt", [("f32seq", "f32"), ("f32", "bf16")])
def test_noalloc_matches_workspace_entry_twice(acc, filt):
    from algotom_amd import _native as nat
    lib = nat.load()
    dev = nat.device()
    acc = {"f32seq": nat.ACC_F32SEQ, "f32": nat.ACC_F32}[acc]
    filt = {"f32": nat.ATH_F32, "bf16": nat.ATH_BF16}[filt]
    rng = np.random.default_rng(11)
    # a 1801 x 2 x 1024 stack (~15 MiB filtered), then a smaller 2D case
    cases = [((1801, 2, 1024), 511.7, 150), ((181, 1, 257), 130.2, 38)]
    for shape, center, pad in cases:
        nrow, nslice, ncol = shape
        need = lib.ath_fbp_workspace_bytes(nrow, nslice, ncol, pad, 1, filt)
        assert 0 < need <= nat.FBP_NOALLOC_MAX_BYTES
        sino = rng.uniform(0.1, 1.0, shape).astype(np.float32)
        angles = np.linspace(0.0, np.pi, nrow, endpoint=False)
        ref = _fused(lib, nat, dev, sino, angles, center, pad, acc, filt,
                     False)
        got = _fused(lib, nat, dev, sino, angles, center, pad, acc, filt,
                     True)
        print("noalloc %s: max|diff| %g" % (shape, np.abs(got - ref).max()))
        assert np.isfinite(ref).all()
        assert np.array_equal(got, ref)

