"""Host-side pieces of the transfer pipeline that need no GPU: chunk
planning, the streaming threshold and the threaded block copy."""

import numpy as np
import pytest

from algotom_amd import _host
from algotom_amd.rec import reconstruction as rec


@pytest.mark.parametrize("nslice,slice_bytes", [
    (64, 4 * (1801 * 2560 + 2560 * 2560)),   # the 64-slice stack: 8 x 8
    (11, 1 << 30),                           # one slice per chunk
    (7, 1),                                  # everything in one chunk
    (130, 3 << 20)])
def test_chunk_ranges_cover_every_slice_once(nslice, slice_bytes):
    ranges = rec._chunk_ranges(nslice, slice_bytes)
    assert ranges[0][0] == 0 and ranges[-1][1] == nslice
    for (a0, a1), (b0, b1) in zip(ranges, ranges[1:]):
        assert a1 == b0 and a1 > a0
    sizes = [b - a for a, b in ranges]
    assert all(s == sizes[0] for s in sizes[:-1]) and sizes[-1] <= sizes[0]
    if sizes[0] >= 8:
        assert sizes[0] % 4 == 0
    assert sizes[0] * slice_bytes <= max(rec._CHUNK_BYTES, slice_bytes)


def test_stack_chunks_are_eight_slices():
    assert rec._chunk_ranges(64, 4 * (1801 * 2560 + 2560 * 2560))[0] == (0, 8)


def test_streaming_threshold():
    small = np.zeros((10, 3, 20), np.float32)
    assert not rec._streams(small, 20)
    big = np.lib.stride_tricks.as_strided(np.zeros(1, np.float32),
                                          (1801, 64, 2560), (0, 0, 0))
    assert rec._streams(big, 2560)
    one = np.lib.stride_tricks.as_strided(np.zeros(1, np.float32),
                                          (1801, 1, 8192), (0, 0, 0))
    assert not rec._streams(one, 8192)      # a single slice never streams


@pytest.mark.parametrize("shape,src_dtype,dst_dtype", [
    ((37, 20, 50), np.uint16, np.float32),   # rows >= workers, cast
    ((3, 700, 900), np.float32, np.float32), # rows < workers: split axis 1
    ((5000,), np.float64, np.float64),       # 1D
    ((2, 3), np.float32, np.float32)])       # below the threading size
def test_threaded_copy_strided_and_cast(shape, src_dtype, dst_dtype):
    rng = np.random.default_rng(len(shape))
    big = (rng.random(tuple(2 * n for n in shape)) * 1000).astype(src_dtype)
    src = big[tuple(slice(None, None, 2) for _ in shape)]
    dst_big = np.zeros(tuple(n + 1 for n in shape), dst_dtype)
    dst = dst_big[tuple(slice(1, None) for _ in shape)]
    _host._copy(dst, src)
    assert np.array_equal(dst, src.astype(dst_dtype))
    assert not dst_big[(0,) * len(shape)]


def test_to_host_of_host_tensor():
    import torch
    t = torch.arange(12, dtype=torch.float32).reshape(3, 4)
    out = _host.to_host(t)
    assert isinstance(out, np.ndarray) and np.array_equal(out, t.numpy())
