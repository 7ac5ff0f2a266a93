"""Host-array inputs streamed through the chunk pipeline (uploads and
downloads of neighbouring slice chunks overlapped with the computation).

Every streamed result must equal, bit for bit, the same call on the whole
stack already on the device: the slices are independent, so the chunking
changes nothing in the arithmetic.  Thresholds are lowered so that small
stacks stream in several chunks, the last one partial.
"""

import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def small_chunks(monkeypatch):
    from algotom_amd.rec import reconstruction as rec
    monkeypatch.setattr(rec, "_STREAM_MIN", 1)
    monkeypatch.setattr(rec, "_CHUNK_BYTES", 3 << 19)
    return rec


def _stack(nrow, nslice, ncol, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.2, 1.0, (nrow, nslice, ncol)).astype(dtype)


def _device(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("method", ["fbp", "bpf", "dfi"])
def test_streamed_equals_device(small_chunks, method):
    rec = small_chunks
    fn = getattr(rec, method + "_reconstruction")
    sino = _stack(181, 11, 256, 1)
    out_item = 8 if method == "dfi" else 4
    ranges = rec._chunk_ranges(11, 4 * 181 * 256 + out_item * 256 * 256)
    assert len(ranges) >= 3 and ranges[-1][1] - ranges[-1][0] < ranges[0][1]
    got = fn(sino, 127.3, apply_log=True)
    want = fn(_device(sino), 127.3, apply_log=True).cpu().numpy()
    assert isinstance(got, np.ndarray) and got.dtype == want.dtype
    assert np.array_equal(got, want)


def test_streamed_fbp_input_kinds(small_chunks, tmp_path):
    """A strided view, float64, uint16 and a memory map."""
    rec = small_chunks
    base = _stack(121, 14, 200, 2)
    want = lambda x: rec.fbp_reconstruction(  # noqa: E731
        _device(x), 99.5, apply_log=False).cpu().numpy()
    view = base[:, ::2, :]
    assert np.array_equal(rec.fbp_reconstruction(view, 99.5, apply_log=False),
                          want(view))
    f64 = base.astype(np.float64)
    assert np.array_equal(rec.fbp_reconstruction(f64, 99.5, apply_log=False),
                          want(f64))
    u16 = (base * 1000).astype(np.uint16)
    assert np.array_equal(rec.fbp_reconstruction(u16, 99.5, apply_log=False),
                          want(u16.astype(np.float32)))
    mm = np.lib.format.open_memmap(str(tmp_path / "s.npy"), mode="w+",
                                   dtype=np.float32, shape=base.shape)
    mm[...] = base
    assert np.array_equal(rec.fbp_reconstruction(mm, 99.5, apply_log=False),
                          want(base))


def test_streamed_log_fixes_caller_array(small_chunks):
    rec = small_chunks
    sino = _stack(121, 9, 160, 3)
    sino[5, 7, 30] = 0.0
    sino[9, 0, 3] = -2.0
    fixed = sino.copy()
    fixed[fixed <= 0] = 1.0
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        got = rec.fbp_reconstruction(sino, 80.0, apply_log=True)
    assert any("values <= 0.0" in str(x.message) for x in w)
    assert sino[5, 7, 30] == 1.0 and sino[9, 0, 3] == 1.0
    want = rec.fbp_reconstruction(_device(fixed), 80.0,
                                  apply_log=True).cpu().numpy()
    assert np.array_equal(got, want)


def test_streamed_volume_360(tmp_path):
    import torch
    from algotom_amd.rec import volume
    rng = np.random.default_rng(4)
    sino = rng.uniform(0.3, 1.0, (241, 7, 180)).astype(np.float32)
    out = np.lib.format.open_memmap(str(tmp_path / "v.npy"), mode="w+",
                                    dtype=np.float32, shape=(300, 7, 300))
    _, cor_h = volume.reconstruct_volume_360(sino, 29.5, chunk=3, out=out)
    want, cor_d = volume.reconstruct_volume_360(_device(sino), 29.5, chunk=3)
    assert cor_h == cor_d and tuple(want.shape) == (300, 7, 300)
    assert np.array_equal(np.asarray(out), want.cpu().numpy())
    torch.cuda.synchronize()


@pytest.mark.parametrize("rows,stage", [(37, 1 << 16), (5, 1 << 12)])
def test_stager_blocks(rows, stage):
    """Uploads and downloads split into many staging blocks, from and into
    strided host arrays, with a dtype cast on upload."""
    import torch
    from algotom_amd import _host
    st = _host.Stager(nbytes=stage)
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(rows)
    host = (rng.random((rows, 3, 40, 50)) * 100).astype(np.uint16)
    src = host[:, 1, ::2, :]                      # strided rows
    dev = torch.empty(src.shape, dtype=torch.float32, device="cuda")
    st.upload(src, dev, stream)
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy(), src.astype(np.float32))
    back = np.zeros((rows, 2, 20, 50), dtype=np.float32)
    dst = back[:, 1]
    st.download(dev * 2, dst, stream)
    assert np.array_equal(back[:, 1], 2 * src.astype(np.float32))
    assert not back[:, 0].any()


@pytest.mark.parametrize("dtype", ["float32", "uint16"])
def test_vertrec_host_prefetch_equals_device(monkeypatch, dtype):
    """vertical_reconstruction_multiple on a host array (prefetched groups,
    several per call) equals the call on the same projections on the
    device."""
    import torch
    from algotom_amd.rec import vertrec
    monkeypatch.setattr(vertrec, "_GROUP_BYTES", 200_000)  # 4 groups
    rng = np.random.default_rng(7)
    proj = (rng.random((97, 12, 80)) * 3000 + 100).astype(dtype)
    proj_view = proj[:, 1:11, :]                  # strided host view
    kw = dict(alpha=20.0, step_index=3, apply_log=True,
              ramp_filter="before", show_progress=False, chunk_size=10)
    got = vertrec.vertical_reconstruction_multiple(proj_view, 30, 36, 39.2,
                                                   **kw)
    dev = torch.from_numpy(np.ascontiguousarray(proj_view).astype(
        np.float32)).cuda()
    want = vertrec.vertical_reconstruction_multiple(dev, 30, 36, 39.2, **kw)
    want = want.cpu().numpy() if torch.is_tensor(want) else want
    assert np.array_equal(np.asarray(got), np.asarray(want))
