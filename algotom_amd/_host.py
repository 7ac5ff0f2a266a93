"""Host <-> device copies of large arrays, and a chunk pipeline that overlaps
them with computation.

Large transfers go through two pinned staging buffers: the DMA of one block
runs while host threads copy the other block between its buffer and the
numpy array.  A plain pageable copy into a fresh array is bound by
first-touch page faults on one thread (1.7 GB: ~400 ms against ~40 ms this
way on MI355X hosts).  The host side accepts any strides, so a slice range
of a larger array (a memory map included) moves without an extra copy.
"""

import os
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

STAGE_BYTES = 128 << 20          # per staging buffer
STAGE_MIN = 64 << 20             # smaller results take a plain copy
_INLINE_BYTES = 4 << 20          # host copies below this stay on one thread
_stagers = {}
_pool = []
_stage_lock = threading.Lock()   # one to_host transfer at a time
_pipes = {}                      # per device: copy streams and stagers
_pipe_lock = threading.Lock()    # one chunk pipeline at a time


def _copy_pool():
    """(executor, n): the host threads that block copies are split over."""
    if not _pool:
        n = max(2, min(16, os.cpu_count() or 2))
        _pool.extend([ThreadPoolExecutor(max_workers=n), n])
    return _pool


def _copy(dst, src):
    """np.copyto(dst, src) (casting as needed), split over the copy pool."""
    if dst.nbytes < _INLINE_BYTES or dst.ndim == 0:
        np.copyto(dst, src, casting="unsafe")
        return
    pool, n = _copy_pool()
    axis = 0 if (dst.shape[0] >= n or dst.ndim == 1) else 1
    m = dst.shape[axis]
    parts = [(i * m // n, (i + 1) * m // n) for i in range(n)]
    parts = [p for p in parts if p[1] > p[0]]

    def one(p):
        sl = (slice(p[0], p[1]),) if axis == 0 else (slice(None),
                                                     slice(p[0], p[1]))
        np.copyto(dst[sl], src[sl], casting="unsafe")

    list(pool.map(one, parts))


def _np_dtype(torch_dtype):
    import torch
    return {torch.float32: np.float32, torch.float64: np.float64,
            torch.float16: np.float16, torch.int32: np.int32,
            torch.uint8: np.uint8, torch.int16: np.int16}[torch_dtype]


class Stager:
    """Two pinned buffers moving blocks of axis 0 between a host array (any
    strides and, for uploads, any real dtype) and a contiguous device
    tensor, on one stream.  One Stager serves one thread at a time."""

    def __init__(self, nbytes=STAGE_BYTES):
        self.nbytes = int(nbytes)
        self.bufs = None
        self.events = [None, None]

    def _plan(self, shape, itemsize):
        row = itemsize * int(np.prod(shape[1:], dtype=np.int64))
        per = max(1, self.nbytes // max(row, 1))
        need = per * row
        if self.bufs is None or self.bufs[0].numel() < need:
            import torch
            self.sync()
            self.bufs = [torch.empty(max(need, self.nbytes), dtype=torch.uint8,
                                     pin_memory=True) for _ in range(2)]
        return [(r0, min(shape[0], r0 + per))
                for r0 in range(0, shape[0], per)], row

    def _view(self, slot, r0, r1, row, shape, torch_dtype):
        return (self.bufs[slot][:(r1 - r0) * row].view(torch_dtype)
                .view((r1 - r0,) + tuple(shape[1:])))

    def sync(self):
        """Wait for the DMAs that still read the staging buffers."""
        for ev in self.events:
            if ev is not None:
                ev.synchronize()
        self.events = [None, None]

    def upload(self, src, dst, stream):
        """Enqueue ``dst <- src`` on ``stream``; returns once the host copies
        are done (the last DMA may still run: order later work after it on
        ``stream``)."""
        import torch
        if tuple(src.shape) != tuple(dst.shape) or not dst.is_contiguous():
            raise ValueError("upload: shape mismatch or non-contiguous target")
        groups, row = self._plan(dst.shape, dst.element_size())
        for g, (r0, r1) in enumerate(groups):
            slot = g % 2
            if self.events[slot] is not None:
                self.events[slot].synchronize()   # its last DMA has read it
            pin = self._view(slot, r0, r1, row, dst.shape, dst.dtype)
            _copy(pin.numpy(), src[r0:r1])
            with torch.cuda.stream(stream):
                dst[r0:r1].copy_(pin, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
            self.events[slot] = ev

    def download(self, src, dst, stream):
        """``dst <- src`` after the work already on ``stream``; returns when
        ``dst`` holds the data."""
        import torch
        if tuple(src.shape) != tuple(dst.shape) or not src.is_contiguous():
            raise ValueError("download: shape mismatch or non-contiguous source")
        if dst.dtype != _np_dtype(src.dtype):
            raise ValueError("download: dtype mismatch")
        self.sync()
        groups, row = self._plan(src.shape, src.element_size())

        def drain(slot, r0, r1, ev):
            ev.synchronize()
            _copy(dst[r0:r1],
                  self._view(slot, r0, r1, row, src.shape, src.dtype).numpy())

        prev = None
        for g, (r0, r1) in enumerate(groups):
            slot = g % 2                          # drained two blocks ago
            with torch.cuda.stream(stream):
                self._view(slot, r0, r1, row, src.shape, src.dtype).copy_(
                    src[r0:r1], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
            if prev is not None:
                drain(*prev)                      # overlaps the DMA above
            prev = (slot, r0, r1, ev)
        if prev is not None:
            drain(*prev)


def to_host(t):
    """A new numpy array holding tensor ``t`` (device or host)."""
    import torch
    nbytes = t.numel() * t.element_size()
    if (not t.is_cuda or nbytes < STAGE_MIN
            or t.dtype not in (torch.float32, torch.float64)):
        return t.detach().cpu().numpy()
    src = t.detach().contiguous()
    out = np.empty(tuple(src.shape), dtype=_np_dtype(src.dtype))
    with _stage_lock:
        st = _stagers.setdefault(src.device.index, Stager())
        st.download(src, out, torch.cuda.current_stream(src.device))
    return out


def stream_chunks(src, out, ranges, compute, in_dtype, device):
    """Chunk loop over axis 1 of host arrays with transfers overlapped.

    ``src`` is (nrow, nslice, ncol) and ``out`` (a, nslice, b), both numpy
    (any strides; memory maps included).  For each (j0, j1) of ``ranges``,
    ``compute(part)`` receives ``src[:, j0:j1, :]`` as a contiguous device
    tensor of ``in_dtype`` (it may modify it) and returns the device result
    for ``out[:, j0:j1, :]``, enqueued on the current stream.  Chunk k+1
    uploads on one stream and chunk k-1 downloads on another while chunk k
    computes; at most two results wait for download at a time.
    """
    if not ranges:
        return out
    with _pipe_lock:
        return _stream_chunks(src, out, ranges, compute, in_dtype, device)


def _stream_chunks(src, out, ranges, compute, in_dtype, device):
    import torch
    main = torch.cuda.current_stream(device)
    key = torch.device(device).index
    if key not in _pipes:
        _pipes[key] = (torch.cuda.Stream(device), torch.cuda.Stream(device),
                       Stager(), Stager())
    up, down, up_stage, down_stage = _pipes[key]
    nrow, ncol = src.shape[0], src.shape[2]
    most = max(j1 - j0 for j0, j1 in ranges)
    bufs = [torch.empty(nrow * most * ncol, dtype=in_dtype, device=device)
            for _ in range(2)]
    done = [None, None]                    # compute event per input buffer

    def part_of(k):
        j0, j1 = ranges[k]
        return bufs[k % 2][:nrow * (j1 - j0) * ncol].view(nrow, j1 - j0, ncol)

    def upload(k, wait):
        j0, j1 = ranges[k]
        if wait is not None:
            up.wait_event(wait)            # chunk k-2 has finished with it
        up_stage.upload(src[:, j0:j1, :], part_of(k), up)
        ev = torch.cuda.Event()
        ev.record(up)
        return ev

    def download(k, res, ev):
        j0, j1 = ranges[k]
        down.wait_event(ev)
        down_stage.download(res, out[:, j0:j1, :], down)

    up_pool = ThreadPoolExecutor(max_workers=1)
    down_pool = ThreadPoolExecutor(max_workers=1)
    pending, fut = [], None
    try:
        fut = up_pool.submit(upload, 0, None)
        for k in range(len(ranges)):
            ev_in = fut.result()
            fut = None
            if k + 1 < len(ranges):
                fut = up_pool.submit(upload, k + 1, done[(k + 1) % 2])
            main.wait_event(ev_in)
            res = compute(part_of(k)).contiguous()
            j0, j1 = ranges[k]
            if tuple(res.shape) != (out.shape[0], j1 - j0, out.shape[2]):
                raise ValueError("stream_chunks: result shape %s"
                                 % (tuple(res.shape),))
            ev = torch.cuda.Event()
            ev.record(main)
            done[k % 2] = ev
            if len(pending) >= 2:
                pending[-2].result()       # bounds the results held
            pending.append(down_pool.submit(download, k, res, ev))
            del res
        for p in pending:
            p.result()
    finally:
        for f in [fut] + pending:
            if f is not None:
                try:
                    f.result()
                except Exception:          # the first error is re-raised
                    pass
        up_pool.shutdown(wait=True)
        down_pool.shutdown(wait=True)
        up.synchronize()
        down.synchronize()
        main.synchronize()
    return out


_prefetch = {}                   # per device: copy stream and stager
_prefetch_lock = threading.Lock()


def prefetch_chunks(src, ranges, dtype, device):
    """Yields (k, tensor) for k, (r0, r1) in enumerate(ranges): a contiguous
    device tensor of ``dtype`` holding src[r0:r1] (numpy, any strides, a
    memory map included).  Chunk k+1 uploads on a copy stream while the
    caller works on chunk k; the caller's work must be enqueued on the
    current stream before it asks for the next chunk, and a yielded tensor
    is valid until then."""
    import torch
    if not ranges:
        return
    with _prefetch_lock:
        main = torch.cuda.current_stream(device)
        key = torch.device(device).index
        if key not in _prefetch:
            _prefetch[key] = (torch.cuda.Stream(device), Stager())
        up, stage = _prefetch[key]
        shape1 = tuple(src.shape[1:])
        row = int(np.prod(shape1, dtype=np.int64))
        most = max(r1 - r0 for r0, r1 in ranges)
        bufs = [torch.empty(most * row, dtype=dtype, device=device)
                for _ in range(2)]
        done = [None, None]

        def view(k):
            r0, r1 = ranges[k]
            return bufs[k % 2][:(r1 - r0) * row].view((r1 - r0,) + shape1)

        def upload(k, wait):
            r0, r1 = ranges[k]
            if wait is not None:
                up.wait_event(wait)        # chunk k-2 is finished with it
            stage.upload(src[r0:r1], view(k), up)
            ev = torch.cuda.Event()
            ev.record(up)
            return ev

        pool = ThreadPoolExecutor(max_workers=1)
        fut = None
        try:
            fut = pool.submit(upload, 0, None)
            for k in range(len(ranges)):
                ev = fut.result()
                fut = None
                if k + 1 < len(ranges):
                    # its buffer was chunk k-1's, finished at done[...]
                    fut = pool.submit(upload, k + 1, done[(k + 1) % 2])
                main.wait_event(ev)
                yield k, view(k)
                done_ev = torch.cuda.Event()
                done_ev.record(main)
                done[k % 2] = done_ev
        finally:
            if fut is not None:
                try:
                    fut.result()
                except Exception:
                    pass
            pool.shutdown(wait=True)
            up.synchronize()
            main.synchronize()
