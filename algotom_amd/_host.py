"""Device-to-host copies of large results.

Results of ``STAGE_MIN`` bytes or more go through two pinned staging buffers
(chunked asynchronous copies on the current stream) and are copied into the
new numpy array by a host thread pool.  A plain pageable copy into a fresh
array is bound by first-touch page faults on one thread (1.7 GB: ~400 ms
against ~40 ms this way on MI355X hosts).
"""

import os
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

STAGE_BYTES = 128 << 20          # per staging buffer
STAGE_MIN = 64 << 20             # smaller results take a plain copy
_stage = {}
_pools = []
_stage_lock = threading.Lock()   # one staged transfer at a time


def _host_pools():
    """(drain, copy, n): one task per staging buffer, and the n workers its
    copies are split over."""
    if not _pools:
        n = max(2, min(16, os.cpu_count() or 2))
        _pools.extend([ThreadPoolExecutor(max_workers=2),
                       ThreadPoolExecutor(max_workers=n), n])
    return _pools


def to_host(t):
    """A new numpy array holding tensor ``t`` (device or host)."""
    import torch
    nbytes = t.numel() * t.element_size()
    if (not t.is_cuda or nbytes < STAGE_MIN
            or t.dtype not in (torch.float32, torch.float64)):
        return t.detach().cpu().numpy()
    with _stage_lock:
        return _to_host_staged(t.detach(), nbytes)


def _to_host_staged(t, nbytes):
    import torch
    src = t.contiguous().view(torch.uint8).reshape(-1)
    out = np.empty(tuple(t.shape), dtype=np.float32
                   if t.dtype == torch.float32 else np.float64)
    dst = out.reshape(-1).view(np.uint8)
    bufs = _stage.get(t.device.index)
    if bufs is None:
        bufs = [torch.empty(STAGE_BYTES, dtype=torch.uint8, pin_memory=True)
                for _ in range(2)]
        _stage[t.device.index] = bufs
    drain, copiers, nthreads = _host_pools()

    def host_copy(buf, i0, i1, ev):
        ev.synchronize()
        bn = buf.numpy()
        m = i1 - i0
        parts = [(j * m // nthreads, (j + 1) * m // nthreads)
                 for j in range(nthreads)]
        list(copiers.map(lambda q: np.copyto(dst[i0 + q[0]:i0 + q[1]],
                                             bn[q[0]:q[1]]), parts))

    stream = torch.cuda.current_stream(t.device)
    pending = [None, None]
    for k, i0 in enumerate(range(0, nbytes, STAGE_BYTES)):
        i1 = min(nbytes, i0 + STAGE_BYTES)
        slot = k % 2
        if pending[slot] is not None:
            pending[slot].result()  # the host has drained this buffer
        bufs[slot][:i1 - i0].copy_(src[i0:i1], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(stream)
        pending[slot] = drain.submit(host_copy, bufs[slot], i0, i1, ev)
    for fut in pending:
        if fut is not None:
            fut.result()
    return out
