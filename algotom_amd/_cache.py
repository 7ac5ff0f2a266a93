"""Thread-safe LRU caches of device buffers (filter plans and the like).

Several host threads may call the reconstruction functions at once, each on
its own stream.  An entry is built under the cache's lock, so two threads
asking for the same key share one build; lookups move the entry to the end
of the LRU order; the oldest entries are dropped once there are more than
``maxsize``.  Dropping an entry only releases its tensors to torch's caching
allocator: every stream other than the one that built an entry is recorded
on its tensors when it uses them (``Tensor.record_stream``), so the memory is
handed out again only after the work enqueued on those streams so far has
finished.
"""

import collections
import threading


class DeviceCache:
    def __init__(self, maxsize):
        self.maxsize = int(maxsize)
        self._lock = threading.RLock()
        self._entries = collections.OrderedDict()

    def __len__(self):
        with self._lock:
            return len(self._entries)

    def clear(self):
        with self._lock:
            self._entries.clear()

    def get(self, key, build, tensors=lambda entry: ()):
        """The entry for ``key``, built by ``build()`` when missing;
        ``tensors(entry)`` lists the device tensors the entry holds."""
        import torch
        with self._lock:
            hit = self._entries.get(key)
            if hit is None:
                value = build()
                homes = [torch.cuda.current_stream(t.device).cuda_stream
                         for t in tensors(value)]
                hit = (value, homes)
                self._entries[key] = hit
                while len(self._entries) > self.maxsize:
                    self._entries.popitem(last=False)
            else:
                self._entries.move_to_end(key)
            value, homes = hit
            for t, home in zip(tensors(value), homes):
                cur = torch.cuda.current_stream(t.device)
                if cur.cuda_stream != home:
                    t.record_stream(cur)
            return value
