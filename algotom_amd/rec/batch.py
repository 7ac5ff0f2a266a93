"""Batched FBP of many sinograms of uneven width, angle count and centre.

The reference reconstructs slices one call at a time (reconstruction.py
:460-490) and datasets in separate SLURM jobs.  Here a batch is planned once
(device buffers, filter plans, job descriptors) and executed as

* one ramp-filter launch per FFT length and input type (every row of every
  sinogram that shares them, one workgroup per row), then
* one persistent back-projection launch per arithmetic over the (sinogram,
  tile) work list of its sinograms, workgroups pulling items from a device
  counter, heaviest sinograms (most angles) first.

``capture()`` records that launch sequence in a HIP graph (torch.cuda.CUDAGraph
on ROCm) so a replay costs one graph launch.  Per sinogram the arithmetic is
that of ``fbp_reconstruction`` with the same arguments.
"""

import numpy as np

from algotom_amd import _native as nat
from algotom_amd.rec import autotune
from algotom_amd.rec import reconstruction as rec


class BatchFBP:
    """Plan and run FBP for a list of sinograms.

    Parameters
    ----------
    shapes : list of tuple
        (nrow, ncol) per 2D sinogram, or (nrow, nslice, ncol) per stack.
    centers : list of float
        Rotation centre per sinogram.
    angles : list of array_like or None
        Angles (radian) per sinogram; None gives the reference default
        deg2rad(linspace(0, 180, nrow)).
    filter_name, pad, pad_mode, apply_log, ratio, gpu :
        As in fbp_reconstruction, shared by the batch.
    precision : {None, "fp32", "fp16", "bf16"}
        None keeps the reference arithmetic.
    tile_rows : int or None
        Back-projection tile height; None takes the tuned value.
    dtypes : list of numpy dtypes or None
        Element type per sinogram (float32 or float64, as fbp_reconstruction
        keeps it through the log and the filter); None: all float32.

    Each sinogram gets exactly the arithmetic ``fbp_reconstruction`` gives
    it: with ``precision=None`` a single sinogram with ``gpu=True`` is
    back-projected with float64 sums and the rest with float32 storage
    after every angle, so the batch runs one persistent back-projection
    launch per arithmetic it contains.
    """

    def __init__(self, shapes, centers, angles=None, filter_name="hann",
                 pad=None, pad_mode="edge", apply_log=False, ratio=1.0,
                 gpu=True, precision=None, tile_rows=None, nblocks=0,
                 dtypes=None):
        import torch
        self.dev = nat.device()
        self.lib = nat.load()
        if pad_mode not in nat.FILTER_PAD_MODES:
            raise NotImplementedError("pad_mode %r" % pad_mode)
        shapes = [tuple(s) for s in shapes]
        n = len(shapes)
        if len(centers) != n:
            raise ValueError("one centre per sinogram is required")
        if angles is None:
            angles = [None] * n
        if dtypes is None:
            dtypes = [np.float32] * n
        in_types = [torch.float64 if np.dtype(d) == np.float64
                    else torch.float32 for d in dtypes]
        self.shapes3 = [(s[0], 1, s[1]) if len(s) == 2 else s for s in shapes]
        self.input_3d = [len(s) == 3 for s in shapes]
        self.apply_log = bool(apply_log)
        self.pad_mode = pad_mode
        self.precision = precision
        if precision is None:
            accs = [nat.ACC_F64 if (gpu and s[1] == 1) else nat.ACC_F32SEQ
                    for s in self.shapes3]
        else:
            accs = [nat.ACC_F32] * n
        store = rec._storage_dtype(precision)
        self.store_code = nat.dtype_code(store)
        main = max(set(accs), key=lambda a: sum(
            s[0] * s[1] * s[2] * s[2]
            for s, b in zip(self.shapes3, accs) if b == a))
        key = autotune.make_key("batched", main, str(store).split(".")[-1],
                                max(s[2] for s in self.shapes3),
                                max(s[1] for s in self.shapes3))
        self.tune_key = key
        valid = {r for _, r in autotune.compiled_configs("batch")}
        if tile_rows is None:
            hit = autotune.lookup(key)
            tile_rows = hit[1] if hit and hit[1] in valid else 32
        self.tile_rows = int(tile_rows)
        self.nblocks = int(nblocks)

        self.inputs, self.filtered, self.outputs = [], [], []
        self._angles, self._keep = [], []
        for (nrow, nslice, ncol), ang, tin in zip(self.shapes3, angles,
                                                  in_types):
            self.inputs.append(torch.empty((nrow, nslice, ncol), dtype=tin,
                                           device=self.dev))
            self.filtered.append(torch.empty((nrow, nslice, ncol),
                                             dtype=store, device=self.dev))
            self.outputs.append(torch.empty((ncol, nslice, ncol),
                                            dtype=torch.float32,
                                            device=self.dev))
            if ang is None:
                ang = np.deg2rad(np.linspace(0.0, 180.0, nrow))
            if len(ang) != nrow:
                raise ValueError("!!!Number of angles is not the same as the "
                                 "row number of the sinogram!!!")
            self._angles.append(rec._angles_device(ang, self.dev))

        # ---- filter groups: one launch per (FFT length, input type)
        self.filter_groups = []
        groups = {}
        for i, (nrow, nslice, ncol) in enumerate(self.shapes3):
            p = min(int(0.15 * ncol), 150) if pad is None else int(pad)
            L = nat.check(self.lib.ath_filter_fft_length(ncol, p),
                          "ath_filter_fft_length")
            code = nat.dtype_code(in_types[i])
            groups.setdefault((L, code), []).append((i, p))
        for (L, code), members in sorted(groups.items()):
            jobs = (nat.FilterJob * len(members))()
            row = 0
            for k, (i, p) in enumerate(members):
                nrow, nslice, ncol = self.shapes3[i]
                taps = rec._window_taps(None, nrow, ncol + 2 * p, filter_name)
                plan, nwin, _ = rec._filter_plan(taps, self.dev, L)
                self._keep.append(plan)
                jobs[k] = nat.FilterJob(self.inputs[i].data_ptr(),
                                        self.filtered[i].data_ptr(),
                                        plan.data_ptr(), nrow, nslice, ncol,
                                        p, nwin, 0, row)
                row += nrow * nslice
            self.filter_groups.append((L, code, len(members), row,
                                       self._to_device(jobs)))

        # ---- back-projection work lists, one per arithmetic: heaviest
        # per-tile cost (most angles) first
        self.bp_groups = []
        for acc in sorted(set(accs)):
            members = [i for i in range(n) if accs[i] == acc]
            order = sorted(members, key=lambda i: -self.shapes3[i][0])
            jobs = (nat.BpJob * len(order))()
            for k, i in enumerate(order):
                nrow, nslice, ncol = self.shapes3[i]
                c = centers[i]
                if c < 0 or c >= ncol:
                    raise ValueError("Center is out of the range [0, {}]"
                                     .format(ncol - 1))
                radius = -1.0
                if ratio is not None:
                    r = ratio
                    if r == 0.0:
                        r = min(c, ncol - c) / (0.5 * ncol)
                    radius = float(r * (ncol // 2))
                jobs[k] = nat.BpJob(
                    self.filtered[i].data_ptr(), self.outputs[i].data_ptr(),
                    self._angles[i].data_ptr(), nrow, nslice, ncol,
                    float(np.float32(c)), radius, 1, 0, 0)
            sb = 1 if all(self.shapes3[i][1] == 1 for i in members) else 4
            self.bp_groups.append({
                "acc": acc, "slice_batch": sb, "jobs": jobs,
                "counter": torch.zeros(1, dtype=torch.int32,
                                       device=self.dev)})
        self.acc = main
        self.updates = sum(s[0] * s[1] * s[2] * s[2] for s in self.shapes3)
        self._retile()
        self.graph = None

    @property
    def total_tiles(self):
        return sum(g["total"] for g in self.bp_groups)

    def _retile(self):
        for g in self.bp_groups:
            tile = 0
            for k in range(len(g["jobs"])):
                jb = g["jobs"][k]
                jb.tile_begin = tile
                tile += self.lib.ath_bp_job_tiles(jb.ncol, jb.nslice,
                                                  g["slice_batch"],
                                                  self.tile_rows)
            g["total"] = tile
            g["jobs_dev"] = self._to_device(g["jobs"])

    def _to_device(self, cstruct):
        import torch
        raw = np.frombuffer(bytes(cstruct), dtype=np.uint8).copy()
        return torch.from_numpy(raw).to(self.dev)

    def set_tile_rows(self, tile_rows):
        """Change the back-projection tile height (drops a captured graph)."""
        self.tile_rows = int(tile_rows)
        self._retile()
        self.graph = None

    def launch(self):
        """Enqueue the whole batch on the current stream."""
        stream = nat.stream_ptr(self.dev)
        for L, code, njobs, rows, jobs in self.filter_groups:
            nat.check(self.lib.ath_ramp_filter_batched(
                nat.ptr(jobs), njobs, rows, code, self.store_code, L,
                nat.FILTER_PAD_MODES[self.pad_mode], int(self.apply_log),
                stream),
                "ath_ramp_filter_batched")
        for g in self.bp_groups:
            nat.check(self.lib.ath_back_projection_batched(
                nat.ptr(g["jobs_dev"]), len(g["jobs"]), g["total"],
                self.store_code, g["acc"], g["slice_batch"], self.tile_rows,
                self.nblocks, nat.ptr(g["counter"]), stream),
                "ath_back_projection_batched")

    def capture(self):
        """Record launch() in a HIP graph; run() then replays it."""
        import torch
        side = torch.cuda.Stream(self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(side):
            self.launch()  # warm-up outside the graph (LDS attributes, code)
        torch.cuda.current_stream(self.dev).wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.launch()
        self.graph = graph
        return graph

    def run(self):
        if self.graph is not None:
            self.graph.replay()
        else:
            self.launch()
        return self.outputs

    def autotune(self, persist=True):
        """Time every tile height on this batch, keep the fastest."""
        def run(cfg):
            self.set_tile_rows(cfg[1])
            self.launch()
        best, timings = autotune.tune(
            self.tune_key, [(1, r) for r in autotune.CANDIDATE_ROWS], run,
            persist)
        self.set_tile_rows(best[1])
        return best, timings

    def load(self, sinograms):
        """Copy sinograms (numpy or tensors) into the input buffers (float64
        ones stay float64); values <= 0 are handled as in
        fbp_reconstruction when apply_log is set."""
        import torch
        for i, s in enumerate(sinograms):
            buf = self.inputs[i]
            if isinstance(s, torch.Tensor):
                t = s
            else:
                a = np.asarray(s)
                want = (np.float64 if buf.dtype == torch.float64
                        else np.float32)
                t = torch.from_numpy(np.ascontiguousarray(a, dtype=want))
            buf.copy_(t.reshape(buf.shape))
            if self.apply_log:
                rec._clamp_nonpositive_device(s, buf)

    def results(self, as_numpy=True):
        out = []
        for o, is3d in zip(self.outputs, self.input_3d):
            r = o if is3d else o[:, 0, :]
            out.append(rec._to_host(r) if as_numpy else r)
        return out


def fbp_reconstruction_batch(sinograms, centers, angles=None,
                             filter_name="hann", pad=None, pad_mode="edge",
                             apply_log=True, ratio=1.0, gpu=True,
                             precision=None):
    """fbp_reconstruction over a list of sinograms of any shapes, in one
    batched execution.  Returns a list of numpy arrays."""
    dtypes = [np.float64 if str(getattr(s, "dtype", "")).endswith("float64")
              else np.float32 for s in sinograms]
    plan = BatchFBP([tuple(s.shape) for s in sinograms], centers, angles,
                    filter_name=filter_name, pad=pad, pad_mode=pad_mode,
                    apply_log=apply_log, ratio=ratio, gpu=gpu,
                    precision=precision, dtypes=dtypes)
    plan.load(sinograms)
    plan.run()
    return plan.results()
