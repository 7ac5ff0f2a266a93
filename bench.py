"""Benchmark: FBP of a 64 x 1801 x 2560 sinogram stack on one MI355X.

Workload (the reference's stack case): 64 seeded random-ellipse sinograms,
1801 angles over [0, pi), 2560 detector columns, transmission
Poisson(1e4 exp(-k p)) / 1e4 with max(k p) = 3, apply_log=True, centre
1280.37, hann window, edge padding, circle mask, output 64 x 2560 x 2560
float32.  One step = one fbp_reconstruction call on the device-resident
stack (log, pad, ramp filter, back-projection, mask, scale).

value = pixel-angle updates per second (GUPS = 1e9 updates/s),
U = 64 * 2560^2 * 1801 = 7.554e11 updates per step.

roofline (back-projection kernel, the dominant one): it is a gather with no
MFMA work and its HBM traffic is 8 % of peak, so the bound is VALU issue.
achieved = the kernel's VALU instructions per launch (SQ_INSTS_VALU, from the
committed rocprofv3 counters of this kernel and workload,
profiles/bp_counters.json) x 64 lanes over the kernel's mean duration from
HIP events on the launch stream; peak = 256 CU x 64 lanes x 2.4 GHz = 39.3 T
lane-op/s; traffic = the same profile's calibrated FETCH_SIZE + WRITE_SIZE
bytes per launch.  The gather-byte model (U x 8 B) and the compulsory-byte
HBM fraction are reported in detail.
cpu_baseline: the oracle's OpenMP C port (numba-typed restatement of the
reference's CPU path) on 4 of the 64 sinograms, on this host's cores
(median of 5 after a warm-up).

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--precision fp32]

--workload refine: differentiable FBP least-squares refinement. 16
sinograms of 901 angles over a 120-degree arc x 1024 columns (seed 4000),
Gaussian noise sigma = 1% of max; loss = ||FBP(s; c) - phantom||^2 +
1e-3 TV(FBP(s; c)); Adam on the sinogram s and the centre c (started 1.5 px
off) through the hand-written adjoint, the centre derivative image computed
in the forward back-projection's launch (checkpoint "derivative"); value =
grad steps/s over --steps steps (default 200).

--workload volume: 128 slices of a 3601-angle x 4096-column half-acquisition
scan over [0, 2 pi) with the rotation axis at column 499.5 (overlap 1000),
Poisson transmission at 1e4 photons (seed 5000); each step converts
360 -> 180 degrees on the GPU (7192 stitched columns) and reconstructs the
128 x 7192 x 7192 volume in chunks of 16 slices (apply_log=True).

--workload volume_refine: the same 128-slice block, converted once to
1801 x 7192 sinograms (line integrals, no noise), refined by least squares
against the phantoms: loss = ||FBP(s; c) - phantom||^2 + 1e-3 TV, Adam on the
sinograms and the centre (started 1.5 px off) through the hand-written
adjoint with the filtered sinogram checkpointed, the volume in chunks of 16
slices per step; value = grad steps/s over --steps steps (default 3).

--workload uneven: 256 sinograms with widths from {1024, 2048, 2560, 4096},
angle counts from {901, 1801, 3601} over [0, pi) and centres within +-20 px
of mid-width (seed 3000), exact projections, reconstructed as one batch: one
filter launch per FFT length and one persistent back-projection launch,
captured in a HIP graph, tile height autotuned; every precision mode is
timed and its error against the all-FP64 oracle reported on sampled
sinograms.

--workload center: centre-of-rotation search on one 1801-angle x
2560-column transmission sinogram (seed 1000, true centre 1280.37):
find_center_based_slice_metric(start=1260, stop=1300, step=0.5, radius=2,
metric="entropy") -- a coarse integer pass over 41 centres plus the reference
slice, then a sub-pixel pass; every pass smooths, logs and filters once and
back-projects all its centres in one batched launch.  value = candidate
slices reconstructed and scored per second.

--workload c1: the plain-FBP CPU workload at full size (721 x 512
Fourier-slice sinogram of a 20-ellipse phantom, committed under
tests/golden/), the CPU back-projector's arithmetic; the GPU time beside the
oracle's C port timed on the host cores.

--workload bpf: BPF (back-projection filtering) of a 16 x 1801 x 2560
transmission stack (seed 1000): log, edge-padded back-projection, the 2D
circular ramp filter (FFT length 8192) of every slice, mask and scale;
value = GUPS of the whole step (U = 16 * 2560^2 * 1801).

--workload vertrec: vertical slices straight from a 1801 x 1024 x 2560
projection stack (uniform random data on the device, seed 0): one
vertical_reconstruction (ramp filter after, reference GPU typing, chunks of
30 projections) and one vertical_reconstruction_multiple of 16 parallel
slices (ramp filter before); value = GUPS of the 16-slice call
(U = 16 * 1024 * 2560 * 1801).
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NSLICE, NROW, NCOL, CENTER = 64, 1801, 2560, 1280.37
PEAK_HBM_GBS = 8000.0
# MI355X VALU issue peak: 256 CUs x 64 lanes per clock x 2.4 GHz
PEAK_VALU_TLANE = 256 * 64 * 2.4e9 / 1e12
METRIC = ("GUPS (pixel·angle updates/s) FBP 2560²×1801 ang; "
          "% HBM roofline; grad steps/s")


def make_stack(dev, nslice=NSLICE, seed0=1000):
    import torch
    from algotom_amd.util import phantom
    angles = np.linspace(0.0, np.pi, NROW, endpoint=False)
    sino = torch.empty((NROW, nslice, NCOL), dtype=torch.float32, device=dev)
    for j in range(nslice):
        prm = phantom.random_ellipses(NCOL, 20, seed=seed0 + j)
        p = phantom.ellipse_sinogram(prm, angles, NCOL, np.float32(CENTER),
                                     device=dev)
        sino[:, j, :] = phantom.transmission(p, seed=seed0 + j)
    torch.cuda.synchronize()
    return sino, angles


def copy_bandwidth(dev, nbytes=1 << 30, reps=5):
    """A plain device copy of nbytes (torch's copy kernel): measured HBM
    bandwidth (read + write bytes / time), and the calibration launch for
    the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def load_counters():
    """rocprofv3 counters of the reference-arithmetic back-projection kernel
    on this workload (profiles/bp_counters.json), or None."""
    path = os.path.join(ROOT, "profiles", "bp_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def time_cpu(fn, runs=5):
    """One warm-up call, then the median (and spread) of `runs` timed calls
    of fn(); returns (median_s, [times], last result)."""
    fn()
    times, res = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        res = fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), times, res


def cpu_baseline(sino_dev, angles, gpu_out, nsample=4):
    """Times the oracle C port on nsample slices (median of 5 after a
    warm-up of the same size); also checks parity of the GPU result on
    those slices."""
    from oracle import port
    sample = np.ascontiguousarray(sino_dev[:, :nsample, :].cpu().numpy())
    threads = port.default_threads()
    dt, times, ref = time_cpu(lambda: port.fbp(sample, CENTER, angles,
                                               nthreads=threads))
    updates = nsample * NCOL * NCOL * NROW
    got = gpu_out[:, :nsample, :].cpu().numpy()
    err = float(np.abs(got.astype(np.float64) - ref).max()
                / np.abs(ref).max())
    return {"value": updates / dt / 1e9, "unit": "GUPS", "cores": threads,
            "kind": "port",
            "sample": "%d of %d sinograms (%dx%d), log+filter+BP+mask+scale; "
                      "median of 5 after a warm-up: %.2f s (%.2f-%.2f s)"
                      % (nsample, NSLICE, NROW, NCOL, dt, min(times),
                         max(times))}, err


def host_roundtrip(sino_dev, angles, reps=2):
    """fbp_reconstruction from host numpy to host numpy (slice chunks whose
    host-to-device and device-to-host copies overlap the filter and
    back-projection of their neighbours): the drop-in's PCIe-inclusive time,
    best of reps."""
    from algotom_amd.rec import reconstruction as rec
    s_np = sino_dev.cpu().numpy()
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        rec.fbp_reconstruction(s_np, CENTER, angles)
        dt = (time.perf_counter() - t0) * 1e3
        best = dt if best is None else min(best, dt)
    return best


def run_uneven(args, dev):
    import torch
    from algotom_amd.rec import batch
    from algotom_amd.util import phantom
    from oracle import port
    rng = np.random.default_rng(3000)
    nsino = 256
    widths = rng.choice([1024, 2048, 2560, 4096], nsino)
    nangles = rng.choice([901, 1801, 3601], nsino)
    centers = [float((w - 1) / 2 + rng.uniform(-20.0, 20.0)) for w in widths]
    angles = [np.linspace(0.0, np.pi, int(n), endpoint=False) for n in nangles]
    shapes = [(int(n), int(w)) for n, w in zip(nangles, widths)]
    sinos = []
    for i, ((nrow, ncol), c) in enumerate(zip(shapes, centers)):
        prm = phantom.random_ellipses(ncol, 20, seed=3000 + i)
        sinos.append(phantom.ellipse_sinogram(
            prm, angles[i], ncol, np.float32(c), device=dev).float())
    torch.cuda.synchronize()
    updates = sum(n * w * w for n, w in shapes)
    # FP64 oracle on the two smallest sinograms
    sample = sorted(range(nsino), key=lambda i: shapes[i][0] * shapes[i][1])[:2]
    refs = {i: port.fbp_f64(sinos[i].cpu().numpy(), centers[i], angles[i],
                            apply_log=False) for i in sample}
    # the plain drop-in: one fbp_reconstruction call per sinogram (device
    # tensors in and out, reference arithmetic, no batching or graph)
    from algotom_amd.rec import reconstruction as rec
    for i in range(min(4, nsino)):
        rec.fbp_reconstruction(sinos[i], centers[i], angles[i],
                               apply_log=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(nsino):
        rec.fbp_reconstruction(sinos[i], centers[i], angles[i],
                               apply_log=False)
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    modes = {}
    for prec in (None, "fp32", "fp16", "bf16"):
        plan = batch.BatchFBP(shapes, centers, angles, apply_log=False,
                              precision=prec)
        plan.load(sinos)
        best, timings = plan.autotune(persist=False)
        plan.capture()
        for _ in range(args.warmup):
            plan.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.run()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        errs = []
        for i in sample:
            got = plan.outputs[i][:, 0, :].cpu().numpy().astype(np.float64)
            errs.append(float(np.abs(got - refs[i]).max()
                              / np.abs(refs[i]).max()))
        modes[prec or "reference"] = {
            "ms_per_batch": ms, "gups": updates / (ms * 1e-3) / 1e9,
            "tile_rows": best[1],
            "autotune_ms": {str(k[1]): v for k, v in timings.items()},
            "max_rel_err_vs_fp64": max(errs)}
        del plan
        torch.cuda.empty_cache()
    key = args.precision or "reference"
    line = {
        "metric": METRIC, "value": modes[key]["gups"], "unit": "GUPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": modes[key]["ms_per_batch"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32" if args.precision is None else
                 "%s-storage/fp32-acc" % args.precision,
        "data": "synthetic",
        "config": {"workload": "fbp_uneven_batch_256", "nsino": nsino,
                   "updates": updates, "graph": True,
                   "parallelism": "single-gpu"},
        "detail": {"modes": modes,
                   "plain_per_call": {"ms": plain_ms,
                                      "gups": updates / (plain_ms * 1e-3)
                                      / 1e9},
                   "fp16_batched_over_plain": modes["fp16"]["gups"]
                   / (updates / (plain_ms * 1e-3) / 1e9),
                   "bf16_batched_over_plain": modes["bf16"]["gups"]
                   / (updates / (plain_ms * 1e-3) / 1e9),
                   "fp16_over_reference_batched": modes["fp16"]["gups"]
                   / modes["reference"]["gups"],
                   "fp64_sample": [list(shapes[i]) for i in sample]},
    }
    print(json.dumps(line))


def run_volume(args, dev):
    import torch
    from algotom_amd.rec import volume
    from algotom_amd.util import phantom
    nsl, nrow, ncol, cor = 128, 3601, 4096, 499.5
    step_a = 2.0 * np.pi / nrow
    angles360 = np.arange(nrow) * step_a
    sino = torch.empty((nrow, nsl, ncol), dtype=torch.float32, device=dev)
    for j in range(nsl):
        prm = phantom.random_ellipses(7192, 20, seed=5000 + j)
        p = phantom.ellipse_sinogram(prm, angles360, ncol, cor, device=dev)
        sino[:, j, :] = phantom.transmission(p, seed=5000 + j)
    torch.cuda.synchronize()
    n180 = nrow // 2 + 1
    angles = np.arange(n180) * step_a
    width = 2 * ncol - 1000
    out = torch.empty((width, nsl, width), dtype=torch.float32, device=dev)
    updates = nsl * width * width * n180

    def run():
        volume.reconstruct_volume_360(sino, cor, chunk=16, angles=angles,
                                      precision=args.precision, out=out)

    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    line = {
        "metric": METRIC, "value": updates / (ms * 1e-3) / 1e9,
        "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32" if args.precision is None else
                 "%s-storage/fp32-acc" % args.precision,
        "data": "synthetic",
        "config": {"workload": "fbp_volume_128x3601x4096_half_acquisition",
                   "stitched_width": width, "nrow_180": n180,
                   "chunk": 16, "updates": updates,
                   "output_GB": width * width * nsl * 4 / 1e9,
                   "parallelism": "single-gpu"},
        "detail": {"finite": bool(torch.isfinite(out).all().item()),
                   "max_abs": float(out.abs().max().item())},
    }
    # host in, host out (7.55 GB in, 26.5 GB out): the chunk pipeline
    # overlaps both transfers with the computation
    del out
    sino_np = sino.cpu().numpy()
    del sino
    torch.cuda.empty_cache()
    out_np = np.empty((width, nsl, width), dtype=np.float32)
    t0 = time.perf_counter()
    volume.reconstruct_volume_360(sino_np, cor, chunk=16, angles=angles,
                                  precision=args.precision, out=out_np)
    host_ms = (time.perf_counter() - t0) * 1e3
    line["detail"]["host_resident_ms"] = host_ms
    line["detail"]["host_resident_gups"] = updates / (host_ms * 1e-3) / 1e9
    print(json.dumps(line))


def run_volume_refine(args, dev):
    """Least-squares refinement of the large volume block: the 128 slices of
    the 3601-angle x 4096-column half-acquisition scan (axis at column
    499.5, seed 5000), converted on the GPU to 1801 x 7192 sinograms; loss =
    ||FBP(s; c) - phantom||^2 + 1e-3 TV(FBP(s; c)) over the 128 x 7192^2
    volume, Adam on the sinograms s and the centre c (started 1.5 px off).
    Every step runs the volume in chunks of 16 slices (forward, with the
    centre derivative from the same back-projection pass kept for the
    backward, then the hand-written adjoint); the centre's gradient sums
    over the chunks before the update."""
    import torch
    from algotom_amd.prep import conversion
    from algotom_amd.rec import autograd
    from algotom_amd.util import phantom
    nsl, nrow, ncol, cor, chunk = 128, 3601, 4096, 499.5, 16
    step_a = 2.0 * np.pi / nrow
    angles360 = np.arange(nrow) * step_a
    sino = torch.empty((nrow, nsl, ncol), dtype=torch.float32, device=dev)
    params = []
    for j in range(nsl):
        prm = phantom.random_ellipses(7192, 20, seed=5000 + j)
        params.append(prm)
        sino[:, j, :] = phantom.ellipse_sinogram(prm, angles360, ncol, cor,
                                                 device=dev)
    stitched, new_cor = conversion.convert_sinogram_360_to_180(sino, cor)
    del sino
    torch.cuda.empty_cache()
    n180, _, width = stitched.shape
    angles = np.arange(n180) * step_a
    target = torch.empty((width, nsl, width), dtype=torch.float32, device=dev)
    for j in range(nsl):
        target[:, j, :] = phantom.ellipse_image_torch(params[j], width, dev)
    geom = autograd.FBPGeometry(n180, width, angles, ratio=1.0,
                                checkpoint="derivative")
    print("volume_refine: data ready (%d x %d x %d sinograms)"
          % (n180, nsl, width), file=sys.stderr, flush=True)
    s_param = stitched.contiguous().requires_grad_(True)
    c_param = torch.tensor(new_cor + 1.5, dtype=torch.float64,
                           requires_grad=True)
    opt = torch.optim.Adam([{"params": [s_param], "lr": 1e-3},
                            {"params": [c_param], "lr": 0.05}])
    npix = float(target.numel())
    tnorm = math.sqrt(sum(float(torch.sum(target[:, j0:j0 + chunk, :]
                                          .double() ** 2))
                          for j0 in range(0, nsl, chunk)))
    rel = []
    count = [0]

    def step(track=False):
        opt.zero_grad(set_to_none=True)
        total = torch.zeros((), dtype=torch.float64, device=dev)
        err = torch.zeros((), dtype=torch.float64, device=dev)
        for j0 in range(0, nsl, chunk):
            j1 = min(nsl, j0 + chunk)
            y = autograd.fbp_differentiable(s_param[:, j0:j1, :], c_param,
                                            geom)
            t = target[:, j0:j1, :]
            loss = (torch.sum((y - t) ** 2)
                    + 1e-3 * autograd.total_variation(y)) / npix
            loss.backward()
            total += loss.detach()
            if track:
                err += torch.sum((y.detach() - t).double() ** 2)
            del y, loss
        opt.step()
        if track:
            rel.append(math.sqrt(float(err.item())) / tnorm)
        count[0] += 1
        print("volume_refine: step %d done, loss %.6g" % (count[0],
                                                         float(total.item())),
              file=sys.stderr, flush=True)
        return total

    first = float(step(track=True).item())
    for _ in range(max(args.warmup - 1, 0)):
        step()
    torch.cuda.synchronize()
    nsteps = args.steps if args.steps_given else 3
    t0 = time.perf_counter()
    for _ in range(nsteps):
        loss = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    updates = nsl * width * width * n180
    line = {
        "metric": METRIC, "value": nsteps / dt, "unit": "grad steps/s",
        "n_gpus": 1, "steps": nsteps, "warmup": args.warmup,
        "ms_per_step": dt / nsteps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "fbp_refine_volume_128x1801x7192",
                   "nslice": nsl, "nrow_180": n180, "width": width,
                   "chunk": chunk, "checkpoint": "derivative",
                   "parallelism": "single-gpu"},
        "detail": {"loss_first": first, "loss_last": float(loss.item()),
                   "rel_error_first_forward": rel[0],
                   "center_start": new_cor + 1.5,
                   "center_end": float(c_param.item()),
                   "center_true": new_cor,
                   "updates_per_forward": updates,
                   "peak_hbm_GB": torch.cuda.max_memory_allocated(dev) / 1e9},
    }
    print(json.dumps(line))


def run_refine(args, dev):
    import torch
    from algotom_amd.rec import autograd
    from algotom_amd.util import phantom
    nsl, nrow, ncol = 16, 901, 1024
    angles = np.linspace(0.0, 2.0 * np.pi / 3.0, nrow)
    c_true = (ncol - 1) / 2.0 + 0.37
    sinos, targets = [], []
    rng = np.random.default_rng(4000)
    for j in range(nsl):
        prm = phantom.random_ellipses(ncol, 20, seed=4000 + j)
        targets.append(torch.from_numpy(phantom.ellipse_image(prm, ncol)))
        p = phantom.ellipse_sinogram(prm, angles, ncol, np.float32(c_true),
                                     device=dev)
        noise = torch.from_numpy(rng.standard_normal(p.shape)).to(dev)
        sinos.append((p + 0.01 * float(p.max()) * noise).float())
    sino = torch.stack(sinos, dim=1).contiguous()
    target = torch.stack(targets, dim=1).to(dev)
    geom = autograd.FBPGeometry(nrow, ncol, angles, ratio=1.0,
                                checkpoint="derivative")
    s_param = sino.clone().requires_grad_(True)
    c_param = torch.tensor(c_true + 1.5, dtype=torch.float64,
                           requires_grad=True)
    opt = torch.optim.Adam([{"params": [s_param], "lr": 1e-3},
                            {"params": [c_param], "lr": 0.05}])
    npix = float(target.numel())

    def step():
        opt.zero_grad(set_to_none=True)
        y = autograd.fbp_differentiable(s_param, c_param, geom)
        loss = (torch.sum((y - target) ** 2)
                + 1e-3 * autograd.total_variation(y)) / npix
        loss.backward()
        opt.step()
        return loss

    first = float(step().item())
    for _ in range(max(args.warmup - 1, 0)):
        step()
    torch.cuda.synchronize()
    nsteps = args.steps if args.steps_given else 200
    t0 = time.perf_counter()
    for _ in range(nsteps):
        loss = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    line = {
        "metric": METRIC, "value": nsteps / dt, "unit": "grad steps/s",
        "n_gpus": 1, "steps": nsteps, "warmup": args.warmup,
        "ms_per_step": dt / nsteps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "fbp_refine_16x901x1024_arc120",
                   "nslice": nsl, "nrow": nrow, "ncol": ncol,
                   "checkpoint": "derivative", "parallelism": "single-gpu"},
        "detail": {"loss_first": first, "loss_last": float(loss.item()),
                   "center_start": c_true + 1.5,
                   "center_end": float(c_param.item()),
                   "center_true": c_true,
                   "updates_per_forward": nsl * nrow * ncol * ncol},
    }
    print(json.dumps(line))


def run_center(args, dev):
    import torch
    from algotom_amd.rec import center
    from algotom_amd.util import phantom
    nrow, ncol, c_true = NROW, NCOL, CENTER
    angles = np.deg2rad(np.linspace(0.0, 180.0, nrow))
    prm = phantom.random_ellipses(ncol, 20, seed=1000)
    p = phantom.ellipse_sinogram(prm, angles, ncol, c_true, device=dev)
    sino = phantom.transmission(p, seed=1000).contiguous()
    torch.cuda.synchronize()
    kw = dict(step=0.5, radius=2, metric="entropy", angles=angles)
    ncoarse = len(np.arange(1260, 1301, 1.0))
    nfine = len(np.arange(0.0, 2 * 2.0 + 0.5, 0.5))
    nslices = ncoarse + 1 + nfine + 1   # + the nmin/nmax reference slices
    for _ in range(args.warmup):
        found = center.find_center_based_slice_metric(sino, 1260, 1300, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        found = center.find_center_based_slice_metric(sino, 1260, 1300, **kw)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    base = None
    if not args.no_cpu_baseline:
        # one candidate of the reference's loop (reconstruction.py:1349-1355:
        # FBP at one centre, then its entropy) with the oracle's C port
        from oracle import center_np, port
        s_host = sino.cpu().numpy()
        pre = center_np.preprocess(s_host, 2, True)
        threads = port.default_threads()
        t0 = time.perf_counter()
        rec0 = port.fbp(pre, c_true, angles=angles, apply_log=False)
        center_np.slice_metric(rec0, "entropy", np.min(rec0), np.max(rec0))
        dt = time.perf_counter() - t0
        base = {"value": 1.0 / dt, "unit": "slices/s", "cores": threads,
                "kind": "port",
                "sample": "one 1801x2560 candidate slice: FBP (C port, "
                          "float64 sums) + histogram entropy"}
    line = {
        "metric": METRIC, "value": nslices / (ms * 1e-3),
        "unit": "slices/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (f64 sums)",
        "data": "synthetic",
        "config": {"workload": "center_search_1801x2560_entropy",
                   "start": 1260, "stop": 1300, "step": 0.5, "radius": 2,
                   "slices_per_search": nslices,
                   "updates_per_search": nslices * ncol * ncol * nrow,
                   "parallelism": "single-gpu"},
        "cpu_baseline": base,
        "detail": {"found": float(found), "true_center": c_true,
                   "gups": nslices * ncol * ncol * nrow / (ms * 1e-3) / 1e9},
    }
    print(json.dumps(line))


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3, out


def run_bpf(args, dev):
    import torch
    from algotom_amd.rec import reconstruction as rec
    nsl = 16
    sino, angles = make_stack(dev, nslice=nsl)
    updates = nsl * NCOL * NCOL * NROW
    ms, out = _timed(lambda: rec.bpf_reconstruction(sino, CENTER, angles),
                     args.steps, args.warmup)
    one = sino[:, 0, :].contiguous()
    ms2d, _ = _timed(lambda: rec.bpf_reconstruction(one, CENTER, angles),
                     args.steps, args.warmup)
    # one 1801 x 4096 sinogram: 2D filter FFT length 16384 (two-level columns)
    from algotom_amd.util import phantom
    wide = 4096
    prm = phantom.random_ellipses(wide, 20, seed=77)
    pw = phantom.ellipse_sinogram(prm, angles, wide, 0.5 * (wide - 1),
                                  device=dev)
    sw = phantom.transmission(pw, seed=77)
    msw, _ = _timed(lambda: rec.bpf_reconstruction(sw, 0.5 * (wide - 1),
                                                   angles),
                    args.steps, args.warmup)
    line = {
        "metric": METRIC, "value": updates / (ms * 1e-3) / 1e9,
        "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (f64 filter)",
        "data": "synthetic",
        "config": {"workload": "bpf_stack_16x1801x2560", "nslice": nsl,
                   "nrow": NROW, "ncol": NCOL, "center": CENTER,
                   "filter_fft_length": 8192, "parallelism": "single-gpu"},
        "cpu_baseline": None,
        "detail": {"ms_per_slice": ms / nsl,
                   "single_sinogram_ms (float64 output)": ms2d,
                   "single_sinogram_1801x4096_ms": msw,
                   "finite": bool(torch.isfinite(out).all().item())},
    }
    print(json.dumps(line))


def run_c1(args, dev):
    """The plain-FBP CPU workload at full size: the committed 721 x 512
    Fourier-slice sinogram of a 20-ellipse phantom (tests/golden/
    c1_fourier_slice.npz), centre 256.0, hann, edge padding, apply_log off,
    the CPU back-projector's arithmetic (gpu=False).  value = GUPS of
    fbp_reconstruction on the device-resident sinogram; cpu_baseline = the
    oracle's C port of the same call on this host's cores (median of 5
    after a warm-up), its parity and both errors against the fixture's
    float64 FBP in detail."""
    import torch
    from algotom_amd.rec import reconstruction as rec
    from oracle import port
    path = os.path.join(ROOT, "tests", "golden", "c1_fourier_slice.npz")
    with np.load(path, allow_pickle=False) as d:
        sino_np, angles, ref64 = d["sinogram"], d["angles"], d["fbp_f64"]
    nrow, ncol = sino_np.shape
    updates = nrow * ncol * ncol
    sino = torch.from_numpy(sino_np).to(dev)
    kw = dict(apply_log=False, gpu=False)
    ms, out = _timed(lambda: rec.fbp_reconstruction(sino, 256.0, angles,
                                                    **kw),
                     max(args.steps, 20), args.warmup)
    got = out.cpu().numpy()
    scale = np.abs(ref64).max()
    detail = {"gpu_max_rel_err_vs_fp64": float(np.abs(got - ref64).max()
                                               / scale)}
    t0 = time.perf_counter()
    rec.fbp_reconstruction(sino_np, 256.0, angles, **kw)
    detail["host_numpy_ms"] = (time.perf_counter() - t0) * 1e3
    base = None
    if not args.no_cpu_baseline:
        threads = port.default_threads()
        dt, times, ref = time_cpu(lambda: port.fbp(sino_np, 256.0, angles,
                                                   nthreads=threads, **kw))
        detail["cpu_max_rel_err_vs_fp64"] = float(np.abs(ref - ref64).max()
                                                  / scale)
        detail["gpu_vs_cpu_port_max_rel"] = float(np.abs(got - ref).max()
                                                  / np.abs(ref).max())
        detail["cpu_port_ms"] = dt * 1e3
        base = {"value": updates / dt / 1e9, "unit": "GUPS",
                "cores": threads, "kind": "port",
                "sample": "the whole workload (721x512, filter+BP+mask+"
                          "scale); median of 5 after a warm-up: %.1f ms "
                          "(%.1f-%.1f ms)" % (dt * 1e3, min(times) * 1e3,
                                             max(times) * 1e3)}
    line = {
        "metric": METRIC, "value": updates / (ms * 1e-3) / 1e9,
        "unit": "GUPS", "n_gpus": 1, "steps": max(args.steps, 20),
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (committed fixture)",
        "config": {"workload": "fbp_c1_721x512_fourier_slice",
                   "nrow": nrow, "ncol": ncol, "center": 256.0,
                   "apply_log": False, "filter": "hann",
                   "arithmetic": "back_projection_cpu (gpu=False)",
                   "parallelism": "single-gpu"},
        "cpu_baseline": base, "detail": detail,
    }
    print(json.dumps(line))


def run_dfi(args, dev):
    """DFI of 16 sinograms 1801 x 2560 (float32 transmission, log on,
    angles=None, hann, pad_rate 0.25: padded width 3841, Bluestein length
    8192); value = slices reconstructed per second."""
    import torch
    from algotom_amd.rec import reconstruction as rec
    nsl = 16
    sino, _ = make_stack(dev, nslice=nsl)
    ms, out = _timed(lambda: rec.dfi_reconstruction(sino, CENTER),
                     args.steps, args.warmup)
    one = sino[:, 0, :].contiguous()
    ms1, out1 = _timed(lambda: rec.dfi_reconstruction(one, CENTER),
                       args.steps, args.warmup)
    # one 1801 x 4096 sinogram: padded width 6145, Bluestein length 16384
    from algotom_amd.util import phantom
    wide = 4096
    prm = phantom.random_ellipses(wide, 20, seed=77)
    pw = phantom.ellipse_sinogram(prm, np.linspace(0.0, np.pi, NROW,
                                                   endpoint=False), wide,
                                  0.5 * (wide - 1), device=dev)
    sw = phantom.transmission(pw, seed=77)
    msw, _ = _timed(lambda: rec.dfi_reconstruction(sw, 0.5 * (wide - 1)),
                    args.steps, args.warmup)
    cpu = None
    detail = {"ms_per_slice": ms / nsl, "single_sinogram_ms": ms1,
              "single_sinogram_1801x4096_ms": msw,
              "finite": bool(torch.isfinite(out).all().item())}
    # host numpy in and out (float64 result): the chunk pipeline
    sino_np = sino.cpu().numpy()
    rec.dfi_reconstruction(sino_np, CENTER)
    t0 = time.perf_counter()
    rec.dfi_reconstruction(sino_np, CENTER)
    detail["host_numpy_ms_per_slice"] = (time.perf_counter() - t0) * 1e3 / nsl
    if not args.no_cpu_baseline:
        from oracle import dfi_np
        s_np = one.cpu().numpy()
        t0 = time.perf_counter()
        ref = dfi_np.scipy_dfi_slice(s_np, CENTER)
        dt = time.perf_counter() - t0
        got = out1.cpu().numpy()
        m = dfi_np.fbp_np.circle_mask(NCOL, 1.0)
        detail["cpu_parity_max_rel"] = float(
            np.abs(got - ref * m).max() / np.abs(ref * m).max())
        cpu = {"value": 1.0 / dt, "unit": "slices/s", "cores": 1,
               "kind": "port",
               "sample": "1 of %d sinograms (%dx%d) through the reference's "
                         "library calls (scipy.ndimage shift and "
                         "map_coordinates, numpy FFTs; reconstruction.py"
                         ":849-895), %.1f s" % (nsl, NROW, NCOL, dt)}
    line = {
        "metric": METRIC, "value": nsl / (ms * 1e-3),
        "unit": "slices/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (f32 input)",
        "data": "synthetic",
        "config": {"workload": "dfi_stack_16x1801x2560", "nslice": nsl,
                   "nrow": NROW, "ncol": NCOL, "center": CENTER,
                   "padded_width": 3841, "bluestein_length": 8192,
                   "parallelism": "single-gpu"},
        "cpu_baseline": cpu, "detail": detail,
    }
    print(json.dumps(line))


def run_vertrec(args, dev):
    import torch
    from algotom_amd.rec import vertrec
    nproj, height, width = NROW, 1024, NCOL
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    proj = torch.rand((nproj, height, width), generator=gen, device=dev)
    proj = proj * 0.5 + 0.25
    cen = CENTER
    ms1, _ = _timed(lambda: vertrec.vertical_reconstruction(
        proj, width // 2 - 100, cen, alpha=30.0, apply_log=True,
        ramp_filter="after", show_progress=False), args.steps, args.warmup)
    nsl = 16
    ms16, out = _timed(lambda: vertrec.vertical_reconstruction_multiple(
        proj, 1000, 1000 + 5 * (nsl - 1), cen, alpha=30.0, step_index=5,
        apply_log=True, ramp_filter="before", show_progress=False),
        args.steps, args.warmup)
    u1 = height * width * nproj
    line = {
        "metric": METRIC, "value": nsl * u1 / (ms16 * 1e-3) / 1e9,
        "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms16,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (f64 chunk sums)", "data": "synthetic",
        "config": {"workload": "vertrec_1801x1024x2560", "nproj": nproj,
                   "height": height, "width": width, "slices": nsl,
                   "chunk_size": 30, "parallelism": "single-gpu"},
        "cpu_baseline": None,
        "detail": {"single_slice_after_filter_ms": ms1,
                   "single_slice_gups": u1 / (ms1 * 1e-3) / 1e9,
                   "multiple_16_before_filter_ms": ms16,
                   "finite": bool(torch.isfinite(out).all().item())},
    }
    # the same 16-line call on the projections as a host numpy array
    # (18.9 GB): groups upload through pinned buffers on a copy stream
    # while the previous group is processed
    del out
    proj_np = proj.cpu().numpy()
    del proj
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    vertrec.vertical_reconstruction_multiple(
        proj_np, 1000, 1000 + 5 * (nsl - 1), cen, alpha=30.0, step_index=5,
        apply_log=True, ramp_filter="before", show_progress=False)
    line["detail"]["multiple_16_host_numpy_ms"] = \
        (time.perf_counter() - t0) * 1e3
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default=None,
                    choices=[None, "fp32", "fp16", "bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-numpy round trip (profiling runs)")
    ap.add_argument("--workload", default="stack",
                    choices=["stack", "c1", "uneven", "refine", "volume",
                             "volume_refine",
                             "center", "bpf", "vertrec", "dfi"])
    args = ap.parse_args()
    args.steps_given = any(a.startswith("--steps") for a in sys.argv[1:])
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path "
              "(slices would shard across GPUs with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from algotom_amd import _native as nat
    from algotom_amd.rec import reconstruction as rec

    dev = nat.device()
    if args.workload == "uneven":
        run_uneven(args, dev)
        return
    if args.workload == "refine":
        run_refine(args, dev)
        return
    if args.workload == "volume_refine":
        run_volume_refine(args, dev)
        return
    if args.workload == "volume":
        run_volume(args, dev)
        return
    if args.workload == "center":
        run_center(args, dev)
        return
    if args.workload == "dfi":
        return run_dfi(args, dev)
    if args.workload == "c1":
        return run_c1(args, dev)
    if args.workload == "bpf":
        run_bpf(args, dev)
        return
    if args.workload == "vertrec":
        run_vertrec(args, dev)
        return
    sino, angles = make_stack(dev)
    updates = NSLICE * NCOL * NCOL * NROW
    pad = min(int(0.15 * NCOL), 150)
    taps = rec._window_taps(None, NROW, NCOL + 2 * pad, "hann")
    acc = nat.ACC_F32SEQ if args.precision is None else nat.ACC_F32
    store = rec._storage_dtype(args.precision)
    mask_radius = float(1.0 * (NCOL // 2))

    def step(events=None):
        # exactly fbp_reconstruction's device path, split so the
        # back-projection kernel can be bracketed by events
        rec._clamp_nonpositive(sino)
        filt = rec._filter_stack(sino, pad, "edge", True, taps, store)
        if events:
            events[0].record()
        out = rec._backproject_stack(filt, angles, CENTER, acc,
                                     mask_radius=mask_radius,
                                     apply_scale=True)
        if events:
            events[1].record()
        return out

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        out = step(ev[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    bp_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    ms = dt / args.steps * 1e3
    gather_gbs = updates * 8.0 / (bp_ms * 1e-3) / 1e9

    copy_gbs = copy_bandwidth(dev)
    host_ms = None if args.no_host else host_roundtrip(sino, angles)
    base, err = (None, None)
    if not args.no_cpu_baseline:
        base, err = cpu_baseline(sino, angles, out)
    cnt = load_counters() if args.precision is None else None
    roof = {"bound": "valu", "achieved": None, "peak": PEAK_VALU_TLANE,
            "unit": "T VALU lane-op/s", "frac": None, "traffic": None}
    if cnt:
        roof["achieved"] = (cnt["valu_insts_per_launch"] * 64.0
                            / (bp_ms * 1e-3) / 1e12)
        roof["frac"] = roof["achieved"] / PEAK_VALU_TLANE
        roof["traffic"] = cnt["traffic_bytes"]
    compulsory = 4.0 * NROW * NSLICE * NCOL + 4.0 * NCOL * NSLICE * NCOL
    line = {
        "metric": METRIC,
        "value": updates / (ms * 1e-3) / 1e9,
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision is None else
                 "%s-storage/fp32-acc" % args.precision,
        "data": "synthetic",
        "config": {"workload": "fbp_stack_64x1801x2560", "nslice": NSLICE,
                   "nrow": NROW, "ncol": NCOL, "center": CENTER,
                   "apply_log": True, "filter": "hann", "pad": pad,
                   "precision": args.precision or "reference",
                   "parallelism": "single-gpu"},
        "roofline": roof,
        "cpu_baseline": base,
        "detail": {"bp_kernel_ms": bp_ms, "filter_and_rest_ms": ms - bp_ms,
                   "hbm_copy_GBs": copy_gbs,
                   "bp_valu_lane_ops_per_update":
                       cnt["valu_lane_ops_per_update"] if cnt else None,
                   "bp_hbm_GBs": (cnt["traffic_bytes"] / (bp_ms * 1e-3) / 1e9
                                  if cnt else None),
                   "bp_hbm_frac_of_peak": (cnt["traffic_bytes"]
                                           / (bp_ms * 1e-3) / 1e9
                                           / PEAK_HBM_GBS if cnt else None),
                   "compulsory_bytes": compulsory,
                   "compulsory_frac_of_hbm_peak":
                       compulsory / (bp_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                   "gather_byte_model_GBs": gather_gbs,
                   "bp_gups": updates / (bp_ms * 1e-3) / 1e9,
                   "host_numpy_roundtrip_ms": host_ms,
                   "host_numpy_roundtrip_gups":
                       updates / (host_ms * 1e-3) / 1e9 if host_ms else None,
                   "max_rel_err_vs_oracle_sample": err},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
