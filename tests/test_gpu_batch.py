"""Batched uneven FBP (one filter launch per FFT length, one persistent
back-projection launch, optional HIP graph).

Bar: every sinogram of the batch is bit-identical to fbp_reconstruction on
the same input and arguments (same arithmetic per pixel), for the reference
modes and the reduced-precision modes; graph replays with eager work and new
inputs in between each match the eager result; autotuning changes timings,
not bits.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch_inputs(n, seed, widths=(64, 100, 257, 512), nangles=(37, 90, 181)):
    rng = np.random.default_rng(seed)
    sinos, centers, angles = [], [], []
    for _ in range(n):
        ncol = int(rng.choice(widths))
        nrow = int(rng.choice(nangles))
        centers.append(float((ncol - 1) / 2 + rng.uniform(-5, 5)))
        angles.append(np.linspace(0.0, np.pi, nrow, endpoint=False))
        sinos.append(rng.uniform(0.05, 1.0, (nrow, ncol)).astype(np.float32))
    return sinos, centers, angles


@pytest.mark.parametrize("precision", [None, "fp32", "fp16", "bf16"])
def test_batch_matches_single_calls(precision):
    from algotom_amd.rec import batch, reconstruction as rec
    sinos, centers, angles = _batch_inputs(12, 1)
    plan = batch.BatchFBP([s.shape for s in sinos], centers, angles,
                          apply_log=True, precision=precision)
    plan.load(sinos)
    plan.run()
    got = plan.results()
    for s, c, a, g in zip(sinos, centers, angles, got):
        ref = rec.fbp_reconstruction(s.copy(), c, a, apply_log=True,
                                     precision=precision)
        assert np.array_equal(g, ref), (s.shape, c)


def test_batch_with_fft_32768_rows():
    # a width past the 16384-point filter joins the batch: its rows take the
    # batched two-half kernel and match the single call bit for bit
    from algotom_amd.rec import batch, reconstruction as rec
    rng = np.random.default_rng(9)
    shapes = [(31, 8100), (45, 300), (29, 9001)]
    sinos = [rng.uniform(0.05, 1.0, s).astype(np.float32) for s in shapes]
    centers = [4049.7, 151.2, 4500.0]
    angles = [np.linspace(0.0, np.pi, s[0], endpoint=False) for s in shapes]
    plan = batch.BatchFBP(shapes, centers, angles, apply_log=True)
    plan.load(sinos)
    plan.run()
    for s, c, a, g in zip(sinos, centers, angles, plan.results()):
        ref = rec.fbp_reconstruction(s.copy(), c, a, apply_log=True)
        assert np.array_equal(g, ref), s.shape


@pytest.mark.parametrize("gpu", [True, False])
def test_batch_stacks_and_2d_mixed(gpu):
    # each member gets fbp_reconstruction's own arithmetic: float64 sums for
    # a single sinogram with gpu=True, float32 per angle for stacks; a
    # float64 sinogram stays float64 through the log and the filter
    from algotom_amd.rec import batch, reconstruction as rec
    rng = np.random.default_rng(2)
    sinos = [rng.uniform(0.1, 1, (40, 3, 70)).astype(np.float32),
             rng.uniform(0.1, 1, (25, 90)).astype(np.float32),
             rng.uniform(0.1, 1, (60, 5, 64)).astype(np.float32),
             rng.uniform(0.1, 1, (33, 81)),
             rng.uniform(0.1, 1, (33, 2, 81))]
    centers = [34.2, 44.0, 31.5, 40.3, 39.6]
    got = batch.fbp_reconstruction_batch([s.copy() for s in sinos], centers,
                                         apply_log=True, gpu=gpu)
    for s, c, g in zip(sinos, centers, got):
        ref = rec.fbp_reconstruction(s.copy(), c, apply_log=True, gpu=gpu)
        assert g.shape == ref.shape
        assert np.array_equal(g, ref), (s.shape, s.dtype, gpu)


def test_graph_replays_match_eager():
    import torch
    from algotom_amd.rec import batch
    sinos, centers, angles = _batch_inputs(9, 3)
    plan = batch.BatchFBP([s.shape for s in sinos], centers, angles,
                          apply_log=True)
    plan.load(sinos)
    plan.launch()
    torch.cuda.synchronize()
    eager = plan.results()
    plan.capture()
    other, _, _ = _batch_inputs(9, 3)
    for rep in range(4):
        # eager work between replays: new inputs, a different computation
        data = sinos if rep % 2 == 0 else [o * 0.5 + 0.2 for o in other]
        plan.load(data)
        junk = torch.randn(1 << 20, device="cuda")
        junk.mul_(2.0).sum().item()
        plan.run()
        torch.cuda.synchronize()
        got = plan.results()
        if rep % 2 == 0:
            for g, e in zip(got, eager):
                assert np.array_equal(g, e)
        else:
            plan.graph, saved = None, plan.graph
            plan.launch()
            torch.cuda.synchronize()
            ref = plan.results()
            plan.graph = saved
            for g, e in zip(got, ref):
                assert np.array_equal(g, e)


def test_autotune_changes_speed_not_bits(tmp_path, monkeypatch):
    import torch
    from algotom_amd.rec import autotune, batch
    monkeypatch.setenv("ALGOTOM_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    autotune.reset()
    sinos, centers, angles = _batch_inputs(6, 4)
    plan = batch.BatchFBP([s.shape for s in sinos], centers, angles,
                          apply_log=True, tile_rows=16)
    plan.load(sinos)
    plan.run()
    torch.cuda.synchronize()
    before = plan.results()
    best, timings = plan.autotune()
    assert len(timings) == 3 and best in timings
    assert autotune.lookup(plan.tune_key) == best
    plan.run()
    torch.cuda.synchronize()
    for a, b in zip(before, plan.results()):
        assert np.array_equal(a, b)
    autotune.reset()
