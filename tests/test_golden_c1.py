"""The committed plain-FBP CPU workload (tests/golden/c1_fourier_slice.npz,
written by tests/golden/make_c1.py): pinned by what the mathematics fixes.

* the Fourier-slice projector (simulation.py:330-380 restated) sends a
  centred disk to its closed-form chord lengths 2 sqrt(R^2 - t^2), and every
  projection of the phantom keeps the phantom's mass (the zero-frequency
  sample lies on every radial line);
* the fixture is what the generator writes (shapes, types, values);
* its float64 FBP reconstructs the phantom within the reference test's
  threshold (tests/test_rec/test_reconstruction.py:104, median <= 0.2),
  in fact to 2e-3, and matches the C port's reference arithmetic.
"""

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_c1  # noqa: E402

FIXTURE = os.path.join(HERE, "golden", "c1_fourier_slice.npz")


@pytest.fixture(scope="module")
def c1():
    with np.load(FIXTURE, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_fourier_slice_disk_closed_form():
    size, radius = 128, 30.0
    ic = size // 2
    y, x = np.mgrid[:size, :size]
    disk = (((x - ic) ** 2 + (y - ic) ** 2) <= radius ** 2).astype(
        np.float32)
    angles = np.linspace(0.0, np.pi, 12, endpoint=False)
    sino = make_c1.fourier_slice_sinogram(disk, angles)
    t = np.arange(size) - ic
    chord = 2.0 * np.sqrt(np.clip(radius ** 2 - t ** 2, 0.0, None))
    # a pixelated disk against the exact chord: edge pixels differ by O(1)
    assert np.abs(sino - chord).mean() <= 0.05 * chord.max()
    assert np.abs(sino[:, ic] - chord[ic]).max() <= 0.03 * chord[ic]
    # the DC sample is exact on every line; the crop drops a ringing tail
    assert np.allclose(sino.sum(axis=1), disk.sum(), rtol=1e-5)


def test_fixture_is_the_generator_output(c1):
    assert c1["sinogram"].shape == (721, 512)
    assert c1["sinogram"].dtype == np.float32
    assert c1["phantom"].dtype == np.float32
    assert c1["fbp_f64"].dtype == np.float64
    assert float(c1["center"]) == 256.0
    assert np.array_equal(c1["angles"],
                          np.linspace(0.0, np.pi, 721, endpoint=False))
    mat = make_c1.phantom()
    assert np.array_equal(mat, c1["phantom"])
    sino = make_c1.fourier_slice_sinogram(mat, c1["angles"]).astype(
        np.float32)
    assert np.abs(sino - c1["sinogram"]).max() <= 1e-6 * np.abs(sino).max()


def test_projections_keep_the_mass(c1):
    mass = float(c1["phantom"].astype(np.float64).sum())
    sums = c1["sinogram"].astype(np.float64).sum(axis=1)
    # exact on the uncropped lines; the crop drops ringing tails (~2e-5)
    assert np.abs(sums - mass).max() <= 1e-4 * mass


def test_fbp_reconstructs_the_phantom(c1):
    from oracle import fbp_np, port
    f64 = c1["fbp_f64"]
    mask = fbp_np.circle_mask(512, 1.0).astype(bool)
    diff = np.abs(f64 - c1["phantom"])[mask]
    assert np.median(diff) <= 2e-3
    ref = port.fbp(c1["sinogram"], 256.0, c1["angles"], apply_log=False,
                   gpu=True)
    assert np.abs(ref - f64).max() <= 1e-5 * np.abs(f64).max()
