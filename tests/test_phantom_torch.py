"""The torch form of the ellipse phantom (used to build large targets on the
device) equals the numpy form exactly."""

import numpy as np

from algotom_amd.util import phantom


def test_ellipse_image_torch_matches_numpy():
    for seed, size in [(0, 97), (1, 128), (2, 255)]:
        prm = phantom.random_ellipses(size, 20, seed=seed)
        want = phantom.ellipse_image(prm, size)
        got = phantom.ellipse_image_torch(prm, size).numpy()
        assert got.dtype == np.float32
        assert np.array_equal(got, want)
