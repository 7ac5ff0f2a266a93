"""Writes the committed fixtures under tests/golden/ (run from the repo root):

    python tests/golden/make_golden.py

windows.npz          scipy.signal.windows.<name>(width) (symmetric), the
                     third-party routine reference.py:67-80 calls; keys
                     "<name>_<width>".
reference_case.npz   the reference's FBP test input
                     (tests/test_rec/test_reconstruction.py:50-65) and the
                     oracle's outputs for it (2D float64 accumulation and
                     float32-per-angle accumulation, ratio 1.0), so a GPU run
                     can check without scipy.ndimage.
"""

import os
import sys

import numpy as np
from scipy.signal import windows as sw

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NAMES = ["hann", "bartlett", "blackman", "hamming", "nuttall", "parzen",
         "triang"]
WIDTHS = [5, 64, 65, 82, 664, 1324, 2860]


def main():
    wins = {"%s_%d" % (n, w): getattr(sw, n)(w) for n in NAMES for w in WIDTHS}
    np.savez_compressed(os.path.join(HERE, "windows.npz"), **wins)

    from __graft_entry__ import _reference_test_case
    from oracle import port
    mat, sino, angles, center = _reference_test_case()
    f64 = port.fbp(sino, center, angles, apply_log=False, gpu=True)
    f32 = port.fbp(sino, center, angles, apply_log=False, gpu=False)
    np.savez_compressed(os.path.join(HERE, "reference_case.npz"),
                        phantom=mat, sinogram=sino, angles=angles,
                        center=np.float64(center), fbp_acc_f64=f64,
                        fbp_acc_f32seq=f32)


if __name__ == "__main__":
    main()
