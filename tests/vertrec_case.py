"""The synthetic case of the reference's vertrec tests
(tests/test_rec/test_vertrec.py:34-81): a 40 x 61 x 61 phantom (a blurred
disk with three square holes, empty first row) projected over 91 angles
in [0, 360] by rotating and summing; the first 46 projections cover 180
degrees."""

import functools

import numpy as np


@functools.lru_cache(maxsize=1)
def vertrec_case():
    import scipy.ndimage as ndi
    from algotom_amd.util import utility
    width, height = 61, 40
    mat = utility.make_circle_mask(width, 0.8)
    for rad, i_s, j_s in ((10, 30, 20), (5, 20, 30), (5, 20, 50)):
        mat[i_s:i_s + rad, j_s:j_s + rad] = 0
    mat = ndi.gaussian_filter(mat, 1)
    phantom = np.zeros((height, width, width), dtype=np.float32)
    phantom[1:] = mat
    n360 = 91
    deg = np.linspace(0.0, 360.0, n360, dtype=np.float32)
    projs = np.zeros((n360, height, width), dtype=np.float32)
    for i, a in enumerate(deg):
        projs[i] = np.sum(ndi.rotate(phantom, -a, axes=(1, 2), order=1,
                                     reshape=False), axis=1)
    angles_360 = np.deg2rad(deg)
    n = n360 // 2 + 1
    return {"phantom": phantom, "projs_360": projs, "angles_360": angles_360,
            "projs": projs[:n], "angles": angles_360[:n], "num_proj": n,
            "center": width // 2, "width": width, "height": height}
