"""Host helpers of the FBP path mirrored from algotom/util/utility.py."""

import numpy as np


def make_circle_mask(width, ratio):
    """Create a circle mask (algotom/util/utility.py:160-182).

    Parameters
    -----------
    width : int
        Width of a square array.
    ratio : float
        Ratio between the diameter of the mask and the width of the array.

    Returns
    ------
    array_like
         Square float32 array: 1 where x^2 + y^2 <= (ratio * (width//2))^2
         on the integer grid centred at width//2, else 0.  The GPU
         back-projector applies the same test in its store.
    """
    c = width // 2
    radius = ratio * c
    y, x = np.ogrid[-c:width - c, -c:width - c]
    mask = np.zeros((width, width), dtype=np.float32)
    mask[x * x + y * y <= radius * radius] = 1.0
    return mask
