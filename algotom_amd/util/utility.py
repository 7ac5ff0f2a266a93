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


def mapping(mat, x_mat, y_mat, order=1, mode="reflect"):
    """Resample a 2D array at (row = y_mat, column = x_mat) with
    scipy.ndimage.map_coordinates (algotom/util/utility.py:131-157).
    Host-side: used to build filter windows."""
    import scipy.ndimage as ndi
    coords = np.stack((np.ravel(y_mat), np.ravel(x_mat)))
    return ndi.map_coordinates(mat, coords, order=order,
                               mode=mode).reshape(np.shape(x_mat))


def transform_1d_window_to_2d(win_1d, order=1, mode="reflect"):
    """Turn a 1d window into a radially symmetric 2d window
    (algotom/util/utility.py:925-963): the value at radius r (float32) is
    the 1d window interpolated at centre + r, with the sign of r taken from
    the polar angle so both halves of the 1d window are used.

    Returns a square (len(win_1d), len(win_1d)) float64 array.
    """
    n0 = len(win_1d)
    n = n0 + 1 if n0 % 2 == 0 else n0
    c = n // 2
    xs = 1.0 * np.flipud(np.arange(n)) - c
    ys = 1.0 * np.arange(n) - c
    x_mat, y_mat = np.meshgrid(xs, ys)
    radius = np.float32(np.clip(np.sqrt(x_mat ** 2 + y_mat ** 2), 0, c))
    theta = np.arctan2(y_mat, x_mat)
    radius[theta < 0] *= -1
    radius = np.float32(np.clip(radius + c, 0, n - 1))
    theta = np.clip(np.float32(theta * (n - 1.0) / np.pi), 0, n - 1)
    rows = np.tile(win_1d, (n, 1))
    return mapping(rows, radius, theta, order=order, mode=mode)[0:n0, 0:n0]
