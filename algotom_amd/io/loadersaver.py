"""The few algotom.io.loadersaver helpers the reconstruction path writes its
results with (algotom/io/loadersaver.py): folder naming and 2D image
output (TIFF keeps the array's float type; other formats are scaled to
uint8).  Host-side file I/O, no GPU work."""

from pathlib import Path

import numpy as np


def make_folder(file_path):
    """Create the folder of ``file_path`` (or the folder itself, for a path
    without a suffix) if it does not exist (loadersaver.py:275-294)."""
    path = Path(file_path).resolve()
    folder_path = path.parent if path.suffix else path
    if not folder_path.exists():
        try:
            folder_path.mkdir(parents=True, exist_ok=True)
        except Exception as e:
            raise ValueError(f"Can't create : {folder_path}. Error: {e}")


def make_folder_name(folder_path, name_prefix="Output", zero_prefix=5):
    """A new folder name under ``folder_path`` that does not exist yet:
    ``<name_prefix>_<n>`` with n zero-padded to ``zero_prefix`` digits,
    starting after the number of existing ``<name_prefix>_*`` entries
    (loadersaver.py:297-326)."""
    folder_path = Path(folder_path)
    prefix = name_prefix + "_"
    num = len(list(folder_path.glob(f"{prefix}*"))) + 1
    name = prefix + f"{num:0{zero_prefix}}"
    while (folder_path / name).is_dir():
        num += 1
        name = prefix + f"{num:0{zero_prefix}}"
    return name


def make_file_name(file_path):
    """``file_path``, or ``<stem>_NNNN<suffix>`` with the first free number
    if it exists (loadersaver.py:329-359)."""
    file_path = Path(file_path)
    if file_path.exists():
        n = 0
        while True:
            cand = file_path.parent / f"{file_path.stem}_{n:04d}{file_path.suffix}"
            if not cand.exists():
                return str(cand)
            n += 1
    return str(file_path)


def save_image(file_path, mat, overwrite=True):
    """Save a 2D array (numpy, or a torch tensor, copied to the host) as an
    image (loadersaver.py:382-418): TIFF keeps the dtype, other formats get
    uint8(255 (mat - min) / (max - min)).  Returns the path written."""
    from PIL import Image
    if hasattr(mat, "detach"):
        mat = mat.detach().cpu().numpy()
    mat = np.asarray(mat)
    file_path = Path(file_path).resolve()
    if file_path.suffix not in (".tif", ".tiff"):
        mat = np.uint8(255.0 * (mat - np.min(mat)) /
                       (np.max(mat) - np.min(mat)))
    elif "complex" in str(mat.dtype):
        raise ValueError(f"Can't save to tiff with format: {mat.dtype}")
    image = Image.fromarray(mat)
    if not overwrite:
        file_path = Path(make_file_name(file_path))
    make_folder(file_path)
    try:
        image.save(file_path)
    except Exception as error:
        raise ValueError(f"Couldn't write to file: {file_path}. "
                         f"Error {error}")
    return str(file_path)
