"""algotom_amd.io.loadersaver (algotom/io/loadersaver.py:275-418) -- CPU."""

import numpy as np

from algotom_amd.io import loadersaver as losa


def test_make_folder_name_skips_existing(tmp_path):
    assert losa.make_folder_name(tmp_path, "Find_center", 3) == \
        "Find_center_001"
    (tmp_path / "Find_center_001").mkdir()
    (tmp_path / "Find_center_003").mkdir()
    # two entries match the prefix: start at 3, which exists, then 4
    assert losa.make_folder_name(tmp_path, "Find_center", 3) == \
        "Find_center_004"
    assert losa.make_folder_name(tmp_path) == "Output_00001"


def test_save_image_tif_keeps_float_and_png_scales(tmp_path):
    from PIL import Image
    a = np.random.default_rng(0).random((6, 9)).astype(np.float32)
    p = losa.save_image(str(tmp_path / "sub" / "x.tif"), a)
    assert np.array_equal(np.asarray(Image.open(p)), a)
    q = losa.save_image(str(tmp_path / "x.png"), a)
    want = np.uint8(255.0 * (a - a.min()) / (a.max() - a.min()))
    assert np.array_equal(np.asarray(Image.open(q)), want)
    r = losa.save_image(str(tmp_path / "sub" / "x.tif"), a, overwrite=False)
    assert r.endswith("x_0000.tif")
