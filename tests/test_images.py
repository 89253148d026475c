"""Image output on the CPU: the oracle's byte conversion pinned by hand-computed values,
and the library's PNG writer (host code, dr_png_write) read back by an independent
decoder (Pillow)."""
import ctypes

import numpy as np
import pytest

import dumbray_amd as D
import oracle


def test_oracle_frame_bytes_known_values():
    c = np.array([0.0, 1.0, 0.5, -0.25, 2.0, np.nan, 1.0 / 255.0, 0.99999, 254.9 / 255.0, np.inf, -np.inf, 3e-3],
                 np.float32)
    # 255*0.5 = 127.5 -> 127; 255*(1/255 in fp32) = 1.0000000 -> 1; 0.99999 -> 254.99745 -> 254;
    # 254.9/255 -> 254.9 -> 254; 3e-3 -> 0.765 -> 0
    want = np.array([0, 255, 127, 0, 255, 0, 1, 254, 254, 255, 0, 0], np.uint8)
    np.testing.assert_array_equal(oracle.frame_bytes(c), want)


def test_oracle_blend_is_running_mean_for_constant_frames():
    f = None
    imgs = [np.full((2, 2, 4), v, np.float32) for v in (0.25, 0.75, 0.5, 1.5)]
    for k, im in enumerate(imgs, 1):
        f = oracle.blend(f, im, k)
    np.testing.assert_allclose(f, np.full((2, 2, 4), 0.75), rtol=1e-6)


@pytest.mark.parametrize("w,h", [(1, 1), (13, 7), (64, 36)])
def test_png_writer_round_trip(tmp_path, w, h):
    from PIL import Image

    rng = np.random.default_rng(w * 100 + h)
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    path = str(tmp_path / "x.png")
    assert D.lib().dr_png_write(path.encode(), rgba.ctypes.data_as(ctypes.c_void_p), w, h) == 0
    im = Image.open(path)
    assert im.mode == "RGBA" and im.size == (w, h)
    np.testing.assert_array_equal(np.asarray(im), rgba)


def test_png_writer_rejects_bad_arguments(tmp_path):
    rgba = np.zeros((2, 2, 4), np.uint8)
    assert D.lib().dr_png_write(str(tmp_path / "x.png").encode(), rgba.ctypes.data_as(ctypes.c_void_p), 0, 2) != 0
    assert D.lib().dr_png_write(b"/nonexistent_dir/x.png", rgba.ctypes.data_as(ctypes.c_void_p), 2, 2) != 0
