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


# ---------------------------------------------------------------- PNG reading (dr_png_read)
def _read(path):
    w, h = ctypes.c_int(), ctypes.c_int()
    assert D.lib().dr_png_info(path.encode(), ctypes.byref(w), ctypes.byref(h)) == 0, D.lib().dr_last_error()
    out = np.zeros((h.value, w.value, 4), np.uint8)
    assert D.lib().dr_png_read(path.encode(), out.ctypes.data_as(ctypes.c_void_p), w.value, h.value) == 0
    return out


@pytest.mark.parametrize("mode", ["RGBA", "RGB", "L", "LA", "P", "P+tRNS", "1"])
def test_png_reader_matches_pillow(tmp_path, mode):
    from PIL import Image

    rng = np.random.default_rng(hash(mode) % 1000)
    rgba = rng.integers(0, 256, (13, 17, 4), dtype=np.uint8)
    im = Image.fromarray(rgba, "RGBA")
    if mode == "P+tRNS":
        im = im.convert("RGB").quantize(16)
        im.info["transparency"] = bytes([0, 128, 255] + [200] * 13)
        kw = {"transparency": im.info["transparency"]}
    else:
        im = im.convert(mode)
        kw = {}
    path = str(tmp_path / "t.png")
    im.save(path, **kw)
    want = np.asarray(Image.open(path).convert("RGBA"))
    np.testing.assert_array_equal(_read(path), want)


def _chunk(t, data):
    import struct
    import zlib
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xffffffff)


def _png(w, h, depth, ctype, rows_bytes, interlace=0):
    import struct
    import zlib
    ihdr = struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace)
    return b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(rows_bytes)) + _chunk(b"IEND", b"")


def test_png_reader_16bit_keeps_the_high_byte(tmp_path):
    rng = np.random.default_rng(7)
    px = rng.integers(0, 65536, (5, 6, 3), dtype=np.uint16)
    raw = b"".join(b"\x00" + px[y].astype(">u2").tobytes() for y in range(5))
    p = tmp_path / "x.png"
    p.write_bytes(_png(6, 5, 16, 2, raw))
    got = _read(str(p))
    np.testing.assert_array_equal(got[..., :3], (px >> 8).astype(np.uint8))
    assert np.all(got[..., 3] == 255)


def test_png_reader_filters_and_adam7(tmp_path):
    """Every scanline filter (Sub, Up, Average, Paeth) and Adam7 interlacing, written here
    with their textbook definitions, decode to the original pixels."""
    rng = np.random.default_rng(3)
    W, H = 11, 9
    img = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)

    def pae(a, b, c):
        p = a + b - c
        pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
        return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)

    def filtered(sub, ftypes):
        h, w = sub.shape[:2]
        out, prev = b"", np.zeros(w * 4, np.int64)
        for y in range(h):
            row = sub[y].reshape(-1).astype(np.int64)
            ft = ftypes[y % len(ftypes)]
            enc = np.zeros_like(row)
            for i in range(len(row)):
                a = row[i - 4] if i >= 4 else 0
                b = prev[i]
                c = prev[i - 4] if i >= 4 else 0
                pred = [0, a, b, (a + b) // 2, pae(a, b, c)][ft]
                enc[i] = (row[i] - pred) % 256
            out += bytes([ft]) + enc.astype(np.uint8).tobytes()
            prev = row
        return out

    p = tmp_path / "f.png"
    p.write_bytes(_png(W, H, 8, 6, filtered(img, [0, 1, 2, 3, 4])))
    np.testing.assert_array_equal(_read(str(p)), img)
    # Adam7: passes of (x0, y0, dx, dy)
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
    raw = b""
    for x0, y0, dx, dy in passes:
        sub = img[y0::dy, x0::dx]
        if sub.size:
            raw += filtered(sub, [4, 3, 2, 1, 0])
    p.write_bytes(_png(W, H, 8, 6, raw, interlace=1))
    np.testing.assert_array_equal(_read(str(p)), img)


def test_png_reader_rejects_garbage(tmp_path):
    p = tmp_path / "bad.png"
    p.write_bytes(b"not a png")
    w, h = ctypes.c_int(), ctypes.c_int()
    assert D.lib().dr_png_info(str(p).encode(), ctypes.byref(w), ctypes.byref(h)) != 0


@pytest.mark.parametrize("case", ["palette_depth0", "grey_depth3", "too_many_pixels", "empty_idat", "short_rows",
                                  "zero_length_tail"])
def test_png_reader_rejects_malformed_headers_and_data(tmp_path, case):
    """Malformed inputs found by fuzzing the reader: each is an error, not a crash."""
    import zlib
    rows = b"\x00" + b"\x01\x02"           # one 2-pixel row of palette indices (depth 8)
    plte = _chunk(b"PLTE", bytes(range(9)))
    sig, end = b"\x89PNG\r\n\x1a\n", _chunk(b"IEND", b"")

    def hdr(w, h, depth, ctype):
        import struct
        return _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    data = {
        "palette_depth0": sig + hdr(2, 1, 0, 3) + plte + _chunk(b"IDAT", zlib.compress(rows)) + end,
        "grey_depth3": sig + hdr(2, 1, 3, 0) + _chunk(b"IDAT", zlib.compress(rows)) + end,
        "too_many_pixels": sig + hdr(65536, 65536, 8, 0) + _chunk(b"IDAT", zlib.compress(b"")) + end,
        "empty_idat": sig + hdr(2, 1, 8, 3) + plte + _chunk(b"IDAT", b"") + end,
        "short_rows": sig + hdr(2, 4, 8, 3) + plte + _chunk(b"IDAT", zlib.compress(rows)) + end,
        "zero_length_tail": sig + hdr(2, 1, 8, 7) + _chunk(b"tEXt", b""),
    }[case]
    p = tmp_path / "m.png"
    p.write_bytes(data)
    w, h = ctypes.c_int(), ctypes.c_int()
    assert D.lib().dr_png_info(str(p).encode(), ctypes.byref(w), ctypes.byref(h)) != 0
    assert D.lib().dr_last_error()


def test_png_reader_large_compressible_image(tmp_path):
    """A 1500 x 1100 image of one colour with a few odd pixels (6.6 MB of scanlines from a
    few KB of deflate data) decodes to Pillow's pixels."""
    from PIL import Image
    img = np.zeros((1100, 1500, 4), np.uint8)
    img[...] = (10, 200, 30, 255)
    img[7, 3] = (1, 2, 3, 4)
    img[1099, 1499] = (250, 251, 252, 253)
    path = str(tmp_path / "big.png")
    Image.fromarray(img, "RGBA").save(path)
    np.testing.assert_array_equal(_read(path), img)
