"""Image writers (saveImage, main.cpp:86-120) against the reference's own
ImageSaver / stb_image_write.

Golden: tests/golden/image_golden.json holds the size and checksum of every
file the reference writes (flipped on write, as main sets it) for the images
of tests/golden/image_cases.py -- noise with varying alpha, gradients, flat
runs, a render-like image, a 1000-pixel-wide strip; the bytes of the PNG
(filter choice + zlib stream), BMP (pink alpha compositing), TGA (run-length
packets), JPEG (quality 100) and PFM written here must be identical.
Independent checks: decoding the PNG / BMP / TGA with Pillow gives back the
8-bit truncation floor(255 f) of the film, rows bottom-up; the PFM holds the
floats.
"""
import io
import json
import pathlib
import sys

import numpy as np
import pytest

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLD))
from image_cases import image_cases, to_u8  # noqa: E402


def checksum(b):
    b = bytes(b) + b"\0" * (-len(b) % 4)
    w = np.frombuffer(b, "<u4").astype(np.uint64)
    return "%016x" % int((w * (np.arange(len(w), dtype=np.uint64) * np.uint64(2) + np.uint64(1))).sum(dtype=np.uint64))


CASES = image_cases()
GOLDEN = {e["name"]: e for e in json.load(open(GOLD / "image_golden.json"))}


@pytest.mark.parametrize("ext", ["png", "bmp", "tga", "jpg", "pfm"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_bytes_match_reference_writer(pkg, name, ext):
    b = pkg.encode_image(ext, CASES[name])
    want = GOLDEN[name][ext]
    assert (len(b), checksum(b)) == (want[0], want[1])


@pytest.mark.parametrize("ext", ["png", "bmp", "tga"])
@pytest.mark.parametrize("name", ["noise_alpha", "render", "flat_runs"])
def test_decoded_pixels(pkg, name, ext):
    from PIL import Image
    img = CASES[name]
    im = Image.open(io.BytesIO(pkg.encode_image(ext, img)))
    u8 = to_u8(img)[::-1]                           # rows bottom-up
    got = np.asarray(im.convert("RGBA" if ext != "bmp" else "RGB"))
    if ext == "bmp":                                # composited over (255, 0, 255) by alpha
        bg = np.array([255, 0, 255])
        want = bg + np.trunc((u8[..., :3].astype(np.int64) - bg) * u8[..., 3:].astype(np.int64) / 255.0).astype(np.int64)
        assert np.array_equal(got, want.astype(np.uint8))
    else:
        assert np.array_equal(got, u8)


def test_pfm_and_jpeg_decode(pkg):
    from PIL import Image
    img = CASES["render"]
    b = pkg.encode_image("pfm", img)
    head = b"PF\n300 200\n-1.000000\n"
    assert b.startswith(head)
    assert np.array_equal(np.frombuffer(b[len(head):], "<f4").reshape(200, 300, 3), img[..., :3])
    j = np.asarray(Image.open(io.BytesIO(pkg.encode_image("jpg", img))).convert("RGB")).astype(np.int64)
    assert np.abs(j - to_u8(img)[::-1, :, :3]).mean() < 2.0     # quality 100, 4:4:4


def test_truncation_and_errors(pkg, tmp_path):
    v = np.array([0.0, 0.5 / 255, 1.0 / 255, 0.999, 1.0], np.float32)
    img = np.ones((1, 5, 4), np.float32)
    img[0, :, 0] = v
    b = pkg.encode_image("tga", img)
    from PIL import Image
    got = np.asarray(Image.open(io.BytesIO(b)).convert("RGBA"))[0, :, 0]
    assert got.tolist() == [0, 0, 1, 254, 255]
    p = tmp_path / "x.png"
    pkg.save_image(str(p), img)
    assert p.read_bytes() == pkg.encode_image("png", img)
    with pytest.raises(RuntimeError, match="extension is not supported"):
        pkg.save_image(str(tmp_path / "x.gif"), img)
    with pytest.raises(RuntimeError, match="extension is not supported"):
        pkg.save_image(str(tmp_path / "x.PNG"), img)
    with pytest.raises(RuntimeError, match="-8"):
        pkg.save_image(str(tmp_path / "missing_dir" / "x.png"), img)
