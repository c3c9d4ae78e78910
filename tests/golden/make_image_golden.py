#!/usr/bin/env python3
"""Regenerates tests/golden/image_golden.json: the files the reference's
ImageSaver (stb_image_write, flipped on write as main sets it) writes for
the images of image_cases.py, summarised by size and checksum, written by
oracle/_ref/asset_harness (built by oracle/build_ref.sh).

Usage: python3 tests/golden/make_image_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from image_cases import image_cases, to_u8  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "asset_harness")


def main():
    with tempfile.TemporaryDirectory() as tmp:
        args = []
        for name, img in image_cases().items():
            h, w = img.shape[:2]
            to_u8(img).tofile(os.path.join(tmp, name + ".rgba8"))
            img.tofile(os.path.join(tmp, name + ".f32"))
            args += [name, str(w), str(h)]
        out = os.path.join(tmp, "out.json")
        subprocess.run([HARNESS, out, "write", tmp] + args, check=True, stdout=subprocess.DEVNULL)
        res = json.load(open(out))
    with open(os.path.join(HERE, "image_golden.json"), "w") as fh:
        json.dump(res, fh, indent=1)
    print(f"{len(res)} images")


if __name__ == "__main__":
    main()
