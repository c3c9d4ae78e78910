"""Float RGBA test images for the image writers (seeded; shared by
make_image_golden.py and tests/test_image_io.py)."""
import numpy as np


def image_cases():
    rng = np.random.default_rng(90210)
    out = {}
    out["tiny"] = rng.random((1, 1, 4), dtype=np.float32)
    noise = rng.random((23, 37, 4), dtype=np.float32)
    out["noise_alpha"] = noise
    y, x = np.mgrid[0:48, 0:64].astype(np.float32)
    g = np.stack([x / 63, y / 47, (x + y) / 110, np.ones_like(x)], -1).astype(np.float32)
    out["gradient"] = g
    flat = np.zeros((10, 50, 4), np.float32)
    flat[..., 3] = 1
    flat[:, :20, 0] = 0.5
    flat[:, 20:, 1] = 0.25
    flat[3, 7] = (0.9, 0.1, 0.3, 1.0)
    flat[6, 30:33] = rng.random((3, 4), dtype=np.float32)
    flat[8, 45:] = (1.0, 1.0, 1.0, 1.0)
    out["flat_runs"] = flat
    yy, xx = np.mgrid[0:200, 0:300].astype(np.float32)
    shade = 0.5 + 0.5 * np.sin(xx / 23.0) * np.cos(yy / 17.0)
    render = np.stack([shade, shade ** 2, 1 - shade, np.ones_like(shade)], -1)
    render[..., :3] = np.clip(render[..., :3] + rng.normal(0, 0.02, (200, 300, 3)), 0, 1)
    render[60:90, 100:250, :3] = 0.0
    out["render"] = render.astype(np.float32)
    wide = np.ones((3, 1000, 4), np.float32)
    wide[..., :3] = (np.arange(1000, dtype=np.float32)[None, :, None] % 97) / 96.0
    out["wide"] = wide
    for k, v in out.items():
        v[..., :3] = np.clip(v[..., :3], 0.0, 1.0)
        out[k] = np.ascontiguousarray(v, np.float32)
    return out


def to_u8(img):
    """static_cast<unsigned char>(255.0f * f) of [0, 1] floats."""
    return (np.float32(255.0) * img).astype(np.uint8)
