"""OMV_EXT_LIGHT_SPLAT in the oracle: light-tracer (t = 1) splats.

The splat image (film with the extension minus film without it, same seed:
the camera estimate is identical) is pinned against expectations it must
reproduce, not against a restatement of its own formula:

  * "direct": a point light under a diffuse PBR ceiling.  The splats of the
    first light vertex estimate the reference estimator's own image of the
    scene with B = 1, q = 2 (only its direct connections are non-zero) --
    this fixes the particle weight I / p(w) of the + octant emission, the
    pinhole Jacobian f^2 cos_y / (d^2 cos^3 theta), the pixel mapping and the
    1 / (W H spp) normalisation;
  * "mirror": the light reaches a diffuse floor only through a near-perfect
    OpenMATERIAL mirror.  The second-vertex splats estimate the direct image
    of the mirror-image light -- this fixes the particle-weight update
    beta * w * |cos_out| along the walk;
  * two planted mistakes in the oracle (a cos(theta) dropped from the
    Jacobian, the 1/p of the emission octant dropped) each fail the first
    comparison.
"""
import numpy as np
import pytest

SPP = 1024


def _render(ref, desc, view, params, ext, spp, seed=1):
    s = ref.Scene()
    desc.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    p = ref.default_params(samples_per_pixel=spp, seed=seed, extensions=ext, **params)
    img, st = s.render(p, cam, threads=8)
    return img[..., :3].astype(np.float64), st


def _blocks(img, n=4):
    h, w = img.shape[:2]
    return img.reshape(n, h // n, n, w // n, 3).sum(axis=(1, 3, 4))


def _splat_and_want(wl, ref, name):
    sc = wl.light_splat_scenes()
    d, v, p = sc[name]
    with_splat, st = _render(ref, d, v, p, 4, SPP)
    plain, st0 = _render(ref, d, v, p, 0, SPP)
    if name == "direct":
        want = plain
    else:
        r, rv, rp = sc["mirror_ref"]
        want, _ = _render(ref, r, rv, rp, 0, 64)
    # the splat visibility rays are counted with the shadow rays
    assert st["shadow_rays"] > st0["shadow_rays"]
    assert st["camera_rays"] == st0["camera_rays"] and st["light_rays"] == st0["light_rays"]
    return with_splat - plain, want


@pytest.mark.parametrize("name,tot,blk", [("direct", 0.01, 0.05), ("mirror", 0.02, 0.08)])
def test_splat_matches_expectation(wl, ref, name, tot, blk):
    splat, want = _splat_and_want(wl, ref, name)
    assert np.all(np.isfinite(splat)) and splat.min() >= 0.0
    assert abs(splat.sum() / want.sum() - 1.0) < tot, splat.sum() / want.sum()
    ratio = _blocks(splat) / _blocks(want)
    assert np.all(np.abs(ratio - 1.0) < blk), ratio


@pytest.mark.parametrize("mask", [1, 2])
def test_planted_mistakes_fail(wl, ref, mask):
    ref.lib().ref_splat_plant(mask)
    try:
        splat, want = _splat_and_want(wl, ref, "direct")
    finally:
        ref.lib().ref_splat_plant(0)
    ratio = _blocks(splat) / _blocks(want)
    assert abs(splat.sum() / want.sum() - 1.0) > 0.05 or np.any(np.abs(ratio - 1.0) > 0.05), ratio


def test_splat_off_without_lights(wl, ref):
    desc, view = wl.mixed_test_scene(hdr=True, lights=False)
    a, _ = _render(ref, desc, view, dict(camera_bounces=3, light_bounces=2, max_path_length=4), 4, 2)
    b, _ = _render(ref, desc, view, dict(camera_bounces=3, light_bounces=2, max_path_length=4), 0, 2)
    assert np.array_equal(a, b)
