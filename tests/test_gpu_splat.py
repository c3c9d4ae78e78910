"""OMV_EXT_LIGHT_SPLAT on the GPU.

Bars:
  * the GPU splat image (film with the extension minus film without it)
    against the oracle's on the same seeds: >= 98 % of the pixels carrying
    splats within 1e-3 relative, image sums within 1e-3 (splats are summed
    with float atomics in arbitrary order, and rare paths branch differently
    between device and host transcendentals);
  * the expectations the oracle is pinned by, at 16x the oracle's samples:
    direct splats vs the reference estimator's direct image of the same
    scene, mirror splats vs the mirror-image light's direct image;
  * the two schedules and a chunked render agree to 1e-5 relative (the
    camera estimate is bit-identical; splat sums differ in atomic order),
    with identical ray counts;
  * scenes without point lights are unaffected; batches refuse splat jobs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(pkg, view):
    return pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)


def _gpu(pkg, desc, view, params, ext, spp, seed=1, schedule=None, tuning=None):
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    r = pkg.Renderer(pkg.default_params(samples_per_pixel=spp, seed=seed, extensions=ext, **params))
    r.set_scene(s)
    if schedule is not None:
        r.set_schedule(schedule)
    if tuning:
        r.set_tuning(**tuning)
    img = r.render(_cam(pkg, view))
    return img[..., :3].astype(np.float64), r.stats()


def _ref(ref, desc, view, params, ext, spp, seed=1):
    s = ref.Scene()
    desc.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    img, st = s.render(ref.default_params(samples_per_pixel=spp, seed=seed, extensions=ext, **params), cam, threads=8)
    return img[..., :3].astype(np.float64), st


def _blocks(img, n=4):
    h, w = img.shape[:2]
    return img.reshape(n, h // n, n, w // n, 3).sum(axis=(1, 3, 4))


@pytest.mark.parametrize("name", ["direct", "mirror"])
def test_splat_image_matches_oracle(pkg, wl, ref, name):
    d, v, p = wl.light_splat_scenes()[name]
    g = _gpu(pkg, d, v, p, 4, 256)[0] - _gpu(pkg, d, v, p, 0, 256)[0]
    c = _ref(ref, d, v, p, 4, 256)[0] - _ref(ref, d, v, p, 0, 256)[0]
    assert abs(g.sum() / c.sum() - 1.0) < 1e-3, (g.sum(), c.sum())
    sig = c > 1e-3 * c.max()
    rel = np.abs(g - c)[sig] / c[sig]
    assert np.mean(rel < 1e-3) >= 0.98, np.quantile(rel, [0.5, 0.9, 0.99])


@pytest.mark.parametrize("name,tot,blk", [("direct", 0.005, 0.03), ("mirror", 0.01, 0.05)])
def test_splat_expectation_high_spp(pkg, wl, name, tot, blk):
    sc = wl.light_splat_scenes()
    d, v, p = sc[name]
    spp = 16384
    splat = _gpu(pkg, d, v, p, 4, spp)[0] - _gpu(pkg, d, v, p, 0, spp)[0]
    if name == "direct":
        want = _gpu(pkg, d, v, p, 0, 1024)[0]
    else:
        r, rv, rp = sc["mirror_ref"]
        want = _gpu(pkg, r, rv, rp, 0, 1024)[0]
    assert abs(splat.sum() / want.sum() - 1.0) < tot, splat.sum() / want.sum()
    ratio = _blocks(splat) / _blocks(want)
    assert np.all(np.abs(ratio - 1.0) < blk), ratio


PARAMS = dict(camera_bounces=4, light_bounces=4, max_path_length=5)


@pytest.mark.parametrize("name", ["mirror", "mixed"])
def test_splat_schedules_and_chunks_agree(pkg, wl, name):
    # the mirror scene: deep light walks (mirror <-> floor) that splat at
    # every vertex; the mixed scene: every material, masked re-traces
    if name == "mirror":
        desc, view, _ = wl.light_splat_scenes()["mirror"]
    else:
        desc, view = wl.mixed_test_scene(hdr=True, lights=True, masked=True)
        desc.lights.append(((-2.5, -0.9, -2.0), (30.0, 30.0, 30.0), -1.0))   # its + octant faces the objects
    a, sa = _gpu(pkg, desc, view, PARAMS, 4, 8, schedule=pkg.SCHEDULE_WAVEFRONT)
    b, sb = _gpu(pkg, desc, view, PARAMS, 4, 8, schedule=pkg.SCHEDULE_PER_SAMPLE)
    c, sc_ = _gpu(pkg, desc, view, PARAMS, 4, 8, tuning=dict(chunk_samples=5001, chunk_streams=2))
    plain, _ = _gpu(pkg, desc, view, PARAMS, 0, 8)
    assert np.all(np.isfinite(a))
    assert not np.allclose(a, plain)   # the splats land on the image
    for x in (b, c):
        np.testing.assert_allclose(x, a, rtol=1e-5, atol=1e-6 * np.abs(a).max())
    for k in ("rays", "camera_rays", "light_rays", "shadow_rays"):
        assert sa[k] == sb[k] == sc_[k], (k, sa[k], sb[k], sc_[k])


def test_splat_off_without_lights(pkg, wl):
    desc, view = wl.mixed_test_scene(hdr=True, lights=False)
    a, _ = _gpu(pkg, desc, view, PARAMS, 4, 2)
    b, _ = _gpu(pkg, desc, view, PARAMS, 0, 2)
    assert np.array_equal(a, b)


def test_batch_refuses_splat_jobs(pkg, wl):
    desc, view = wl.mixed_test_scene(hdr=True, lights=True)
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    b = pkg.Batch()
    with pytest.raises(RuntimeError, match="SPLAT"):
        b.add_job(s, _cam(pkg, view), pkg.default_params(samples_per_pixel=1, extensions=4))
