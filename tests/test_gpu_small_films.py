"""Tiny and ragged films (1 x 1, one row, one column, odd sizes) through the
wavefront and per-sample schedules, the false-colour pass and the post pass:
far fewer samples than one traversal grid holds, and post windows that fall
off every edge.  Bars as in the full-size parity tests: films of the two
schedules bit-identical, ids and tie flags bit-exact against the oracle,
radiance per pixel within 1e-3 relative, post floats within 2 ulp."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (1, 7), (13, 1), (3, 5), (2, 2)]
PARAMS = dict(samples_per_pixel=3, camera_bounces=4, light_bounces=3, max_path_length=5, seed=21)


def _scene(pkg, wl, w, h):
    desc, view = wl.mixed_test_scene(hdr=True, lights=True)
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    # a narrow field of view on the middle of the scene, so the few pixels see geometry
    cam = pkg.Camera(w, h, view.position, view.forward, view.up, 12.0)
    return desc, s, cam


@pytest.mark.parametrize("w,h", SIZES)
def test_small_film_radiance(pkg, wl, ref, w, h):
    print(f"small film {w}x{h}", flush=True)
    desc, s, cam = _scene(pkg, wl, w, h)
    films = []
    for sched in (pkg.SCHEDULE_WAVEFRONT, pkg.SCHEDULE_PER_SAMPLE):
        r = pkg.Renderer(pkg.default_params(**PARAMS))
        r.set_scene(s)
        r.set_schedule(sched)
        films.append(r.render(cam).copy())
        assert r.stats()["samples"] == w * h * PARAMS["samples_per_pixel"]
    assert films[0].shape == (h, w, 4)
    assert np.array_equal(films[0], films[1])
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    rimg = rs.render(ref.default_params(**PARAMS), cam.desc(), threads=4)[0]
    a, b = films[0][..., :3].astype(np.float64), rimg[..., :3].astype(np.float64)
    assert np.all(np.isfinite(a))
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    assert err.max() < 1e-3, err.max()
    for denoise in (False, True):
        f, _ = r.post(cam, gamma=0.5, denoise=denoise)
        rf, _ = ref.post(films[1], gamma=0.5, denoise=denoise)
        ulp = np.abs(f.view(np.int32).astype(np.int64) - rf.view(np.int32).astype(np.int64))
        assert ulp.max() <= 2, (denoise, ulp.max())


@pytest.mark.parametrize("w,h", SIZES)
def test_small_film_ids(pkg, wl, ref, w, h):
    print(f"small film ids {w}x{h}", flush=True)
    desc, _, _ = _scene(pkg, wl, w, h)
    desc.fc_subject = "pid"
    desc.background = None
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    _, _, cam = _scene(pkg, wl, w, h)
    r = pkg.Renderer(pkg.default_params(seed=7))
    r.set_scene(s)
    film, ids, tie = r.render_fc(cam)
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    rfilm, rids, rtie = rs.render_fc(ref.default_params(seed=7), cam.desc())
    assert np.array_equal(ids, rids) and np.array_equal(tie, rtie) and np.array_equal(film, rfilm)
