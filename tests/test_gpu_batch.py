"""Batched jobs (one persistent launch over many scenes/cameras/parameters).

Bar: every job's film is bit-identical to Renderer.render's for the same
scene, camera and parameters (same Philox streams, same per-pixel summation
order), whatever the tuning (one persistent per-pixel kernel, or the jobs'
wavefront sequences over several streams) and whether the launch is replayed
from a captured graph; one job is also checked against the CPU oracle at the
radiance tolerance of the plain path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_batch(pkg, wl, n_jobs=10):
    assets, jobs = wl.batched_lightbox(jobs=n_jobs, views_per_asset=2, tri_range=(500, 6000), res_range=(24, 56),
                                       spp_range=(1, 6), depth_range=(2, 6), hdr_w=128, hdr_h=64)
    scenes = []
    for d in assets:
        s = pkg.Scene()
        d.apply(s, pkg.pbr_desc)
        scenes.append(s)
    return assets, jobs, scenes


def _plain_films(pkg, jobs, scenes):
    films = []
    for a, view, params in jobs:
        cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(scenes[a])
        films.append(r.render(cam))
    return films


def _batch(pkg, jobs, scenes):
    b = pkg.Batch()
    cams = []
    for a, view, params in jobs:
        cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
        cams.append(cam)
        b.add_job(scenes[a], cam, pkg.default_params(**params))
    return b


@pytest.mark.parametrize("tuning", [dict(blocks_per_cu=0, tile_waves=1, use_graph=True),
                                    dict(blocks_per_cu=1, tile_waves=4, use_graph=False),
                                    dict(blocks_per_cu=2, tile_waves=2, use_graph=True),
                                    dict(streams=1, use_graph=False),
                                    dict(streams=3, use_graph=True),
                                    dict(streams=4, use_graph=True)])
def test_batch_equals_plain(pkg, wl, tuning):
    _, jobs, scenes = _small_batch(pkg, wl)
    plain = _plain_films(pkg, jobs, scenes)
    b = _batch(pkg, jobs, scenes)
    b.set_tuning(**tuning)
    for rep in range(3):   # graph replays keep giving the eager result
        b.run()
        for j in range(len(jobs)):
            assert np.array_equal(b.film(j), plain[j]), f"job {j} differs on run {rep}"
    st = b.stats()
    assert st["samples"] == sum(v.width * v.height * p["samples_per_pixel"] for _, v, p in jobs)
    assert st["stack_overflows"] == 0 and st["rays"] > 0


def test_batch_mixed_lights_and_zero_spp(pkg, wl):
    """Jobs with and without point lights, and an spp = 0 job (film rgb 0,
    alpha 1 as PinholeCamera::stop leaves it), in one launch."""
    desc, view = wl.mixed_test_scene(hdr=True, lights=True)
    desc2, _ = wl.mixed_test_scene(hdr=True, lights=False)
    s1, s2 = pkg.Scene(), pkg.Scene()
    desc.apply(s1, pkg.pbr_desc)
    desc2.apply(s2, pkg.pbr_desc)
    jobs = [(0, 40, 30, dict(samples_per_pixel=3, camera_bounces=4, light_bounces=3, max_path_length=5, seed=1)),
            (1, 33, 17, dict(samples_per_pixel=2, camera_bounces=3, light_bounces=3, max_path_length=4, seed=2)),
            (0, 16, 16, dict(samples_per_pixel=0, camera_bounces=3, light_bounces=3, max_path_length=4, seed=3))]
    scenes = [s1, s2]
    b = pkg.Batch()
    plain = []
    for si, w, h, p in jobs:
        cam = pkg.Camera(w, h, view.position, view.forward, view.up, view.yfov)
        r = pkg.Renderer(pkg.default_params(**p))
        r.set_scene(scenes[si])
        plain.append(r.render(cam))
        b.add_job(scenes[si], cam, pkg.default_params(**p))
    for streams in (0, 2):
        b.set_tuning(streams=streams)
        b.run()
        for j in range(len(jobs)):
            assert np.array_equal(b.film(j), plain[j]), (streams, j)
        z = b.film(2)
        assert np.all(z[..., :3] == 0) and np.all(z[..., 3] == 1)


def test_batch_job_against_oracle(pkg, wl, ref):
    assets, jobs, scenes = _small_batch(pkg, wl, n_jobs=4)
    b = _batch(pkg, jobs, scenes)
    b.run()
    a, view, params = jobs[1]
    rs = ref.Scene()
    assets[a].apply(rs, ref.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    rimg, _ = rs.render(ref.default_params(**params), cam.desc(), threads=8)
    img = b.film(1)
    err = np.abs(img[..., :3].astype(np.float64) - rimg[..., :3]) / np.maximum(np.abs(rimg[..., :3]), 1e-3)
    assert float(np.mean(err.max(-1) < 1e-3)) >= 0.99


@pytest.mark.skip(reason="one run of this test ended in a host-side segmentation fault inside omv_batch_autotune "
                         "(after the HDRI NEE / light-splat changes); no cause was found by reading the code, nor by the "
                         "CPU stand-in of the runtime under AddressSanitizer (tests/test_batch_hostsim.py).  That sweep "
                         "captured a graph per candidate, re-recording fork/join events across destroyed graphs; it was "
                         "replaced by the plain-launch sweep of test_batch_autotune_plain_sweep and stays off the GPU")
def test_batch_autotune(pkg, wl):
    _, jobs, scenes = _small_batch(pkg, wl, n_jobs=6)
    plain = _plain_films(pkg, jobs, scenes)
    b = _batch(pkg, jobs, scenes)
    best = b.autotune(reps=1)
    # either a multi-stream wavefront configuration or a per-pixel kernel one
    assert best["streams"] in (1, 2, 4, 8) or (best["blocks_per_cu"] >= 1 and best["tile_waves"] in (1, 2, 4)), best
    b.run()
    for j in range(len(jobs)):
        assert np.array_equal(b.film(j), plain[j])


def test_batch_autotune_plain_sweep(pkg, wl):
    """The autotuner times every candidate (wavefront sequences over 1, 2, 4
    streams -- no more than the 6 jobs -- and the per-pixel kernel's
    blocks/CU x tile sizes) with plain launches, keeps the fastest, and the
    batch then runs the choice captured into a graph (fresh fork / join
    events for the capture); every job's film equals the plain render's."""
    _, jobs, scenes = _small_batch(pkg, wl, n_jobs=6)
    plain = _plain_films(pkg, jobs, scenes)
    b = _batch(pkg, jobs, scenes)
    best = b.autotune(reps=1)
    print(f"autotune choice {best}", flush=True)
    assert best["use_graph"] == 1
    assert best["streams"] in (1, 2, 4) or (best["streams"] == 0 and best["blocks_per_cu"] >= 1 and best["tile_waves"] in (1, 2, 4)), best
    for rep in range(2):
        b.run()
        for j in range(len(jobs)):
            assert np.array_equal(b.film(j), plain[j]), f"job {j} differs on run {rep}"
