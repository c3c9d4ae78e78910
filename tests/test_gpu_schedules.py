"""The two GPU schedules of omv_renderer_render give the same film.

The wavefront schedule (ray queues, one traversal kernel per bounce) and the
per-sample schedule (one lane per sample) evaluate the same per-sample
arithmetic in the same order, so films must be equal bit for bit, and the
ray counts (camera, light, connection, total) must be equal too.  Cases cover
lights on/off, HDRI, masked materials (re-trace loop), the missing material,
the OM rough-conductor extension, instancing, and parameter edges
(B = 0, b = 0, q = 1, spp = 1).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(pkg, desc, view, **params):
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    out = []
    for sched in (pkg.SCHEDULE_WAVEFRONT, pkg.SCHEDULE_PER_SAMPLE):
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(s)
        r.set_schedule(sched)
        assert r.schedule() == sched
        img = r.render(cam)
        out.append((img, r.stats()))
    return out


def _same(out, visit_tol=0.1):
    (a, sa), (b, sb) = out
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, b), f"films differ at {np.argwhere(a != b)[:5].tolist()}"
    for k in ("samples", "rays", "camera_rays", "light_rays", "shadow_rays"):
        assert sa[k] == sb[k], (k, sa[k], sb[k])
    # node / triangle counts: both schedules traverse each ray entry by entry
    # with the same early-outs; checked loosely, the hits and rays exactly
    for k in ("nodes_visited", "tris_tested"):
        assert abs(sa[k] - sb[k]) <= visit_tol * sb[k], (k, sa[k], sb[k])
    assert sa["stack_overflows"] == 0


CASES = {
    "mixed_lights_hdr_masked": (dict(hdr=True, lights=True, masked=True), dict(samples_per_pixel=4, camera_bounces=5, light_bounces=4,
                                                                              max_path_length=6, seed=3)),
    "mixed_lights_only": (dict(hdr=False, lights=True), dict(samples_per_pixel=3, camera_bounces=4, light_bounces=4, max_path_length=5, seed=9)),
    "mixed_hdr_only_missing": (dict(hdr=True, lights=False, missing=True), dict(samples_per_pixel=3, camera_bounces=6, light_bounces=2,
                                                                               max_path_length=6, seed=1)),
    "mixed_om_rough": (dict(hdr=True, lights=True), dict(samples_per_pixel=2, camera_bounces=4, light_bounces=3, max_path_length=5, seed=2,
                                                         extensions=1, om_roughness=0.25)),
    "edge_B0": (dict(hdr=True, lights=True), dict(samples_per_pixel=2, camera_bounces=0, light_bounces=3, max_path_length=4, seed=4)),
    "edge_b0": (dict(hdr=True, lights=True), dict(samples_per_pixel=2, camera_bounces=3, light_bounces=0, max_path_length=4, seed=5)),
    "edge_q1": (dict(hdr=True, lights=True), dict(samples_per_pixel=2, camera_bounces=3, light_bounces=3, max_path_length=1, seed=6)),
    "edge_spp1_deep": (dict(hdr=True, lights=True, masked=True), dict(samples_per_pixel=1, camera_bounces=10, light_bounces=10,
                                                                      max_path_length=8, seed=7)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_schedules_agree_mixed(pkg, wl, name):
    kw, params = CASES[name]
    desc, view = wl.mixed_test_scene(**kw)
    view.width, view.height = 64, 48
    _same(_both(pkg, desc, view, **params))


@pytest.mark.parametrize("place", ["above", "below", "one_axis", "both"])
def test_light_placement_keeps_the_film(pkg, wl, place):
    """Light rays draw directions in the positive octant only, so a light
    above every child of the TLAS root on some axis has all its rays culled
    before their directions are drawn (the wavefront schedule), while the
    per-sample schedule traces every one.  Lights above the scene, below it
    (rays traced, light subpaths grow), above it on x only, and one of each:
    films and ray counts must agree bit for bit."""
    lights = {"above": [((3.0, 4.0, 3.5), (40.0, 40.0, 40.0), -1.0)],
              "below": [((-1.2, -0.6, -1.0), (30.0, 20.0, 10.0), -1.0)],
              "one_axis": [((6.0, -0.5, -0.5), (30.0, 30.0, 30.0), 20.0)],
              "both": [((3.0, 4.0, 3.5), (40.0, 40.0, 40.0), -1.0), ((-1.2, -0.6, -1.0), (30.0, 20.0, 10.0), -1.0)]}[place]
    desc, view = wl.mixed_test_scene(hdr=True, lights=False, masked=True)
    desc.lights.extend(lights)
    view.width, view.height = 64, 48
    out = _both(pkg, desc, view, samples_per_pixel=3, camera_bounces=4, light_bounces=4, max_path_length=6, seed=13)
    _same(out)
    if place in ("below", "both"):
        assert out[0][1]["light_rays"] > view.width * view.height * 3   # light subpaths past their first ray


def test_schedules_agree_torus_knot(pkg, wl):
    desc, view, params = wl.hdri_torus_knot(width=96, height=96, spp=4, tex=128, hdr_w=256, hdr_h=128, segments=200, tube=16)
    _same(_both(pkg, desc, view, **params))


def test_schedules_agree_icosphere_rough(pkg, wl):
    desc, view, params = wl.plain_bdpt_icosphere()
    view.width, view.height = 64, 64
    _same(_both(pkg, desc, view, **dict(params, samples_per_pixel=4, extensions=1, om_roughness=0.3)))


def test_schedules_agree_instanced(pkg, wl):
    desc, view, params = wl.large_instanced(width=120, height=68, spp=2, count=24, tris_per=2000, hdr_w=256, hdr_h=128)
    _same(_both(pkg, desc, view, **params))


@pytest.mark.parametrize("streams", [1, 2])
@pytest.mark.parametrize("lights", [False, True])
def test_wavefront_multi_chunk(pkg, wl, lights, streams):
    """Several wavefront chunks (chunk_samples not a multiple of spp): chunk
    boundaries fall on whole pixels and the film is unchanged, also with the
    chunks alternating between two streams."""
    desc, view = wl.mixed_test_scene(hdr=True, lights=lights)
    view.width, view.height = 200, 130
    params = dict(samples_per_pixel=5, camera_bounces=3, light_bounces=2, max_path_length=4, seed=8)
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    out = []
    for sched in (pkg.SCHEDULE_WAVEFRONT, pkg.SCHEDULE_PER_SAMPLE):
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(s)
        r.set_schedule(sched)
        if sched == pkg.SCHEDULE_WAVEFRONT:
            r.set_tuning(chunk_samples=20001, chunk_streams=streams)
        out.append((r.render(cam), r.stats()))
    _same(out)


@pytest.mark.parametrize("tuning", [dict(trace_keep=1), dict(trace_keep=32), dict(trace_keep=64),
                                    dict(trace_keep=48, trace_blocks_per_cu=1, chunk_samples=1000),
                                    dict(traversal=1), dict(traversal=1, trace_blocks_per_cu=3),
                                    dict(traversal=2), dict(traversal=2, tail_round=2, chunk_samples=999),
                                    dict(sort_rays=2), dict(sort_rays=2, chunk_samples=1000, traversal=1),
                                    dict(chunk_streams=2, chunk_samples=1000), dict(chunk_streams=2, chunk_samples=777, chunk_stagger=1, trace_keep=32),
                                    dict(chunk_streams=2, chunk_samples=1000, chunk_stagger=9, sort_rays=2)])
def test_wavefront_tuning_agrees(pkg, wl, tuning):
    """Launch tuning (persistent traversal with refill, fewer resident blocks,
    small chunks) changes the schedule only: the film and the counts equal the
    per-sample kernel's."""
    desc, view = wl.mixed_test_scene(hdr=True, lights=True, masked=True)
    view.width, view.height = 64, 48
    params = dict(samples_per_pixel=3, camera_bounces=5, light_bounces=4, max_path_length=6, seed=12)
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    out = []
    for sched in (pkg.SCHEDULE_WAVEFRONT, pkg.SCHEDULE_PER_SAMPLE):
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(s)
        r.set_schedule(sched)
        if sched == pkg.SCHEDULE_WAVEFRONT:
            r.set_tuning(**tuning)
        out.append((r.render(cam), r.stats()))
    _same(out)
    assert out[0][1]["trace_launches"] > 0 and out[0][1]["trace_ms"] > 0


def test_wavefront_tuning_knot(pkg, wl):
    desc, view, params = wl.hdri_torus_knot(width=96, height=96, spp=4, tex=128, hdr_w=256, hdr_h=128, segments=200, tube=16)
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    imgs = []
    for keep in (0, 40):
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(s)
        r.set_tuning(trace_keep=keep)
        imgs.append(r.render(cam))
    assert np.array_equal(imgs[0], imgs[1])


@pytest.mark.parametrize("sched", ["wavefront", "per_sample"])
def test_render_device_equals_render(pkg, wl, sched):
    """omv_renderer_render_device leaves the film in a device buffer; it is
    the film omv_renderer_render reads back, bit for bit, with both
    schedules and with several chunks on two streams."""
    torch = pytest.importorskip("torch")
    desc, view = wl.mixed_test_scene(hdr=True, lights=True)
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(53, 37, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(samples_per_pixel=4, camera_bounces=4, light_bounces=4, max_path_length=4, seed=9))
    r.set_scene(s)
    r.set_schedule(pkg.SCHEDULE_WAVEFRONT if sched == "wavefront" else pkg.SCHEDULE_PER_SAMPLE)
    if sched == "wavefront":
        r.set_tuning(chunk_samples=2000)   # several chunks over two streams
    want = r.render(cam)
    film = torch.full((cam.height, cam.width, 4), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    r.render_device(cam, film.data_ptr())
    got = film.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the camera keeps the film of the last host render
    assert np.array_equal(cam.film(), want)


@pytest.mark.parametrize("case", ["lights", "nee", "splat", "instanced"])
def test_occlusion_early_out_keeps_the_film(pkg, wl, case):
    """Visibility rays ending at the first deciding hit (tuning occlusion 0,
    the default) give the film of the closest-hit visibility (occlusion 1)
    bit for bit, with both schedules, and visit fewer nodes."""
    if case == "instanced":
        desc, view, _ = wl.large_instanced(width=64, height=36, spp=2, depth=4, count=24, tris_per=600, hdr_w=128, hdr_h=64)
        params = dict(samples_per_pixel=2, camera_bounces=4, light_bounces=4, max_path_length=4, seed=5)
    else:
        desc, view = wl.mixed_test_scene(hdr=True, lights=case != "nee", masked=False)
        params = dict(samples_per_pixel=3, camera_bounces=4, light_bounces=4, max_path_length=4, seed=5,
                      extensions={"lights": 0, "nee": 2, "splat": 4}[case])
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(min(view.width, 64), min(view.height, 48), view.position, view.forward, view.up, view.yfov)
    films = {}
    for sched in (pkg.SCHEDULE_WAVEFRONT, pkg.SCHEDULE_PER_SAMPLE):
        for occ in (0, 1):
            r = pkg.Renderer(pkg.default_params(**params))
            r.set_scene(s)
            r.set_schedule(sched)
            r.set_tuning(occlusion=occ)
            films[(sched, occ)] = (r.render(cam), r.stats())
    for sched in (pkg.SCHEDULE_WAVEFRONT, pkg.SCHEDULE_PER_SAMPLE):
        (a, sa), (b, sb) = films[(sched, 0)], films[(sched, 1)]
        if case == "splat":   # splat sums are float atomics: order-dependent
            np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)
        else:
            assert np.array_equal(a, b), (sched, np.argwhere(a != b)[:5].tolist())
        for k in ("rays", "shadow_rays"):
            assert sa[k] == sb[k], (k, sa[k], sb[k])
        assert sa["nodes_visited"] <= sb["nodes_visited"]


@pytest.mark.parametrize("size", ["small", "large"])
def test_xcd_banded_handout_keeps_the_film(pkg, wl, size):
    """The traversal kernels' XCD-banded hand-out of queue entries (tuning
    xcd_bands: 0 = default, four bands per XCD on scenes of at most 4
    instances, n = n bands, 1 = plain grid stride) visits every entry exactly
    once: films bit for bit and ray counts are those of the plain stride, for
    band counts that divide the grid (2, 16, 32 -- the default count on a
    few-instance scene -- and the default) and one that does not (3: plain
    stride).  The large
    case (512 x 512 x 8 = 2.1M samples) reaches the largest chunk size."""
    desc, view = wl.mixed_test_scene(hdr=True, lights=True, masked=False)
    spp, w, h = (3, 96, 64) if size == "small" else (8, 512, 512)
    params = dict(samples_per_pixel=spp, camera_bounces=4, light_bounces=4, max_path_length=4, seed=9)
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(w, h, view.position, view.forward, view.up, view.yfov)
    out = {}
    for bands in (1, 0, 2, 3, 16, 32):
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(s)
        r.set_tuning(xcd_bands=bands)
        out[bands] = (r.render(cam), r.stats())
    want, ws = out[1]
    for bands, (got, gs) in out.items():
        assert np.array_equal(got, want), (bands, np.argwhere(got != want)[:5].tolist())
        for k in ("rays", "camera_rays", "light_rays", "shadow_rays", "samples"):
            assert gs[k] == ws[k], (bands, k, gs[k], ws[k])


@pytest.mark.parametrize("case", ["lights", "nee", "hdri", "masked", "splat", "instanced", "chunks"])
def test_tail_launch_keeps_the_film(pkg, wl, case):
    """The late camera rounds run as one launch (tuning tail_round = n: each
    path alive at round n carried to its end by one lane) give the film of
    the round-by-round wavefront (tail_round = 1) bit for bit, with the same
    ray counts, from every starting round; also with several chunks on two
    streams (the chunk hand-over event is recorded in the tail case too)."""
    depth = 5
    params = dict(samples_per_pixel=3, camera_bounces=depth, light_bounces=depth, max_path_length=depth, seed=11)
    if case == "instanced":
        desc, view, _ = wl.large_instanced(width=64, height=36, spp=2, depth=depth, count=24, tris_per=600, hdr_w=128, hdr_h=64)
    elif case == "hdri":
        desc, view = wl.mixed_test_scene(hdr=True, lights=False, masked=False)
    else:
        desc, view = wl.mixed_test_scene(hdr=True, lights=case != "nee", masked=case == "masked")
        params["extensions"] = {"nee": 2, "splat": 4}.get(case, 0)
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(min(view.width, 64), min(view.height, 48), view.position, view.forward, view.up, view.yfov)
    out = {}
    for tail in (1, 2, 3, depth):
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(s)
        kw = dict(tail_round=tail)
        if case == "chunks":
            kw.update(chunk_samples=cam.width * 7 * params["samples_per_pixel"])
        r.set_tuning(**kw)
        out[tail] = (r.render(cam), r.stats())
    want, ws = out[1]
    for tail, (got, gs) in out.items():
        if case == "splat":   # splat sums are float atomics: order-dependent
            np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)
        else:
            assert np.array_equal(got, want), (tail, np.argwhere(got != want)[:5].tolist())
        for k in ("rays", "camera_rays", "light_rays", "shadow_rays", "samples"):
            assert gs[k] == ws[k], (tail, k, gs[k], ws[k])


@pytest.mark.parametrize("case", ["nee", "splat", "instanced"])
def test_refilling_traversal_keeps_the_film(pkg, wl, case):
    """The refilling traversal (trace_keep > 0: lanes take the next ray of
    their own list once few lanes of the wave are busy), also in the fused
    launch with environment and connection rays, gives the plain
    traversal's film and ray counts."""
    if case == "instanced":
        desc, view, _ = wl.large_instanced(width=64, height=36, spp=2, depth=6, count=24, tris_per=600, hdr_w=128, hdr_h=64)
        params = dict(samples_per_pixel=2, camera_bounces=6, light_bounces=6, max_path_length=6, seed=13)
    else:
        desc, view = wl.mixed_test_scene(hdr=True, lights=True, masked=True)
        params = dict(samples_per_pixel=3, camera_bounces=5, light_bounces=5, max_path_length=5, seed=13,
                      extensions={"nee": 2, "splat": 4}[case])
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(min(view.width, 64), min(view.height, 48), view.position, view.forward, view.up, view.yfov)
    out = {}
    for keep in (0, 16, 48, 64):
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(s)
        r.set_tuning(trace_keep=keep)
        out[keep] = (r.render(cam), r.stats())
    want, ws = out[0]
    for keep, (got, gs) in out.items():
        if case == "splat":   # splat sums are float atomics: order-dependent
            np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)
        else:
            assert np.array_equal(got, want), (keep, np.argwhere(got != want)[:5].tolist())
        for k in ("rays", "camera_rays", "light_rays", "shadow_rays", "samples"):
            assert gs[k] == ws[k], (keep, k, gs[k], ws[k])


@pytest.mark.parametrize("case", ["masked", "lights", "nee", "instanced", "knot", "chunks"])
def test_packet_traversal_keeps_the_film(pkg, wl, case):
    """Round-1 camera rays traced as wave packets (tuning traversal = 2: one
    wave-uniform stack of (entry, lane mask), records read once per wave)
    give the per-ray traversal's film and ray counts bit for bit: with the
    alpha-mask re-trace loop, instancing, a single-mesh scene, NEE, and
    chunks whose sample counts leave the last wave part-filled."""
    depth = 4
    params = dict(samples_per_pixel=3, camera_bounces=depth, light_bounces=depth, max_path_length=depth, seed=17)
    if case == "instanced":
        desc, view, _ = wl.large_instanced(width=64, height=36, spp=2, depth=depth, count=24, tris_per=600, hdr_w=128, hdr_h=64)
    elif case == "knot":
        desc, view, params = wl.hdri_torus_knot(width=40, height=33, spp=16, tex=128, hdr_w=256, hdr_h=128, segments=200, tube=16)
    else:
        desc, view = wl.mixed_test_scene(hdr=True, lights=case != "nee", masked=case in ("masked", "chunks"))
        params["extensions"] = 2 if case == "nee" else 0
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(min(view.width, 61), min(view.height, 47), view.position, view.forward, view.up, view.yfov)
    out = {}
    for trav in (1, 2):
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(s)
        kw = dict(traversal=trav)
        if case == "chunks":
            kw.update(chunk_samples=cam.width * 5 * params["samples_per_pixel"] + 7)
        r.set_tuning(**kw)
        out[trav] = (r.render(cam), r.stats())
    # a lane visits the leaves in the packet's order (its leader lane's
    # nearest first), not its own, so its closest t culls a little less:
    # node and triangle counts are checked to 25 %
    _same([out[2], out[1]], visit_tol=0.25)
