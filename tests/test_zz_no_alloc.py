"""Allocation-free render entry point (omv_renderer_render_noalloc).

Bar: its film, written into a caller-owned device buffer, is bit-identical
to omv_renderer_render's for the same scene, camera and parameters, and its
ray counts are equal; two calls in a row (the second on a smaller film)
both hold, since the fixed scratch keeps the last call's contents.  Host-side
refusals (light-splat extension, light subpaths longer than the scratch
holds) return an error before anything is launched.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _setup(pkg, wl, hdr=True, lights=True, **pk):
    desc, view = wl.mixed_test_scene(hdr=hdr, lights=lights)
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    params = dict(samples_per_pixel=4, camera_bounces=4, light_bounces=4, max_path_length=4, seed=11)
    params.update(pk)
    return s, view, params


def _cam(pkg, view, w, h):
    return pkg.Camera(w, h, view.position, view.forward, view.up, view.yfov)


def _plain(pkg, scene, params, cam):
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(scene)
    film = r.render(cam)
    return film, r.stats()


def _noalloc(pkg, r, cam, stats=True):
    film = torch.full((cam.height, cam.width, 4), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = r.render_noalloc(cam, film.data_ptr(), torch.cuda.current_stream().cuda_stream, stats=stats)
    torch.cuda.synchronize()
    return film.cpu().numpy(), st


RAY_KEYS = ("samples", "rays", "camera_rays", "light_rays", "shadow_rays", "stack_overflows")


@pytest.mark.parametrize("case", ["lights_hdr", "nee_no_lights", "lights_no_hdr"])
def test_noalloc_equals_render_twice(pkg, wl, case):
    if case == "lights_hdr":
        scene, view, params = _setup(pkg, wl, hdr=True, lights=True)
    elif case == "nee_no_lights":
        scene, view, params = _setup(pkg, wl, hdr=True, lights=False, extensions=2)
    else:
        scene, view, params = _setup(pkg, wl, hdr=False, lights=True, samples_per_pixel=3, max_path_length=5,
                                     camera_bounces=5, light_bounces=5)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(scene)
    # first call: 80 x 60; second call on the same GPU scratch: a smaller film
    for (w, h) in ((80, 60), (37, 21)):
        cam = _cam(pkg, view, w, h)
        print(f"noalloc case {case} film {w}x{h}", flush=True)
        want, wst = _plain(pkg, scene, params, cam)
        got, st = _noalloc(pkg, r, cam)
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
            f"{case} {w}x{h}: {int(np.sum(got != want))} values differ from omv_renderer_render"
        for k in RAY_KEYS:
            assert st[k] == wst[k], f"{case} {w}x{h}: {k} {st[k]} != {wst[k]}"


def test_noalloc_enqueue_only(pkg, wl):
    """stats=None: the call only enqueues; after a wait the film is the same."""
    scene, view, params = _setup(pkg, wl)
    cam = _cam(pkg, view, 33, 17)
    want, _ = _plain(pkg, scene, params, cam)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(scene)
    got, st = _noalloc(pkg, r, cam, stats=False)
    assert st is None
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_noalloc_refusals(pkg, wl):
    """Refused on the host, before any launch, with a clear message."""
    scene, view, params = _setup(pkg, wl)
    cam = _cam(pkg, view, 16, 16)
    film = torch.zeros((16, 16, 4), dtype=torch.float32, device="cuda")
    r = pkg.Renderer(pkg.default_params(**dict(params, extensions=4)))
    r.set_scene(scene)
    with pytest.raises(RuntimeError, match="LIGHT_SPLAT"):
        r.render_noalloc(cam, film.data_ptr())
    lim = 68   # OMV_NOALLOC_MAX_LIGHT_VERTICES
    r2 = pkg.Renderer(pkg.default_params(**dict(params, light_bounces=lim, max_path_length=lim + 1, camera_bounces=1)))
    r2.set_scene(scene)
    with pytest.raises(RuntimeError, match="more vertices than the fixed scratch"):
        r2.render_noalloc(cam, film.data_ptr())
    # the largest accepted light subpath still renders and equals the plain render
    p3 = dict(params, light_bounces=lim - 1, max_path_length=lim, camera_bounces=2, samples_per_pixel=1)
    r3 = pkg.Renderer(pkg.default_params(**p3))
    r3.set_scene(scene)
    want, _ = _plain(pkg, scene, p3, cam)
    got, _ = _noalloc(pkg, r3, cam)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
