"""Traversal with deep stacks.

A dense soup of small random triangles fills a cube, and three overlapping
instances of it sit in front of the camera, so the child boxes of almost
every node overlap: a ray pushes up to three siblings per level and its
stack grows past the part kept in LDS (OMV_STACK_LDS entries) into the
global spill area.  The stack's wave-uniform LDS-only path and its spill
path must give the same traversal: closest-hit ids bit-exact against the
oracle (false-colour pass), the wavefront film equal to the per-sample film,
no stack overflow.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _soup(wl, n=300_000, seed=11):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-1.0, 1.0, (n, 1, 3))
    tri = centres + rng.normal(0.0, 0.03, (n, 3, 3))
    d = wl.SceneDesc()
    d.materials.append(("om_nk", (0.4, 0.5, 0.6), (3.0, 3.5, 4.0)))
    d.meshes.append({"vertices": tri.reshape(-1, 3).astype(np.float32), "faces": np.arange(3 * n, dtype=np.uint32).reshape(n, 3),
                     "mat_id": 0})
    for t, a in (((0.0, 0.0, -4.0), 0.0), ((0.2, 0.1, -4.1), 0.6), ((-0.1, 0.2, -3.9), 1.3)):
        d.instances.append((wl.trs(t, angle=a), 0))
    d.lights.append(((2.0, 2.0, 1.0), (40.0, 40.0, 40.0), -1.0))
    d.background = ("const", (0.2, 0.3, 0.4))
    return d


def test_deep_stack_ids_match_the_oracle(pkg, wl, ref):
    d = _soup(wl)
    d.fc_subject = "pid"
    view = wl.View(48, 40, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), yfov=40.0)
    s = pkg.Scene()
    d.apply(s, pkg.pbr_desc)
    info = s.bvh_info()
    print(f"deep stack soup: bvh {info}", flush=True)
    # the builder's stack bound 3 * (TLAS + BLAS depth) + 4 exceeds the LDS part
    assert 3 * info["depth"] + 4 > 28
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(seed=5))
    r.set_scene(s)
    film, ids, tie = r.render_fc(cam)
    assert r.stats()["stack_overflows"] == 0
    rs = ref.Scene()
    d.apply(rs, ref.pbr_desc)
    rfilm, rids, rtie = rs.render_fc(ref.default_params(seed=5), cam.desc())
    assert np.array_equal(ids, rids)
    assert np.array_equal(tie, rtie)
    assert np.array_equal(film, rfilm)


def test_deep_stack_schedules_agree(pkg, wl):
    d = _soup(wl)
    view = wl.View(40, 32, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), yfov=40.0)
    s = pkg.Scene()
    d.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    params = dict(samples_per_pixel=2, camera_bounces=3, light_bounces=3, max_path_length=3, seed=7)
    out = []
    for sched in (pkg.SCHEDULE_WAVEFRONT, pkg.SCHEDULE_PER_SAMPLE):
        r = pkg.Renderer(pkg.default_params(**params))
        r.set_scene(s)
        r.set_schedule(sched)
        out.append((r.render(cam), r.stats()))
    (a, sa), (b, sb) = out
    assert np.array_equal(a, b), f"films differ at {np.argwhere(a != b)[:5].tolist()}"
    for k in ("rays", "camera_rays", "light_rays", "shadow_rays"):
        assert sa[k] == sb[k], (k, sa[k], sb[k])
    assert sa["stack_overflows"] == 0 and sb["stack_overflows"] == 0
