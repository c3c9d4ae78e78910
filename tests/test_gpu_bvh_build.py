"""OMV_BUILD_DEVICE_LBVH: per-mesh BVHs built on the GPU (bvh_device.hip).

Bars: a scene committed with the device build renders the same film bit
for bit as with the host SAH build, with the same ray counts and the same
false-colour ids and tie flags (the closest hit does not depend on the
tree); the traversal-stack bound holds; edge meshes (1, 2, 3 triangles,
all centroids equal, a sliver mesh) build and render like the host build.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(pkg, view, w=None, h=None):
    return pkg.Camera(w or view.width, h or view.height, view.position, view.forward, view.up, view.yfov)


def _render(pkg, desc, cam, mode, params, fc=False):
    s = pkg.Scene()
    s.set_build_mode(mode)
    desc.apply(s, pkg.pbr_desc)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(s)
    if fc:
        out = r.render_fc(cam)
    else:
        out = r.render(cam)
    return out, r.stats(), s.bvh_info()


SCENES = ["mixed", "knot", "instanced"]


def _scene(wl, name):
    if name == "mixed":
        desc, view = wl.mixed_test_scene(hdr=True, lights=True, masked=True)
        return desc, view
    if name == "knot":
        desc, view, _ = wl.hdri_torus_knot(width=96, height=96, tex=64, hdr_w=64, hdr_h=32, segments=300, tube=20)
        return desc, view
    desc, view, _ = wl.large_instanced(width=128, height=72, count=24, tris_per=3000, hdr_w=64, hdr_h=32)
    return desc, view


@pytest.mark.parametrize("name", SCENES)
def test_device_build_renders_same_film(pkg, wl, name):
    desc, view = _scene(wl, name)
    print(f"bvh build case {name}", flush=True)
    cam = _cam(pkg, view, 96, 72)
    params = dict(samples_per_pixel=2, camera_bounces=4, light_bounces=3, max_path_length=5, seed=9)
    a, sa, ia = _render(pkg, desc, cam, pkg.BUILD_HOST_SAH, params)
    b, sb, ib = _render(pkg, desc, cam, pkg.BUILD_DEVICE_LBVH, params)
    assert np.array_equal(a, b), np.argwhere(a != b)[:5].tolist()
    for k in ("rays", "camera_rays", "light_rays", "shadow_rays"):
        assert sa[k] == sb[k], k
    assert sb["stack_overflows"] == 0
    assert ib["tri_records"] == ia["tri_records"] and ib["nodes"] > 0


@pytest.mark.parametrize("name", SCENES)
def test_device_build_fc_ids_same(pkg, wl, name):
    desc, view = _scene(wl, name)
    desc.fc_subject = "pid"
    cam = _cam(pkg, view, 96, 72)
    (fa, ida, ta), _, _ = _render(pkg, desc, cam, pkg.BUILD_HOST_SAH, dict(seed=4), fc=True)
    (fb, idb, tb), _, _ = _render(pkg, desc, cam, pkg.BUILD_DEVICE_LBVH, dict(seed=4), fc=True)
    assert np.array_equal(ida, idb) and np.array_equal(ta, tb) and np.array_equal(fa, fb)
    assert (ida[0] >= 0).mean() > 0.05


def _edge_desc(wl, kind):
    d = wl.SceneDesc()
    d.materials.append(("pbr", dict(base_color=(0.8, 0.6, 0.4, 1.0), metallic=0.0, roughness=0.5)))
    rng = np.random.default_rng(7)
    if kind in ("one", "two", "three"):
        n = {"one": 1, "two": 2, "three": 3}[kind]
        v = rng.uniform(-1, 1, (3 * n, 3)).astype(np.float32)
        f = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    elif kind == "same_centroid":   # every Morton code equal
        n = 50
        ang = rng.uniform(0, 2 * np.pi, n)
        v = np.zeros((3 * n, 3), np.float32)
        for i, a in enumerate(ang):
            c, s = np.cos(a), np.sin(a)
            v[3 * i:3 * i + 3] = [[c, s, 0.3], [-c, -s, 0.3], [0.5 * s, -0.5 * c, -0.6]]
        v -= v.reshape(n, 3, 3).mean(axis=1).repeat(3, axis=0)
        f = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    else:                            # a sliver: a long thin strip of 4000 triangles
        x = np.linspace(-3, 3, 2001, dtype=np.float32)
        v = np.concatenate([np.stack([x, np.zeros_like(x), np.zeros_like(x)], 1),
                            np.stack([x, np.full_like(x, 1e-3), np.zeros_like(x)], 1)]).astype(np.float32)
        i = np.arange(2000)
        f = np.concatenate([np.stack([i, i + 1, 2001 + i], 1), np.stack([i + 1, 2002 + i, 2001 + i], 1)]).astype(np.uint32)
    d.meshes.append({"vertices": v, "faces": f, "mat_id": 0})
    d.instances.append((wl.trs((0.0, 0.0, -4.0)), 0))
    d.instances.append((wl.trs((0.8, 0.3, -5.0), angle=0.7), 0))
    d.lights.append(((2.0, 2.0, 1.0), (40.0, 40.0, 40.0), -1.0))
    d.background = ("const", (0.2, 0.3, 0.4))
    return d


@pytest.mark.parametrize("kind", ["one", "two", "three", "same_centroid", "sliver"])
def test_device_build_edge_meshes(pkg, wl, kind):
    desc = _edge_desc(wl, kind)
    print(f"bvh edge case {kind}", flush=True)
    view = wl.View(64, 48, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0))
    cam = _cam(pkg, view)
    params = dict(samples_per_pixel=2, camera_bounces=3, light_bounces=2, max_path_length=4, seed=3)
    a, sa, _ = _render(pkg, desc, cam, pkg.BUILD_HOST_SAH, params)
    b, sb, _ = _render(pkg, desc, cam, pkg.BUILD_DEVICE_LBVH, params)
    assert np.array_equal(a, b)
    assert sa["rays"] == sb["rays"]


def test_device_build_commit_time(pkg, wl):
    """Commit of one 1.2M-triangle mesh: host SAH vs device build (printed;
    the device build is required to be faster)."""
    v, f, _ = wl.torus_knot(segments=4000, tube=150)
    d = wl.SceneDesc()
    d.materials.append(("om_nk", (0.2, 0.4, 1.1), (3.1, 2.4, 1.9)))
    d.meshes.append({"vertices": v, "faces": f, "mat_id": 0})
    d.instances.append((wl.trs((0.0, 0.0, -4.0)), 0))
    times = {}
    for mode in (pkg.BUILD_HOST_SAH, pkg.BUILD_DEVICE_LBVH, pkg.BUILD_DEVICE_LBVH):
        s = pkg.Scene()
        s.set_build_mode(mode)
        t = time.perf_counter()
        d.apply(s, pkg.pbr_desc)
        times[mode] = time.perf_counter() - t
    print(f"commit {len(f)} triangles: host SAH {times[0]:.3f} s, device LBVH {times[1]:.3f} s", flush=True)
    assert times[1] < times[0]


def _prep_scene(wl, kind):
    """Meshes that exercise commitMesh's normal / tangent paths."""
    d = wl.SceneDesc()
    d.materials.append(("pbr", dict(base_color=(0.8, 0.6, 0.4, 1.0), metallic=0.0, roughness=0.5)))
    d.materials.append(("pbr", dict(base_color=(0.5, 0.5, 0.5, 1.0), metallic=0.2, roughness=0.4, textures={2: 0}, channels={2: 1})))
    d.textures.append(("rgba8", np.full((4, 4, 4), 128, np.uint8)))
    v, f, uv = wl.torus_knot(segments=300, tube=40)
    if kind == "axis":            # no normals, no uv: computed normals, axis tangents
        d.meshes.append({"vertices": v, "faces": f, "mat_id": 0})
    elif kind == "uv":            # normal map on uv channel 1: uv tangents
        d.meshes.append({"vertices": v, "faces": f, "mat_id": 1, "uvs": (uv, uv * 0.5 + 0.25)})
    elif kind == "normals_given":
        n = np.random.default_rng(3).normal(0, 1, v.shape).astype(np.float32)
        d.meshes.append({"vertices": v, "faces": f, "mat_id": 1, "normals": n, "uvs": (uv, uv[::-1].copy())})
    else:                         # degenerate faces, a vertex twice in a face, unused vertices
        ff = f.copy()
        ff[::7, 1] = ff[::7, 0]
        ff[::11, 2] = ff[::11, 1]
        vv = np.concatenate([v, np.zeros((5, 3), np.float32)])
        uu = np.concatenate([uv, np.zeros((5, 2), np.float32)])
        uu[::13] = uu[::13] * 0.0
        d.meshes.append({"vertices": vv, "faces": ff, "mat_id": 1, "uvs": (uu, uu)})
    d.instances.append((wl.trs((0.0, 0.0, -3.0)), 0))
    d.background = ("const", (0.3, 0.3, 0.3))
    return d


@pytest.mark.parametrize("kind", ["axis", "uv", "normals_given", "degenerate"])
def test_device_mesh_preprocessing_bitexact(pkg, wl, kind):
    """commitMesh's normals and tangents made on the device equal the host's
    bit for bit (per-vertex sums in face order)."""
    desc = _prep_scene(wl, kind)
    print(f"mesh preprocessing case {kind}", flush=True)
    out = []
    for mode in (pkg.BUILD_HOST_SAH, pkg.BUILD_DEVICE_LBVH):
        s = pkg.Scene()
        s.set_build_mode(mode)
        desc.apply(s, pkg.pbr_desc)
        out.append(s.mesh_vertices(0))
    (a, ma), (b, mb) = out
    assert ma == mb
    assert np.array_equal(a, b), np.argwhere(a != b)[:5].tolist()
    assert np.all(np.isfinite(a))
