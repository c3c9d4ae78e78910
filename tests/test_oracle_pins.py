"""Pins of four oracle functions that only the oracle itself checked before:

* the false-colour paint rules (FCSubject.h:153-482) against what the
  reference's own subject classes paint (tests/golden/fc_golden.json, written
  by oracle/_ref/fc_harness from the reference sources) and against the
  reference's own expectations (FCSubject.cpp:189-315);
* the mesh-density value (RenderMesh.cpp:197-203, BDPTRenderer.cpp:69-71) and
  the pointwise surface gradient (RenderMesh.cpp:206-212) against closed
  forms on hand-built triangles;
* the watertight ray/triangle test's (t, u, v) against an independent float64
  intersection, and its no-gap property on shared edges and vertices;
* isConnected's occluded branch (BDPTRenderer.cpp:208-235): an occluder
  between the light and the lit point blacks the pixel out, one just beyond
  the lit point changes nothing (closest hit, not any hit up to tfar).

Each check is paired with a planted mistake in the oracle (ref_pin_plant)
that it must catch.
"""
import ctypes as C
import json
import math
import pathlib

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "fc_golden.json"
FC = {"pid": 1, "gid": 2, "mid": 3, "rmp": 4, "mmp": 5, "md": 6, "sg": 7, "mn": 8, "in": 9}
FP = C.POINTER(C.c_float)


def f32(a):
    return np.asarray(a, np.float32)


@pytest.fixture
def plant(ref):
    """Sets a planted-mistake mask for the duration of one check."""
    L = ref.lib()

    class _P:
        def __enter__(self):
            return self

        def set(self, mask):
            L.ref_pin_plant(mask)

        def __exit__(self, *a):
            L.ref_pin_plant(0)

    p = _P()
    yield p
    L.ref_pin_plant(0)


@pytest.fixture(scope="module")
def fcg():
    return json.load(open(GOLDEN))


# ------------------------------------------------------------ paint rules
def _paint_clamped(ref, sub, v):
    out = (C.c_float * 3)()
    ref.lib().ref_fc_paint_clamped(FC[sub], C.c_float(v), out)
    return f32(out[:])


def _paint_enum(ref, id_, table, mid):
    out = (C.c_float * 3)()
    t = np.ascontiguousarray(table, np.uint32)
    ref.lib().ref_fc_paint_enumerated(C.c_uint32(id_), t.ctypes.data_as(C.POINTER(C.c_uint32)), len(t), int(mid), out)
    return f32(out[:])


def _paint_material(ref, kind):
    out = (C.c_float * 3)()
    ref.lib().ref_fc_paint_material(kind, out)
    return f32(out[:])


def _paint_inverted(ref, flag):
    out = (C.c_float * 3)()
    ref.lib().ref_fc_paint_inverted(int(flag), out)
    return f32(out[:])


def _clamped_mismatches(ref, fcg):
    bad = 0
    for sub in ("mmp", "rmp", "md", "sg"):
        for v, col in fcg[sub]:
            bad += not np.array_equal(_paint_clamped(ref, sub, v), f32(col))
        # a subject the material does not support receives the default (zero) argument
        bad += not np.array_equal(_paint_clamped(ref, sub, 0.0), f32(fcg["default_argument"][sub]))
    return bad


def test_clamped_subject_paints_equal_the_reference(ref, fcg):
    assert _clamped_mismatches(ref, fcg) == 0


def test_enumerated_material_name_inverted_paints_equal_the_reference(ref, fcg):
    for sub, mid in (("pid", False), ("gid", False), ("mid", True)):
        table = fcg[sub]["table"]
        for id_, col in fcg[sub]["paints"]:
            if mid and id_ == 0xFFFFFFFF:
                continue   # the missing material: pinned through the render below
            assert np.array_equal(_paint_enum(ref, id_, table, mid), f32(col)), (sub, id_)
    assert np.array_equal(f32(fcg["mid"]["paints"][-1][1]), f32([1000, 0, 1000]))
    for kind, name in ((0, "om"), (1, "pbr"), (2, "mm")):
        assert np.array_equal(_paint_material(ref, kind), f32(fcg["mn"][name])), name
    for flag in (0, 1):
        assert np.array_equal(_paint_inverted(ref, flag), f32(fcg["in"][str(flag)]))


def test_reference_unit_expectations(ref):
    """The reference's own FC unit tests (FCSubject.cpp:189-315), on the oracle."""
    assert np.array_equal(_paint_clamped(ref, "mmp", 0.5), f32([0.5, 0.5, 1.0]))   # light blue
    assert np.array_equal(_paint_clamped(ref, "rmp", 1.0), f32([0.0, 1.0, 0.0]))   # green
    # MeshDensityFC on the triangular prism: the denser face paints redder
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 1, 1]], np.float32)
    F = [(0, 2, 1), (0, 1, 3), (1, 2, 3), (0, 3, 2)]
    area = np.array([np.linalg.norm(np.cross(V[b] - V[a], V[c] - V[a])) / 2 for a, b, c in F], np.float32)
    assert int(np.argmax(area)) == 1
    D = np.float32(1.0) / (np.float32(1e-3) + area)
    assert int(np.argmax(D)) == 0
    D = D / D.max()
    c0, c1 = _paint_clamped(ref, "md", D[0]), _paint_clamped(ref, "md", D[1])
    assert np.all(c0[1:] < c1[1:])
    assert np.array_equal(_paint_material(ref, 0), f32([0, 1, 0]))
    assert np.array_equal(_paint_material(ref, 1), f32([0, 0, 1]))
    assert np.array_equal(_paint_material(ref, 2), f32([1000, 0, 1000]))
    inV = np.array([-1, -1, 0], np.float32)
    for n, want in (((1, 1, 1), [0, 1, 0]), ((-1, 0, 0), [1, 0, 0])):
        assert np.array_equal(_paint_inverted(ref, -np.dot(inV, n) < 0), f32(want))
    # MaterialIDFC: out of the table black, inside some channel > 0
    table = [0x00A0B0C0, 0x00112233, 0x00FF0000]
    assert np.array_equal(_paint_enum(ref, 5, table, True), f32([0, 0, 0]))
    for i in range(2):
        assert np.any(_paint_enum(ref, i, table, True) > 0)
    assert np.array_equal(_paint_enum(ref, 2, table, True), f32([0, 0, 0]))   # red dropped


def test_planted_paint_slip_is_caught(ref, fcg, plant):
    plant.set(1)   # rmp / mmp targets swapped
    assert _clamped_mismatches(ref, fcg) > 0


def test_missing_material_paints_pink_under_mid(ref, wl):
    desc, view = wl.mixed_test_scene(hdr=False, lights=False, missing=True)
    desc.fc_subject = "mid"
    s = ref.Scene()
    desc.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(64, 48, view.position, view.forward, view.up, view.yfov)
    film, ids, _ = s.render_fc(ref.default_params(), cam)
    miss = ids[2] == -1
    hit = ids[0] >= 0
    assert np.any(miss & hit)
    assert np.all(film[miss & hit][:, :3] == f32([1000, 0, 1000]))


# ------------------------------------------------ density and gradient
def _fc_scene(ref, meshes, instances, subject):
    """meshes: list of (vertices, faces, normals or None); one OM material."""
    s = ref.Scene()
    s.allocate(len(meshes), len(instances), 1, 0, 0)
    s.set_material_om_nk(0, (1.0, 1.0, 1.0), (3.0, 3.0, 3.0))
    for i, (v, f, n) in enumerate(meshes):
        s.set_mesh(i, v, f, 0, normals=n)
    for i, (tm, m) in enumerate(instances):
        s.set_instance(i, tm, m)
    s.set_false_color_subject(subject)
    s.commit()
    return s


def _translate(t, scale=1.0):
    M = np.eye(4, dtype=np.float32) * np.float32(scale)
    M[3, 3] = 1.0
    M[:3, 3] = t
    return M


LEGS = (0.3, 0.6, 1.0)


def _density_scene(ref):
    v, f = [], []
    x0 = 0.0
    for a in LEGS:   # right triangles in y = 0, areas a^2 / 2
        b = len(v)
        v += [(x0, 0.0, 0.0), (x0 + a, 0.0, 0.0), (x0, 0.0, -a)]
        f.append((b, b + 1, b + 2))
        x0 += a + 0.5
    inst = [(_translate((0.0, 0.0, 0.0)), 0), (_translate((0.0, 0.0, -4.0), 2.0), 0)]   # areas x1 and x4
    return _fc_scene(ref, [(f32(v), np.array(f, np.uint32), None)], inst, "md")


def _density_check(ref, s):
    cam = ref.camera_desc(160, 160, (2.2, 20.0, -3.0), (0.0, -1.0, 0.0), (0.0, 0.0, -1.0), 30.0)
    film, ids, _ = s.render_fc(ref.default_params(), cam)
    hit = ids[0] >= 0
    areas = np.array([[a * a / 2 for a in LEGS], [4 * a * a / 2 for a in LEGS]], np.float64)
    amin, amax = areas.min(), areas.max()   # densities 1 / area: max = 1 / amin, min = 1 / amax
    A = areas[ids[3][hit], ids[0][hit]]
    a = 1.0 - np.log(A / amin) / np.log(amax / amin)    # BDPTRenderer.cpp:69-71 with d = 1 / A
    want = np.stack([np.ones_like(a), 1.0 - a, 1.0 - a], -1)   # white -> red
    got = film[hit][:, :3].astype(np.float64)
    seen = {(int(i), int(p)) for i, p in zip(ids[3][hit], ids[0][hit])}
    return np.abs(got - want).max(), seen


def test_mesh_density_closed_form(ref):
    s = _density_scene(ref)
    err, seen = _density_check(ref, s)
    assert len(seen) == 6, seen          # every triangle of both instances seen
    assert err < 2e-6, err


def test_planted_linear_density_is_caught(ref, plant):
    s = _density_scene(ref)
    plant.set(2)
    err, _ = _density_check(ref, s)
    assert err > 0.1


def _tilted_normals(theta, phis):
    return f32([(math.sin(theta) * math.cos(p), math.cos(theta), math.sin(theta) * math.sin(p)) for p in phis])


def _gradient_scene(ref, normals, winding_up=True, tm=None):
    # one triangle in y = 0; (0,0,0), (0,0,1), (1,0,0) has face normal +y
    v = f32([(0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0)])
    f = np.array([(0, 1, 2)] if winding_up else [(0, 2, 1)], np.uint32)
    return _fc_scene(ref, [(v, f, normals)], [(tm if tm is not None else np.eye(4, dtype=np.float32), 0)], "sg")


def _gradient_colour(ref, s, pos=(0.3, 5.0, 0.3), fwd=(0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0)):
    cam = ref.camera_desc(48, 48, pos, fwd, up, 10.0)
    film, ids, _ = s.render_fc(ref.default_params(), cam)
    hit = ids[0] >= 0
    assert hit.sum() > 20
    return film[hit][:, :3].astype(np.float64)


def _sg_colour(g):
    return 1.0 - g


FLAT = math.pow(1e-6, 0.286)   # mean term 0, clamped to 1e-6 (RenderMesh.cpp:206-212)


def test_gradient_flat_face(ref):
    col = _gradient_colour(ref, _gradient_scene(ref, f32([(0, 1, 0)] * 3)))
    assert np.abs(col - _sg_colour(FLAT)).max() < 2e-6


def test_gradient_normals_opposite_the_face(ref):
    """Vertex normals against the winding's face normal: each term is 1, g = 1, black."""
    col = _gradient_colour(ref, _gradient_scene(ref, f32([(0, 1, 0)] * 3), winding_up=False))
    assert np.abs(col).max() < 1e-6


def test_gradient_tilted_normals(ref):
    theta = math.radians(50.0)
    col = _gradient_colour(ref, _gradient_scene(ref, _tilted_normals(theta, (0.0, 2.1, 4.2))))
    g = math.pow(math.sin(theta / 2.0), 0.286)   # each term sqrt((1 - cos t) / 2) = sin(t / 2)
    assert np.abs(col - _sg_colour(g)).max() < 2e-6


def _sheared_instance():
    """diag(3, 1, 0.5) after a 30 degree turn about z: the normal matrix (inverse
    transpose) keeps vertex normals on the face normal; the position matrix does not."""
    c, s = math.cos(math.radians(30)), math.sin(math.radians(30))
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    M = np.eye(4)
    M[:3, :3] = np.diag([3.0, 1.0, 0.5]) @ R
    return M.astype(np.float32)


def test_gradient_flat_face_under_non_uniform_scale(ref, plant):
    tm = _sheared_instance()
    s = _gradient_scene(ref, f32([(0, 1, 0)] * 3), tm=tm)
    # look at the transformed triangle along its world face normal
    V = np.array([(0, 0, 0), (0, 0, 1), (1, 0, 0)], np.float64) @ tm[:3, :3].T.astype(np.float64)
    n = np.cross(V[1] - V[0], V[2] - V[0])
    n /= np.linalg.norm(n)
    c = V.mean(0)
    pos = c + 5.0 * n
    up = np.cross(n, (1.0, 0.0, 0.0))
    col = _gradient_colour(ref, s, pos=pos, fwd=-n, up=up)
    assert np.abs(col - _sg_colour(FLAT)).max() < 2e-6
    plant.set(64)   # vertex normals by the position matrix
    col = _gradient_colour(ref, _gradient_scene(ref, f32([(0, 1, 0)] * 3), tm=tm), pos=pos, fwd=-n, up=up)
    assert np.abs(col - _sg_colour(FLAT)).max() > 0.1


# ------------------------------------------------ watertight triangle test
def _tri_mesh_scene(ref, v, f):
    s = ref.Scene()
    s.allocate(1, 1, 1, 0, 0)
    s.set_material_om_nk(0, (1, 1, 1), (3, 3, 3))
    s.set_mesh(0, v, f, 0)
    s.set_instance(0, np.eye(4, dtype=np.float32), 0)
    s.commit()
    return s


def _mt64(O, D, V0, V1, V2):
    """Moller-Trumbore in float64 for all triangles: (t, u, v) with the hit
    point (1-u-v) V0 + u V1 + v V2 (u weighs the second vertex, v the third)."""
    e1, e2 = V1 - V0, V2 - V0
    p = np.cross(D, e2)
    det = np.einsum("ij,ij->i", e1, p)
    inv = 1.0 / det
    sv = O - V0
    u = np.einsum("ij,ij->i", sv, p) * inv
    q = np.cross(sv, e1)
    v = (q @ D) * inv
    t = np.einsum("ij,ij->i", e2, q) * inv
    return t, u, v


def _triangle_cases(seed=7, n_tri=300, n_rays=3000):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-2, 2, (n_tri, 1, 3))
    v = (c + rng.normal(scale=0.35, size=(n_tri, 3, 3))).reshape(-1, 3).astype(np.float32)
    f = np.arange(3 * n_tri, dtype=np.uint32).reshape(-1, 3)
    rays = []
    for i in range(n_rays):
        k = int(rng.integers(n_tri))
        b = rng.dirichlet((1, 1, 1)) if i % 3 else rng.dirichlet((0.2, 0.2, 0.2))   # a third near edges
        P = b @ v[3 * k:3 * k + 3].astype(np.float64)
        D = rng.normal(size=3)
        D /= np.linalg.norm(D)
        O = P - rng.uniform(0.5, 4.0) * D
        rays.append((f32(O), f32(D)))
    return v, f, rays


def _triangle_errors(ref, s, v, rays):
    V = v.astype(np.float64).reshape(-1, 3, 3)
    errs_t, errs_uv, wrong = [], [], 0
    robust = 0
    for O32, D32 in rays:
        O, D = O32.astype(np.float64), D32.astype(np.float64)
        t, u, w = _mt64(O, D, V[:, 0], V[:, 1], V[:, 2])
        m = np.minimum(np.minimum(u, w), 1.0 - u - w)
        ok = (t > 1e-6) & (m >= 0)
        amb = (t > 1e-6) & (np.abs(m) < 1e-5)   # too close to an edge to demand a verdict
        if not ok.any():
            if amb.any():
                continue
            h, _, _ = s.closest_hit(O32, D32)
            robust += 1
            wrong += h is not None
            continue
        k = int(np.argmin(np.where(ok, t, np.inf)))
        others = np.where(ok, t, np.inf)
        others[k] = np.inf
        if m[k] < 1e-4 or (amb & (t < t[k] * (1 + 1e-4))).sum() > (1 if amb[k] else 0) or others.min() < t[k] * (1 + 1e-4):
            continue   # near an edge or a near-tie: no robust verdict
        robust += 1
        h, _, _ = s.closest_hit(O32, D32)
        if h is None or h[4] != k:
            wrong += 1
            continue
        errs_t.append(abs(h[0] - t[k]) / t[k])
        errs_uv.append(max(abs(h[1] - u[k]), abs(h[2] - w[k])))
    return robust, wrong, np.array(errs_t), np.array(errs_uv)


def test_triangle_t_u_v_against_float64(ref):
    v, f, rays = _triangle_cases()
    s = _tri_mesh_scene(ref, v, f)
    robust, wrong, et, euv = _triangle_errors(ref, s, v, rays)
    assert robust > 2000, robust
    assert wrong == 0
    # float32 against float64: ~1e-7 typically, up to ~3e-5 at grazing incidence
    assert np.quantile(et, 0.99) < 2e-6 and et.max() < 1e-4, (np.quantile(et, 0.99), et.max())
    assert np.median(euv) < 1e-6 and np.quantile(euv, 0.99) < 1e-5 and euv.max() < 1e-3, (np.median(euv), euv.max())


@pytest.mark.parametrize("mask", [4, 8])
def test_planted_triangle_slips_are_caught(ref, plant, mask):
    v, f, rays = _triangle_cases(n_rays=400)
    s = _tri_mesh_scene(ref, v, f)
    plant.set(mask)   # 4: (v, u) returned for (u, v); 8: t * (1 + 2^-10)
    _, wrong, et, euv = _triangle_errors(ref, s, v, rays)
    if mask == 4:
        assert euv.max() > 1e-2
    else:
        assert et.max() > 5e-4


def test_no_gaps_on_shared_edges_and_vertices(ref, wl):
    """Rays through the shared edges and vertices of a triangulated grid
    always hit (the watertight property), with the plane's exact t."""
    v, f, _ = wl.plane(6, 6, 1.0, 0.0)
    s = _tri_mesh_scene(ref, v, f)
    rng = np.random.default_rng(3)
    pts = []
    for a, b, c in f:
        for p, q in ((a, b), (b, c), (c, a)):
            pts.append((v[p].astype(np.float64) + v[q]) / 2)   # edge midpoints (exact in float)
        pts.append(v[a].astype(np.float64))
    pts = [P for P in pts if abs(P[0]) < 0.999 and abs(P[2]) < 0.999]   # interior edges and vertices only
    assert len(pts) > 200
    for P in pts:
        for _ in range(3):
            D = np.array([rng.uniform(-0.4, 0.4), -1.0, rng.uniform(-0.4, 0.4)])
            O = f32(P - 2.0 * D)
            h, _, _ = s.closest_hit(O, f32(D))
            assert h is not None, (P, D)
            assert abs(h[0] - float(O[1])) < 1e-5   # |d_y| = 1: t is the height of the origin


# ------------------------------------------------ isConnected
def _connect_scene(ref, wl, extra):
    d = wl.SceneDesc()
    d.materials.append(("pbr", dict(base_color=(0.8, 0.8, 0.8, 1.0), metallic=0.0, roughness=1.0)))
    v, f, _ = wl.plane(1, 1, 5.0, 0.0)
    d.meshes.append({"vertices": v, "faces": f, "mat_id": 0})
    d.instances.append((wl.identity(), 0))
    if extra == "between":      # a 2 x 2 quad just under the light: shadows the whole visible floor
        v2, f2, _ = wl.plane(1, 1, 1.0, 4.5)
    elif extra == "beyond":     # a second floor 0.05 under the lit one, inside tfar = 1.1 len
        v2, f2, _ = wl.plane(1, 1, 5.0, -0.05)
    else:
        v2 = None
    if v2 is not None:
        d.meshes.append({"vertices": v2, "faces": f2, "mat_id": 0})
        d.instances.append((wl.identity(), 1))
    d.lights.append(((0.0, 5.0, 0.0), (50.0, 50.0, 50.0), -1.0))
    s = ref.Scene()
    d.apply(s, ref.pbr_desc)
    return s


def _connect_render(ref, s):
    cam = ref.camera_desc(48, 32, (0.0, 1.0, 6.0), (0.0, -1.0, -6.0), (0.0, 1.0, 0.0), 30.0)
    p = ref.default_params(samples_per_pixel=2, camera_bounces=1, light_bounces=1, max_path_length=2, seed=3)
    img, st = s.render(p, cam)
    return img[..., :3]


def test_is_connected_occluded_and_beyond(ref, wl):
    free = _connect_render(ref, _connect_scene(ref, wl, None))
    lit = free.max(-1) > 0
    assert lit.mean() > 0.3                 # the floor is lit by the connection
    blocked = _connect_render(ref, _connect_scene(ref, wl, "between"))
    assert np.all(blocked == 0.0)           # every connection crosses the occluder
    beyond = _connect_render(ref, _connect_scene(ref, wl, "beyond"))
    assert np.array_equal(beyond, free)     # the closest hit is the lit point itself


def test_planted_visibility_slips_are_caught(ref, wl, plant):
    free = _connect_render(ref, _connect_scene(ref, wl, None))
    plant.set(16)   # visible whatever the closest hit
    assert np.any(_connect_render(ref, _connect_scene(ref, wl, "between")) > 0)
    plant.set(32)   # any hit up to tfar occludes
    assert not np.array_equal(_connect_render(ref, _connect_scene(ref, wl, "beyond")), free)
