"""The C-ABI boundary on the host (no GPU needed): every symbol declared in
include/omv.h is exported, argument errors map to status codes, and the
commit-time mesh preprocessing of the product (scene.cpp) equals the
oracle's bit for bit.  Device work is not exercised here."""
import ctypes as C
import pathlib
import re

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def _declared(header, prefix):
    text = (ROOT / header).read_text()
    return sorted(set(re.findall(rf"\b({prefix}_[a-z0-9_]+)\s*\(", text)))


def test_product_exports_every_declared_symbol(pkg):
    names = _declared("include/omv.h", "omv")
    assert len(names) >= 40
    lib = C.CDLL(str(pkg.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_every_declared_symbol(ref):
    names = _declared("oracle/refcpu.h", "ref")
    lib = ref.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_product_does_not_link_the_oracle(pkg):
    data = pkg.LIB_PATH.read_bytes()
    assert b"refcpu" not in data and b"ref_render" not in data


def test_defaults_match_reference_parameters(pkg, ref):
    p = pkg.default_params()
    # RendererParameters.h:24-96
    assert (p.camera_bounces, p.light_bounces, p.max_path_length, p.samples_per_pixel) == (10, 10, 8, 20)
    assert (p.light_distance_attenuation, p.number_of_cores) == (1, 16)
    assert np.float32(p.light_scale_value) == np.float32(100.0) and np.float32(p.ray_cut_pix_value) == np.float32(0.002)
    assert np.float32(p.light_min_distance) == np.float32(0.01) and np.float32(p.gamma) == np.float32(0.5)
    q = ref.default_params()
    assert bytes(p) == bytes(q)


def test_status_codes(pkg):
    L = pkg.lib()
    s = C.c_void_p()
    assert L.omv_scene_create(C.byref(s)) == 0
    assert L.omv_scene_allocate(s, 1, 1, 1, 0, 0) == 0
    f = np.zeros(3, np.uint32)
    v = np.zeros(9, np.float32)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    arr = (C.POINTER(C.c_float) * 1)()
    # out-of-range ids (the reference silently ignores these, RenderScene.cpp:88-125)
    assert L.omv_scene_set_mesh(s, 5, 3, 1, 0, 0, P(f, C.c_uint32), P(v, C.c_float), 3, None, 3, None, 4, arr, 2) == -2
    assert L.omv_scene_set_instance(s, 3, P(np.eye(4, dtype=np.float32), C.c_float), 0) == -2
    assert L.omv_scene_set_material_pbr(s, 1, None) == -1
    d = pkg.pbr_desc()
    assert L.omv_scene_set_material_pbr(s, 1, C.byref(d)) == -2   # slot 1 is the reserved missing material
    bad = np.array([0, 1, 7], np.uint32)                            # face index past the vertex count
    assert L.omv_scene_set_mesh(s, 0, 3, 1, 0, 0, P(bad, C.c_uint32), P(v, C.c_float), 3, None, 3, None, 4, arr, 2) == -1
    assert b"face index" in L.omv_last_error()
    # nothing valid to render: commit -> OMV_E_SCENE (RenderScene.cpp:194)
    assert L.omv_scene_commit(s, None, 0) == -5
    assert L.omv_scene_destroy(s) == 0
    assert L.omv_scene_allocate(None, 1, 1, 1, 1, 1) == -1


def _host_commit(pkg, desc):
    """Commit on the host; without a device the upload fails with
    OMV_E_DEVICE after the preprocessing, which is what is compared."""
    s = pkg.Scene()
    try:
        desc.apply(s, pkg.pbr_desc)
    except RuntimeError as e:
        assert "(-4)" in str(e) or "failed (0)" in str(e), e
    return s


@pytest.mark.parametrize("which", ["mixed", "knot", "instanced"])
def test_commit_preprocessing_matches_oracle(pkg, ref, wl, which):
    if which == "mixed":
        desc, _ = wl.mixed_test_scene(masked=True)
    elif which == "knot":
        desc, _, _ = wl.hdri_torus_knot(tex=64, hdr_w=64, hdr_h=32, segments=120, tube=12)
    else:
        desc, _, _ = wl.large_instanced(count=6, tris_per=800, hdr_w=64, hdr_h=32)
    s = _host_commit(pkg, desc)
    r = ref.Scene()
    desc.apply(r, ref.pbr_desc)
    for i in range(len(desc.meshes)):
        a, ma = s.mesh_vertices(i)
        b, mb = r.mesh_vertices(i)
        assert ma == mb
        assert np.array_equal(a, b), f"mesh {i}: committed vertex buffers differ"


def test_camera_desc_matches_oracle(pkg, ref):
    for (w, h, pos, fwd, fov) in [(64, 48, (0.3, 0.9, 4.5), (-0.05, -0.25, -1.0), 55.0), (512, 512, (0, 0.4, 5), (0, -0.08, -1), 40.0),
                                  (1920, 1080, (0, 6, 12), (0, -0.35, -1), 50.0)]:
        c = pkg.Camera(w, h, pos, fwd, (0, 1, 0), fov)
        a = c.desc()
        b = ref.camera_desc(w, h, pos, fwd, (0, 1, 0), fov)
        assert bytes(a) == bytes(b)
    cam = pkg.Camera(8, 8)
    assert pkg.lib().omv_camera_set_yfov(cam.h, C.c_float(180.0)) == -1   # PinholeCamera.cpp:169-173


def test_batch_status_codes(pkg):
    """Batch entry points reject bad state and arguments before any device work."""
    L = pkg.lib()
    b, s, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert L.omv_batch_create(C.byref(b)) == 0
    assert L.omv_scene_create(C.byref(s)) == 0
    assert L.omv_camera_create(C.byref(c), 8, 8) == 0
    p = pkg.default_params()
    jid = C.c_int(-1)
    assert L.omv_batch_add_job(b, s, c, C.byref(p), C.byref(jid)) == -3      # scene not committed
    assert L.omv_batch_add_job(b, None, c, C.byref(p), C.byref(jid)) == -1
    n = C.c_int(-1)
    assert L.omv_batch_job_count(b, C.byref(n)) == 0 and n.value == 0
    assert L.omv_batch_run(b) == -3                                           # no jobs
    assert L.omv_batch_autotune(b, 1, None) == -3
    t = pkg.BatchTuning(0, 0, 1)
    assert L.omv_batch_set_tuning(b, C.byref(t)) == -2                        # tile_waves < 1
    t = pkg.BatchTuning(2, 4, 0)
    assert L.omv_batch_set_tuning(b, C.byref(t)) == 0
    t2 = pkg.BatchTuning()
    assert L.omv_batch_get_tuning(b, C.byref(t2)) == 0
    assert (t2.blocks_per_cu, t2.tile_waves, t2.use_graph) == (2, 4, 0)
    out = np.zeros(4, np.float32)
    assert L.omv_batch_get_film(b, 0, out.ctypes.data_as(C.POINTER(C.c_float)), 4) == -2
    for f, h in ((L.omv_batch_destroy, b), (L.omv_scene_destroy, s), (L.omv_camera_destroy, c)):
        assert f(h) == 0


def test_batched_lightbox_workload(wl):
    """256 jobs = 32 assets x 8 lightbox views; ranges as specified; the
    viewpoints follow main.cpp:434-457 (camera at 2p - C looking at C)."""
    assets, jobs = wl.batched_lightbox()
    assert len(assets) == 32 and len(jobs) == 256
    tris = [len(a.meshes[0]["faces"]) for a in assets]
    assert min(tris) >= 800 and max(tris) <= 240000
    for a, view, p in jobs:
        assert 128 <= view.width == view.height <= 512
        assert 8 <= p["samples_per_pixel"] <= 64 and 4 <= p["max_path_length"] <= 8
    assert len({p["seed"] for _, _, p in jobs}) == 256
    v = assets[0].meshes[0]["vertices"]
    lo, hi = v.min(0), v.max(0)
    c = (lo + hi) * np.float32(0.5)
    pos = np.array([j[1].position for j in jobs if j[0] == 0], np.float32)
    corners = (pos + c) * np.float32(0.5)
    assert np.allclose(np.sort(np.abs(corners - c), axis=0), np.sort(np.abs(np.array([hi - c] * 8)), axis=0), rtol=1e-5)
    kinds = {a.materials[0][0] for a in assets}
    assert kinds == {"om_nk", "pbr"}


def test_schedule_switch(pkg):
    L = pkg.lib()
    r = C.c_void_p()
    assert L.omv_renderer_create(C.byref(r)) == 0
    v = C.c_int(-1)
    assert L.omv_renderer_get_schedule(r, C.byref(v)) == 0 and v.value == pkg.SCHEDULE_WAVEFRONT
    assert L.omv_renderer_set_schedule(r, 7) == -1
    assert L.omv_renderer_set_schedule(r, pkg.SCHEDULE_PER_SAMPLE) == 0
    assert L.omv_renderer_get_schedule(r, C.byref(v)) == 0 and v.value == pkg.SCHEDULE_PER_SAMPLE
    assert L.omv_renderer_destroy(r) == 0


def test_tuning_round_trip_and_ranges(pkg):
    """omv_renderer_set_tuning keeps what it accepts and rejects values out of
    range (chunk streams beyond two, a negative stagger, an unknown traversal
    or sort mode) without changing the tuning it holds."""
    L = pkg.lib()
    r = C.c_void_p()
    assert L.omv_renderer_create(C.byref(r)) == 0
    t = pkg.RenderTuning()
    assert L.omv_renderer_get_tuning(r, C.byref(t)) == 0
    assert (t.chunk_streams, t.chunk_stagger, t.chunk_samples) == (0, 0, 0)
    t.chunk_streams, t.chunk_stagger, t.chunk_samples, t.trace_keep = 2, 3, 1000, 32
    assert L.omv_renderer_set_tuning(r, C.byref(t)) == 0
    got = pkg.RenderTuning()
    assert L.omv_renderer_get_tuning(r, C.byref(got)) == 0
    assert (got.chunk_streams, got.chunk_stagger, got.chunk_samples, got.trace_keep) == (2, 3, 1000, 32)
    for field, bad in (("chunk_streams", 3), ("chunk_stagger", -1), ("traversal", 3), ("sort_rays", 3), ("trace_keep", 65)):
        b = pkg.RenderTuning()
        assert L.omv_renderer_get_tuning(r, C.byref(b)) == 0
        setattr(b, field, bad)
        assert L.omv_renderer_set_tuning(r, C.byref(b)) == -2, field   # OMV_E_RANGE
        assert L.omv_renderer_get_tuning(r, C.byref(got)) == 0
        assert (got.chunk_streams, got.chunk_stagger) == (2, 3)
    assert L.omv_renderer_destroy(r) == 0
