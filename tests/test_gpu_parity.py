"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same inputs and the same Philox streams.

Bars:
  * false-colour ids (primitive, mesh, material, instance) and tie flags:
    bit-exact; colours of the integer/boolean/texture subjects bit-exact;
    md / sg colours (logf / powf inside) within 1e-5 absolute.
  * radiance: per-pixel relative error |gpu - cpu| / max(|cpu|, 1e-3) below
    1e-3 for >= 99% of pixels, and the image means within 1e-3 relative.
    Exact equality is not expected: device cosf/sinf/acosf/powf differ from
    glibc in the last ulp, and connections are summed in camera-vertex order.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _render_pair(pkg, ref, desc, view, **params):
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(s)
    img = r.render(cam)
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    rimg, rst = rs.render(ref.default_params(**params), cam.desc(), threads=8)
    return img, rimg, r.stats(), rst


def _radiance_close(img, rimg, frac_min=0.99, rel=1e-3):
    a, b = img[..., :3].astype(np.float64), rimg[..., :3].astype(np.float64)
    assert np.all(np.isfinite(a)), "non-finite GPU radiance"
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    frac = float(np.mean(err.max(axis=-1) < rel))
    ma, mb = a.mean(), b.mean()
    assert frac >= frac_min, f"only {frac:.4f} of pixels within {rel} (worst {err.max():.3e})"
    assert abs(ma - mb) <= rel * max(abs(mb), 1e-6), f"image means differ: {ma} vs {mb}"
    return frac


@pytest.mark.parametrize("subject", ["pid", "gid", "mid", "mn", "in", "rmp", "mmp", "md", "sg", "iid", "ipid"])
def test_fc_pass_bitexact(pkg, wl, ref, subject):
    desc, view = wl.mixed_test_scene(hdr=False, lights=False, masked=True)
    view.width, view.height = 96, 72
    desc.fc_subject = subject
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(seed=11))
    r.set_scene(s)
    film, ids, tie = r.render_fc(cam)
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    rfilm, rids, rtie = rs.render_fc(ref.default_params(seed=11), cam.desc())
    assert np.array_equal(ids, rids)
    assert np.array_equal(tie, rtie)
    assert (ids[0] >= 0).mean() > 0.3   # the view does hit geometry
    if subject in ("md", "sg"):
        np.testing.assert_allclose(film, rfilm, atol=1e-5, rtol=0)
    else:
        assert np.array_equal(film, rfilm)


@pytest.mark.parametrize("subject", ["pid", "ipid", "iid"])
def test_fc_instance_aware_colours(pkg, wl, ref, subject):
    """pid paints every instance of a mesh alike (per-mesh face ids, as the
    reference); ipid numbers faces over instances and iid paints instances:
    the colour at every hit pixel is the seeded table entry of that id."""
    desc, view, _ = wl.large_instanced(width=160, height=90, count=24, tris_per=2000)
    desc.instances = [(tm, m % 3) for tm, m in desc.instances]     # 24 instances of 3 meshes
    desc.fc_subject = subject
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(seed=12))
    r.set_scene(s)
    film, ids, _ = r.render_fc(cam)
    pid, mesh, inst = ids[0], ids[1], ids[3]
    hit = pid >= 0
    fn = np.array([len(m["faces"]) for m in desc.meshes])
    mesh_base = np.concatenate([[0], np.cumsum(fn)])
    inst_fn = np.array([fn[m] for _, m in desc.instances])
    inst_base = np.concatenate([[0], np.cumsum(inst_fn)])
    if subject == "pid":
        key, size = pid, int(fn.sum())
    elif subject == "iid":
        key, size = inst, len(desc.instances)
    else:
        key, size = inst_base[np.maximum(inst, 0)] + (pid - mesh_base[np.maximum(mesh, 0)]), int(inst_fn.sum())
    table = pkg.fc_color_table(12, {"pid": 1, "iid": 10, "ipid": 11}[subject], size)
    c = table[key[hit]]
    want = np.stack([(c >> 16) & 255, (c >> 8) & 255, c & 255], -1).astype(np.float32) / np.float32(255.0)
    assert np.array_equal(film[..., :3][hit], want)
    assert np.all(film[..., :3][~hit] == 0)
    assert np.unique(inst[hit]).size >= 4 and np.unique(mesh[hit]).size == 3   # instances share meshes in view
    if subject == "ipid":   # the same face of two instances of one mesh: one pid, two ipid colours
        pairs = {}
        for p_, i_, c_ in zip(pid[hit], inst[hit], c):
            pairs.setdefault(p_, {})[i_] = c_
        shared = [v for v in pairs.values() if len(v) > 1]
        assert shared and all(len(set(v.values())) == len(v) for v in shared)
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    rfilm, rids, _ = rs.render_fc(ref.default_params(seed=12), cam.desc())
    assert np.array_equal(film, rfilm) and np.array_equal(ids, rids)


def test_fc_ids_large_instanced(pkg, wl, ref):
    desc, view, _ = wl.large_instanced(width=160, height=90, count=24, tris_per=2000)
    desc.fc_subject = "pid"
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(seed=5))
    r.set_scene(s)
    film, ids, tie = r.render_fc(cam)
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    rfilm, rids, rtie = rs.render_fc(ref.default_params(seed=5), cam.desc())
    assert np.array_equal(ids, rids)
    assert np.array_equal(tie, rtie)
    assert np.array_equal(film, rfilm)


def test_constant_background_closed_form(pkg, wl, ref):
    """No geometry in view: every camera path escapes at once, so every pixel
    is exactly the Background radiance (Background.h getRadiance, times a
    throughput of exactly 1)."""
    d = wl.SceneDesc()
    v, f = wl.icosphere(1, 0.1)
    d.materials.append(("om_nk", (0.5, 0.5, 0.5), (3.0, 3.0, 3.0)))
    d.meshes.append({"vertices": v, "faces": f, "mat_id": 0})
    d.instances.append((wl.trs((0.0, 0.0, 50.0)), 0))     # behind the camera
    d.background = ("const", (0.25, 100.0, 7.5))
    view = wl.View(40, 30, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), yfov=60.0)
    img, rimg, st, _ = _render_pair(pkg, ref, d, view, samples_per_pixel=4, camera_bounces=4, light_bounces=4, max_path_length=4)
    expect = np.array([0.25, 100.0, 7.5], np.float32)
    assert np.array_equal(img[..., :3], np.broadcast_to(expect, img[..., :3].shape))
    assert np.array_equal(rimg, img)
    assert st["camera_rays"] == 40 * 30 * 4


@pytest.mark.parametrize("hdr,lights,f32", [(True, True, False), (False, True, False), (True, False, True)])
def test_bdpt_radiance_mixed(pkg, wl, ref, hdr, lights, f32):
    desc, view = wl.mixed_test_scene(hdr=hdr, lights=lights, tex_f32=f32, masked=True)
    img, rimg, st, rst = _render_pair(pkg, ref, desc, view, samples_per_pixel=8, camera_bounces=4, light_bounces=4,
                                      max_path_length=4, seed=21)
    _radiance_close(img, rimg)
    assert st["stack_overflows"] == 0
    # ray counts agree up to the rare paths that diverge at the last ulp
    assert abs(st["camera_rays"] - rst["camera_rays"]) <= 0.002 * rst["camera_rays"] + 4
    assert abs(st["light_rays"] - rst["light_rays"]) <= 0.002 * rst["light_rays"] + 4


def test_bdpt_radiance_lights_above_and_inside(pkg, wl, ref):
    """One light above the whole scene (its light rays all leave it and are
    culled before their directions are drawn) and one inside it (its light
    subpaths bounce): radiance against the oracle.  The mesh with the
    missing material is given a material here: under point lights it makes
    fireflies (5 pixels holding 88 % of the oracle image's sum) whose
    values follow last-ulp differences of their paths, which the image-mean
    bar cannot take."""
    desc, view = wl.mixed_test_scene(hdr=False, lights=False, masked=True, missing=False)
    desc.lights.append(((3.0, 4.0, 3.5), (40.0, 40.0, 40.0), -1.0))
    desc.lights.append(((-1.2, -0.6, -1.0), (30.0, 20.0, 10.0), -1.0))
    img, rimg, st, rst = _render_pair(pkg, ref, desc, view, samples_per_pixel=8, camera_bounces=4, light_bounces=4,
                                      max_path_length=5, seed=17)
    _radiance_close(img, rimg)
    assert st["stack_overflows"] == 0
    assert abs(st["light_rays"] - rst["light_rays"]) <= 0.002 * rst["light_rays"] + 4
    assert rst["light_rays"] > view.width * view.height * 8


def test_bdpt_radiance_torus_knot_small(pkg, wl, ref):
    desc, view, params = wl.hdri_torus_knot(width=64, height=64, spp=4, depth=5, tex=128, hdr_w=256, hdr_h=128, segments=200, tube=20)
    params["samples_per_pixel"] = 4
    img, rimg, st, rst = _render_pair(pkg, ref, desc, view, **params)
    _radiance_close(img, rimg, frac_min=0.98)


def test_bdpt_icosphere_reference_mode_is_black(pkg, wl, ref):
    """OpenMATERIAL mirrors never pass the connection gate and the camera path
    escapes with no background: the reference-semantics image is exactly 0."""
    desc, view, params = wl.plain_bdpt_icosphere()
    view.width, view.height = 64, 64
    params["samples_per_pixel"] = 2
    img, rimg, st, rst = _render_pair(pkg, ref, desc, view, **params)
    assert np.array_equal(img[..., :3], np.zeros_like(img[..., :3]))
    assert np.array_equal(rimg, img)
    assert st["light_rays"] == 64 * 64 * 2


def test_bdpt_icosphere_om_rough_extension(pkg, wl, ref):
    desc, view, params = wl.plain_bdpt_icosphere()
    view.width, view.height = 64, 64
    params.update(samples_per_pixel=4, extensions=1, om_roughness=0.3)
    img, rimg, st, rst = _render_pair(pkg, ref, desc, view, **params)
    assert img[..., :3].max() > 0.0
    _radiance_close(img, rimg)


# ---------------------------------------------------------------- full sizes
def _oracle_rows(ref, desc, cam_desc, params, row0, rows):
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    img, st = rs.render(ref.default_params(**params), cam_desc, row0=row0, rows=rows, threads=16)
    return img[row0:row0 + rows], st


def test_full_size_hdri_torus_knot(pkg, wl, ref):
    """The headline workload at its full size (512x512, 64 spp, depth 8): the
    whole frame on the GPU, a band of rows checked against the oracle, and
    size-independent properties over the frame."""
    desc, view, params = wl.hdri_torus_knot()
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(s)
    img = r.render(cam)
    st = r.stats()
    assert st["samples"] == 512 * 512 * 64 and st["stack_overflows"] == 0
    assert np.isfinite(img).all() and (img[..., :3] >= 0).all() and np.all(img[..., 3] == 1.0)
    assert st["camera_rays"] >= st["samples"]
    band, rst = _oracle_rows(ref, desc, cam.desc(), params, 248, 16)
    _radiance_close(img[248:264], band, frac_min=0.98)


def test_full_size_large_instanced_ids(pkg, wl, ref):
    """2M-triangle, 200-instance scene at 1920x1080: primitive, mesh,
    material and instance ids and tie flags bit-exact over the whole frame."""
    desc, view, params = wl.large_instanced()
    desc.fc_subject = "pid"
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(seed=9))
    r.set_scene(s)
    film, ids, tie = r.render_fc(cam)
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    rfilm, rids, rtie = rs.render_fc(ref.default_params(seed=9), cam.desc())
    assert (ids[0] >= 0).mean() > 0.2
    assert np.array_equal(ids, rids)
    assert np.array_equal(tie, rtie)
    assert np.array_equal(film, rfilm)


def test_full_size_large_instanced_radiance(pkg, wl, ref):
    """Same scene, full BDPT (4 point lights + HDRI, 16 spp, depth 8): the
    frame on the GPU, a band of rows against the oracle."""
    desc, view, params = wl.large_instanced()
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(s)
    img = r.render(cam)
    st = r.stats()
    assert st["stack_overflows"] == 0 and st["shadow_rays"] > 0
    assert np.isfinite(img).all() and (img[..., :3] >= 0).all()
    band, _ = _oracle_rows(ref, desc, cam.desc(), params, 536, 4)
    _radiance_close(img[536:540], band, frac_min=0.98)


@pytest.mark.parametrize("denoise", [False, True])
def test_post_pass_matches_oracle(pkg, wl, ref, denoise):
    """Device gammaCorrection (+ median) + 8-bit against the host post pass on
    the same film: floats within 2 ulp (device powf vs glibc powf), the median
    selects the same window element, 8-bit values equal except where the
    float sits within 2 ulp of a k/255 step."""
    desc, view = wl.mixed_test_scene(hdr=True, lights=True)
    view.width, view.height = 40, 30
    s = pkg.Scene()
    desc.apply(s, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(samples_per_pixel=2, camera_bounces=3, light_bounces=3, max_path_length=4))
    r.set_scene(s)
    film = r.render(cam)
    f, u = r.post(cam, gamma=0.5, denoise=denoise)
    rf, ru = ref.post(film, gamma=0.5, denoise=denoise)
    ulp = np.abs(f.view(np.int32).astype(np.int64) - rf.view(np.int32).astype(np.int64))
    assert ulp.max() <= 2, ulp.max()
    diff = u.astype(int) != ru.astype(int)
    near = np.abs(np.float32(255.0) * rf - np.round(np.float32(255.0) * rf)) <= 2 * np.spacing(np.float32(255.0) * rf)
    assert not np.any(diff & ~near)
    assert (diff.mean() < 0.01)
