"""Pins of the oracle that the reference's own code cannot provide (the
Embree-dependent translation units do not build offline): known-answer
vectors, closed forms, invariants and brute force.  Each check is chosen so
that a plausible slip (dropped term, swapped index, wrong sign) fails it.
"""
import ctypes as C
import math
import shutil
import subprocess

import numpy as np
import pytest


def f32(a):
    return np.asarray(a, np.float32)


# ----------------------------------------------------------------- Philox
# Random123 known-answer vectors for philox4x32-10 (kat_vectors): ctr, key, out
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0), (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def _philox(ref, ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    ref.lib().ref_philox4x32_10(c, k, o)
    return tuple(o)


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_kat(ref, ctr, key, out):
    assert _philox(ref, ctr, key) == out


def test_philox_matches_rocrand_header(ref, tmp_path):
    """Second, independent implementation: rocRAND's Philox4x32-10 header."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    import os
    if not os.path.exists(hipcc) or not os.path.exists("/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"):
        pytest.skip("rocRAND headers not available")
    src = tmp_path / "p.cpp"
    src.write_text(r'''
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct Open : rocrand_device::philox4x32_10_engine {
    uint4 rounds(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
    unsigned ctrs[3][4] = {{0,0,0,0},{7,1,123456,9},{0xdeadbeef,3,2,1}};
    unsigned keys[3][2] = {{0,0},{42,0},{0xabcdef,17}};
    for (int i = 0; i < 3; ++i) {
        Open e;
        uint4 c = make_uint4(ctrs[i][0], ctrs[i][1], ctrs[i][2], ctrs[i][3]);
        uint2 k = make_uint2(keys[i][0], keys[i][1]);
        uint4 o = e.rounds(c, k);
        printf("%u %u %u %u\n", o.x, o.y, o.z, o.w);
    }
}
''')
    exe = tmp_path / "p"
    r = subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1", str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("could not build the rocRAND check: " + r.stderr[-300:])
    lines = subprocess.run([str(exe)], capture_output=True, text=True).stdout.strip().splitlines()
    ctrs = [(0, 0, 0, 0), (7, 1, 123456, 9), (0xDEADBEEF, 3, 2, 1)]
    keys = [(0, 0), (42, 0), (0xABCDEF, 17)]
    for line, c, k in zip(lines, ctrs, keys):
        assert tuple(int(x) for x in line.split()) == _philox(ref, c, k)


def test_u01_range(ref):
    L = ref.lib()
    assert L.ref_u01(0) == 0.0
    assert L.ref_u01(0xFFFFFFFF) == np.float32(1.0 - 2.0 ** -24)
    assert L.ref_u01(0x80000000) == 0.5


# ----------------------------------------------------------------- Fresnel
def _fresnel(ref, n, k, c):
    rp, rs = C.c_float(), C.c_float()
    ref.lib().ref_fresnel(C.c_float(n), C.c_float(k), C.c_float(c), C.byref(rp), C.byref(rs))
    return rp.value, rs.value


@pytest.mark.parametrize("n,k", [(1.5, 0.0), (0.2, 3.4), (1.2, 7.5), (0.47, 2.4), (2.9, 3.1)])
def test_fresnel_normal_incidence(ref, n, k):
    rp, rs = _fresnel(ref, n, k, 1.0)
    closed = ((n - 1) ** 2 + k ** 2) / ((n + 1) ** 2 + k ** 2)   # Born & Wolf
    if k == 0.0:
        assert rp == rs                  # the reference's own test (OpenMaterial.cpp:42-52), n = 1.5
    else:
        assert abs(rp - rs) < 2e-6       # equal in exact arithmetic, rounded apart for complex N
    assert abs(rp - closed) < 2e-6 * max(closed, 1e-3)


@pytest.mark.parametrize("n,k", [(1.5, 0.0), (0.2, 3.4), (1.2, 7.5)])
def test_fresnel_grazing_is_one(ref, n, k):
    rp, rs = _fresnel(ref, n, k, 0.0)
    assert abs(rp - 1.0) < 1e-6 and abs(rs - 1.0) < 1e-6


@pytest.mark.parametrize("c", [0.95, 0.7, 0.4, 0.1])
def test_fresnel_dielectric_closed_form(ref, c):
    n = 1.5
    s2 = 1 - c * c
    root = math.sqrt(n * n - s2)
    rs_ = ((c - root) / (c + root)) ** 2
    rp_ = ((n * n * c - root) / (n * n * c + root)) ** 2
    rp, rs = _fresnel(ref, n, 0.0, c)
    assert abs(rs - rs_) < 3e-6 and abs(rp - rp_) < 3e-6
    assert rp < rs   # p-polarisation reflects less away from normal incidence


def test_fresnel_conductor_against_complex128(ref):
    rng = np.random.default_rng(3)
    for _ in range(200):
        n, k, c = rng.uniform(0.1, 3.0), rng.uniform(0.0, 8.0), rng.uniform(0.0, 1.0)
        N2 = complex(n, k) ** 2
        root = np.sqrt(N2 - (1 - c * c))
        rp_ = abs((N2 * c - root) / (N2 * c + root)) ** 2
        rs_ = abs((c - root) / (c + root)) ** 2
        rp, rs = _fresnel(ref, n, k, c)
        assert abs(rp - rp_) < 2e-5 and abs(rs - rs_) < 2e-5


# ----------------------------------------------------------------- warps
def test_warp_moments(ref):
    """Uniform hemisphere: E[z] = 1/2; cosine-weighted (sqrt warp): E[z] = 2/3;
    all unit length.  A swapped sin/cos or a dropped sqrt shifts these."""
    L = ref.lib()
    rng = np.random.default_rng(0)
    u = rng.random((20000, 2)).astype(np.float32)
    hz, cz = [], []
    out = np.zeros(3, np.float32)
    for a, b in u:
        L.ref_uniform_hemisphere(C.c_float(a), C.c_float(b), ref.ptr(out))
        assert abs(np.linalg.norm(out) - 1) < 1e-5
        hz.append(out[2])
        L.ref_cosine_sqrt_hemisphere(C.c_float(a), C.c_float(b), ref.ptr(out))
        cz.append(out[2])
    assert abs(np.mean(hz) - 0.5) < 0.01
    assert abs(np.mean(cz) - 2.0 / 3.0) < 0.01


def test_ggx_hemisphere_cdf(ref):
    """GGXHemisphere(alpha2): cos^2 = (1-u)/(1-(1-alpha2)u) exactly inverts
    the GGX normal CDF; alpha2 = 1 reduces to a uniform distribution of cos^2."""
    L = ref.lib()
    out = np.zeros(3, np.float32)
    for u1 in (0.0, 0.25, 0.5, 0.9):
        L.ref_ggx_hemisphere(C.c_float(0.1), C.c_float(u1), C.c_float(1.0), ref.ptr(out))
        assert abs(out[2] ** 2 - (1 - u1)) < 2e-6
        L.ref_ggx_hemisphere(C.c_float(0.1), C.c_float(u1), C.c_float(0.04), ref.ptr(out))
        assert abs(out[2] ** 2 - (1 - u1) / (1 - 0.96 * u1)) < 3e-6


def test_uniform_sphere_rejection(ref):
    L = ref.lib()
    out = np.zeros(3, np.float32)
    # first triple rejected (|v|^2 > 1), second accepted
    u = f32([0.9, 0.9, 0.9, 0.3, 0.4, 0.5])
    assert L.ref_uniform_sphere(ref.ptr(u), 6, ref.ptr(out)) == 6
    v = f32([0.3, 0.4, 0.5])
    assert np.allclose(out, v / np.linalg.norm(v), atol=1e-7)


# ----------------------------------------------------- mesh preprocessing
def _commit_one_mesh(ref, v, f, uv=None, mat="om"):
    s = ref.Scene()
    s.allocate(1, 1, 1, 0, 0)
    if mat == "om":
        s.set_material_om_nk(0, (1, 1, 1), (3, 3, 3))
    else:
        s.set_material_pbr(0, ref.pbr_desc())
    s.set_mesh(0, v, f, 0, uvs=[uv] if uv is not None else ())
    s.set_instance(0, np.eye(4, dtype=np.float32), 0)
    s.commit()
    return s.mesh_vertices(0)[0]


def test_normals_flat_quad(ref):
    v = f32([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    V = _commit_one_mesh(ref, v, f)
    assert np.allclose(V[:, 3:6], [[0, 0, 1]] * 4, atol=1e-7)
    # OM has no normal-map channel: axis tangent = (0,1,0) x n ... |n_z| >= 0.95 -> Y axis: (0,1,0)x(0,0,1) = (1,0,0)
    assert np.allclose(V[:, 6:10], [[1, 0, 0, 1]] * 4, atol=1e-7)


def test_normals_unweighted_average(ref):
    """computeNormals sums unit face normals (not area-weighted): a vertex
    shared by a big and a small face of different orientation gets the
    bisector."""
    v = f32([[0, 0, 0], [10, 0, 0], [0, 10, 0], [0, 0, 0.1]])
    f = np.array([[0, 1, 2], [0, 3, 1]], np.uint32)   # normals +z and +y
    V = _commit_one_mesh(ref, v, f)
    b = np.array([0, 1, 1]) / math.sqrt(2)
    assert np.allclose(V[0, 3:6], b, atol=1e-6) and np.allclose(V[1, 3:6], b, atol=1e-6)
    assert np.allclose(V[2, 3:6], [0, 0, 1], atol=1e-6) and np.allclose(V[3, 3:6], [0, 1, 0], atol=1e-6)


def test_tangents_from_uv_and_handedness(ref):
    v = f32([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    uv = f32([[0, 0], [1, 0], [1, 1], [0, 1]])
    V = _commit_one_mesh(ref, v, f, uv, mat="pbr")
    assert np.allclose(V[:, 6:10], [[1, 0, 0, 1]] * 4, atol=1e-6)
    uvm = f32([[1, 0], [0, 0], [0, 1], [1, 1]])      # u mirrored: tangent -x
    V = _commit_one_mesh(ref, v, f, uvm, mat="pbr")
    # mirrored u gives det < 0 for both faces: they are skipped, so the
    # tangent falls back to the axis rule with w = +1 (RenderMesh.cpp:359, 393-396)
    assert np.allclose(V[:, 6:10], [[1, 0, 0, 1]] * 4, atol=1e-6)
    uvv = f32([[0, 1], [1, 1], [1, 0], [0, 0]])      # v mirrored: det < 0 too
    V = _commit_one_mesh(ref, v, f, uvv, mat="pbr")
    assert np.allclose(V[:, 9], 1.0)


# ------------------------------------------------ traversal and ids
def test_bvh_equals_brute_force(ref, wl):
    desc, view = wl.mixed_test_scene(hdr=False, lights=False)
    s = ref.Scene()
    desc.apply(s, ref.pbr_desc)
    rng = np.random.default_rng(1)
    hits = 0
    for _ in range(3000):
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        a, _, _ = s.closest_hit(o, d, bvh=True)
        b, _, _ = s.closest_hit(o, d, bvh=False)
        assert a == b
        hits += a is not None
    assert hits > 300


def test_tie_flag_on_shared_vertex(ref, wl):
    """A ray through a vertex shared by several coplanar triangles has the
    same t on each: the hit is flagged, and the smallest primitive id wins."""
    v, f, _ = wl.plane(2, 2, 1.0, 0.0)
    s = ref.Scene()
    s.allocate(1, 1, 1, 0, 0)
    s.set_material_om_nk(0, (1, 1, 1), (3, 3, 3))
    s.set_mesh(0, v, f, 0)
    s.set_instance(0, np.eye(4, dtype=np.float32), 0)
    s.commit()
    h, _, _ = s.closest_hit((0.0, 1.0, 0.0), (0.0, -1.0, 0.0))
    assert h is not None and h[5] == 1 and h[0] == 1.0
    cands = [p for p in range(len(f)) if 4 in f[p]]
    assert h[4] == min(cands)
    h2, _, _ = s.closest_hit((0.3, 1.0, 0.2), (0.0, -1.0, 0.0))
    assert h2[5] == 0


def test_fc_global_face_ids(ref, wl):
    desc, view = wl.mixed_test_scene(hdr=False, lights=False)
    desc.fc_subject = "pid"
    s = ref.Scene()
    desc.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(64, 48, view.position, view.forward, view.up, view.yfov)
    film, ids, tie = s.render_fc(ref.default_params(), cam)
    base = np.cumsum([0] + [len(m["faces"]) for m in desc.meshes])
    hit = ids[0] >= 0
    mesh = ids[1][hit]
    local = ids[0][hit] - base[mesh]
    assert np.all(local >= 0) and np.all(local < np.array([len(desc.meshes[m]["faces"]) for m in mesh]))
    # instance -> mesh mapping, material ids (missing material reports -1)
    inst_mesh = np.array([m for _, m in desc.instances])
    assert np.array_equal(inst_mesh[ids[3][hit]], mesh)
    mats = np.array([m["mat_id"] if m["mat_id"] < len(desc.materials) else -1 for m in desc.meshes])
    assert np.array_equal(mats[mesh], ids[2][hit])
    # the pid colour is the seeded colour table's entry of the id (params.seed)
    import ctypes as C
    n = int(base[-1])
    table = np.zeros(n, np.uint32)
    assert ref.lib().ref_fc_color_table(ref.default_params().seed, 1, n, table.ctypes.data_as(C.POINTER(C.c_uint32))) == 0
    for y, x in np.argwhere(hit)[:50]:
        cid = int(table[ids[0][y, x]])
        exp = f32([((cid >> 16) & 255) / 255.0, ((cid >> 8) & 255) / 255.0, (cid & 255) / 255.0])
        assert np.array_equal(film[y, x, :3], exp)


# ------------------------------------------------ estimator closed forms
def test_constant_background_every_pixel(ref, wl):
    d = wl.SceneDesc()
    v, f = wl.icosphere(1, 0.1)
    d.materials.append(("om_nk", (0.5, 0.5, 0.5), (3.0, 3.0, 3.0)))
    d.meshes.append({"vertices": v, "faces": f, "mat_id": 0})
    d.instances.append((wl.trs((0.0, 0.0, 50.0)), 0))
    d.background = ("const", (0.5, 2.0, 8.0))
    s = ref.Scene()
    d.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(16, 12, (0, 0, 0), (0, 0, -1))
    img, st = s.render(ref.default_params(samples_per_pixel=4), cam)
    assert np.array_equal(img[..., :3], np.broadcast_to(f32([0.5, 2.0, 8.0]), img[..., :3].shape))


def test_mirror_sphere_throughput_is_fresnel(ref, wl):
    """OpenMATERIAL mirror sphere under a constant background: a camera ray
    hitting near normal incidence returns bg * F(n, k, ~1) (brdf = F/|cos|
    and ct = |cos|, OpenMaterial.cpp:147-155, BDPTRenderer.cpp:161-162)."""
    n, k = (0.3, 0.9, 1.4), (3.0, 5.0, 7.0)
    d = wl.SceneDesc()
    v, f = wl.icosphere(5, 1.0)
    d.materials.append(("om_nk", n, k))
    d.meshes.append({"vertices": v, "faces": f, "mat_id": 0})
    d.instances.append((wl.identity(), 0))
    d.background = ("const", (100.0, 100.0, 100.0))
    s = ref.Scene()
    d.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(9, 9, (0, 0, 6), (0, 0, -1), yfov=2.0)
    img, _ = s.render(ref.default_params(samples_per_pixel=8), cam)
    F = np.array([((a - 1) ** 2 + b ** 2) / ((a + 1) ** 2 + b ** 2) for a, b in zip(n, k)])
    assert np.allclose(img[4, 4, :3], 100.0 * F, rtol=2e-3)


def test_single_connection_closed_form(ref, wl):
    """A point light straight above a diffuse PBR plane, camera looking down:
    the light subpath leaves into the +octant (upwards) and escapes, the
    camera subpath hits the plane and escapes after one bounce, so the pixel
    is exactly one light connection: evaluateDirect * cos^2 * I * 100 / d
    (BDPTRenderer.cpp:266-296 with attenuation 1/d, A = 1)."""
    d = wl.SceneDesc()
    pv, pf, puv = wl.plane(1, 1, 50.0, 0.0)
    color = (0.6, 0.5, 0.4)
    d.materials.append(("pbr", dict(base_color=color + (1.0,), metallic=0.0, roughness=1.0)))
    d.meshes.append({"vertices": pv, "faces": pf, "mat_id": 0, "uvs": [puv]})
    d.instances.append((wl.identity(), 0))
    d.lights.append(((0.0, 1.0, 0.0), (2.0, 3.0, 4.0), -1.0))
    s = ref.Scene()
    d.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(3, 3, (0.0, 2.0, 0.0), (0.0, -1.0, 0.0001), yfov=1.0)
    p = ref.default_params(samples_per_pixel=1, camera_bounces=1, light_bounces=1, max_path_length=2, seed=4)
    img, st = s.render(p, cam)
    # camera ray nearly straight down hits (x, 0, z) ~ (0, 0, 0): V = L = +y
    I = np.array([2.0, 3.0, 4.0]) * 100.0 / 1.0
    c = 1.0
    # evaluateDirect with V = L = N, metal 0, alpha = 1: F = 0.04, D = 1/pi, G2 = 1
    F = 0.04
    fd = (1 - F) * np.array(color) / math.pi
    D = 1.0 / math.pi
    Vn = 1.0
    G1 = 1.0 / (1.0 + 0.5 * (math.sqrt(1.0 + (Vn / math.sqrt(1 - Vn * Vn + 1e-30)) ** -2) - 1.0))
    fs = 0.25 * F * D * G1 * G1 / (G1 + G1 - G1 * G1)
    expect = (fd + fs) * c * c * I
    assert np.allclose(img[1, 1, :3], expect, rtol=2e-3), (img[1, 1, :3], expect)
    assert st["shadow_rays"] == 9   # one connection per pixel


def test_float_vs_double_oracle(ref, wl):
    # without the missing material: its (1000, 0, 1000) albedo makes the image
    # mean a handful of enormous paths, which float/double divergence reshuffles
    desc, view = wl.mixed_test_scene(hdr=True, lights=True, missing=False)
    s = ref.Scene()
    desc.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(32, 24, view.position, view.forward, view.up, view.yfov)
    p = ref.default_params(samples_per_pixel=16, camera_bounces=4, light_bounces=4, max_path_length=4, seed=2)
    a, _ = s.render(p, cam, fp64=False)
    b, _ = s.render(p, cam, fp64=True)
    assert np.isfinite(a).all() and np.isfinite(b).all()
    rel = abs(a[..., :3].mean() - b[..., :3].mean()) / b[..., :3].mean()
    assert rel < 0.02


# ------------------------------------------------ OMV_EXT_OM_ROUGH extension
def test_om_rough_tends_to_the_mirror(ref, wl):
    """As the roughness of the rough-conductor extension goes to zero, the
    throughput F G |I.H| / (|I.N| |N.H|) tends to the mirror's F: the near
    normal-incidence pixel of the mirror-sphere scene returns bg * F."""
    n, k = (0.3, 0.9, 1.4), (3.0, 5.0, 7.0)
    d = wl.SceneDesc()
    v, f = wl.icosphere(5, 1.0)
    d.materials.append(("om_nk", n, k))
    d.meshes.append({"vertices": v, "faces": f, "mat_id": 0})
    d.instances.append((wl.identity(), 0))
    d.background = ("const", (100.0, 100.0, 100.0))
    s = ref.Scene()
    d.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(9, 9, (0, 0, 6), (0, 0, -1), yfov=2.0)
    img, _ = s.render(ref.default_params(samples_per_pixel=32, extensions=1, om_roughness=0.02), cam)
    F = np.array([((a - 1) ** 2 + b ** 2) / ((a + 1) ** 2 + b ** 2) for a, b in zip(n, k)])
    assert np.allclose(img[4, 4, :3], 100.0 * F, rtol=1e-2)


def test_om_rough_lights_the_icosphere(ref, wl):
    desc, view, params = wl.plain_bdpt_icosphere()
    s = ref.Scene()
    desc.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(32, 32, view.position, view.forward, view.up, view.yfov)
    params["samples_per_pixel"] = 4
    black, _ = s.render(ref.default_params(**params), cam)
    lit, st = s.render(ref.default_params(extensions=1, om_roughness=0.3, **params), cam)
    assert np.all(black[..., :3] == 0.0)
    assert lit[..., :3].max() > 0.0 and st["shadow_rays"] > 0


def test_median3_matches_numpy(ref):
    """SimpleDenoiserMedian<1>: interior = per-channel median of the 3x3
    window (numpy's median of 9 values is the 5th order statistic), border
    unchanged."""
    rng = np.random.default_rng(4)
    img = rng.random((13, 17, 4)).astype(np.float32)
    out = img.copy()
    ref.lib().ref_median3(ref.ptr(out), 17, 13)
    want = img.copy()
    for y in range(1, 12):
        for x in range(1, 16):
            want[y, x, :3] = np.median(img[y - 1:y + 2, x - 1:x + 2, :3].reshape(9, 3), axis=0)
            want[y, x, 3] = 1.0
    assert np.array_equal(out, want)


def test_post_u8_truncates(ref):
    f = np.array([0.0, 1.0, 0.5, 0.999, 1.0 / 255.0, 0.00392], np.float32)
    u = np.zeros(f.size, np.uint8)
    ref.lib().ref_to_u8(ref.ptr(f), f.size, u.ctypes.data_as(C.POINTER(C.c_uint8)))
    assert list(u) == [int(np.float32(255.0) * v) for v in f]
