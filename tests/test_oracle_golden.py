"""The CPU oracle against golden vectors produced by the reference's own code
(tests/golden/reference_golden.json, written by oracle/_ref/golden_harness
from the reference sources; see tests/golden/make_golden.sh).

Every comparison is exact (float32 bit equality) unless a line says
otherwise: the oracle restates the reference's arithmetic in the same order
with FMA contraction off, and both use glibc's libm.
"""
import ctypes as C

import numpy as np
import pytest


def f32(a):
    return np.asarray(a, np.float32)


def test_evaluate_direct(ref, golden):
    L = ref.lib()
    bad = 0
    for g in golden["evaluate_direct"]:
        out = np.zeros(3, np.float32)
        L.ref_evaluate_direct(ref.ptr(f32(g["I"])), ref.ptr(f32(g["O"])), ref.ptr(f32(g["N"])), ref.ptr(f32(g["color"])),
                              C.c_float(g["metal"]), C.c_float(g["alpha"]), ref.ptr(out))
        exp = f32(g["brdf"])
        if not np.array_equal(out, exp):
            bad += 1
            assert np.allclose(out, exp, rtol=2e-6, atol=0), (g, out)
    assert bad == 0, f"{bad} evaluateDirect cases differ in the last ulp"


def test_sample_indirect(ref, golden):
    L = ref.lib()
    for g in golden["sample_indirect"]:
        brdf, O = np.zeros(3, np.float32), np.zeros(3, np.float32)
        u = f32(g["uniforms"] + [0.5] * 8)
        used = L.ref_sample_indirect(ref.ptr(f32(g["I"])), ref.ptr(f32(g["N"])), ref.ptr(f32(g["T"])), ref.ptr(f32(g["color"])),
                                     C.c_float(g["metal"]), C.c_float(g["alpha"]), C.c_float(g["eta"]), g["transmissive"],
                                     ref.ptr(u), ref.ptr(brdf), ref.ptr(O))
        assert used == g["used"], g
        assert np.array_equal(O, f32(g["O"])), (g, O)
        assert np.array_equal(brdf, f32(g["brdf"])), (g, brdf)


def test_warps(ref, golden):
    L = ref.lib()
    for g in golden["warps"]:
        u = f32(g["uniforms"] + [0.5] * 4)
        out = np.zeros(3, np.float32)
        kind = g["kind"]
        if kind == "uniformHemisphere":
            L.ref_uniform_hemisphere(C.c_float(u[0]), C.c_float(u[1]), ref.ptr(out))
            used = 2
        elif kind == "uniformSphere":
            used = L.ref_uniform_sphere(ref.ptr(u), len(g["uniforms"]), ref.ptr(out))
        elif kind == "GGXHemisphere":
            L.ref_ggx_hemisphere(C.c_float(u[0]), C.c_float(u[1]), C.c_float(g["alpha2"]), ref.ptr(out))
            used = 2
        else:
            L.ref_cosine_sqrt_hemisphere(C.c_float(u[0]), C.c_float(u[1]), ref.ptr(out))
            used = 2
        assert used == g["used"], g
        assert np.array_equal(out, f32(g["v"])), (g, out)
        if kind == "uniformSphere":
            assert np.all(out >= 0.0)   # positive octant only (RandomSampler.h:75-84)


def test_distance_random(ref, golden):
    L = ref.lib()
    for g in golden["distance_random"]:
        w = f32(g["w"])
        pdf = C.c_float()
        idx = L.ref_distance_random(len(w), ref.ptr(w), C.c_float(g["u"]), C.byref(pdf))
        assert idx == g["idx"], g
        assert np.float32(pdf.value) == np.float32(g["pdf"]), g


def test_get_ior(pkg, ref, golden, ior_tables):
    L = ref.lib()
    for g in golden["ior"]:
        arr, keep = pkg.ior_tables(ior_tables[g["material"]])
        n, k = C.c_float(), C.c_float()
        rc = L.ref_get_ior(len(ior_tables[g["material"]]), arr, C.c_float(g["T"]), C.c_float(g["wl"]), C.byref(n), C.byref(k))
        if not g["ok"]:
            assert rc != 0, g
            continue
        assert rc == 0, g
        assert np.float32(n.value) == np.float32(g["n"]) and np.float32(k.value) == np.float32(g["k"]), (g, n.value, k.value)


def test_texture(ref, golden):
    L = ref.lib()
    for tex in golden["texture"]:
        img = np.asarray(tex["img"], np.uint8)
        for q in tex["queries"]:
            out = np.zeros(4, np.float32)
            L.ref_texture_rgba8(tex["w"], tex["h"], ref.ptr(img, C.c_uint8), C.c_float(q["u"]), C.c_float(q["v"]), ref.ptr(out))
            assert np.array_equal(out, f32(q["out"])), (tex["w"], tex["h"], q, out)


def _rgbe_decode(rgbe):
    """stb_image's RGBE conversion: mantissa * 2^(e - 136), zero when e == 0."""
    a = np.asarray(rgbe, np.float64).reshape(-1, 4)
    f = np.where(a[:, 3] != 0, np.ldexp(1.0, (a[:, 3] - 136).astype(int)), 0.0)
    return (a[:, :3] * f[:, None]).astype(np.float32)


def test_hdr_background(ref, golden):
    L = ref.lib()
    rot = f32([[1, 0, 0], [0, 0, 1], [0, -1, 0]]).ravel()
    for h in golden["hdr"]:
        rgb = _rgbe_decode(h["rgbe"])
        for q in h["queries"]:
            out = np.zeros(3, np.float32)
            L.ref_hdr_radiance(h["w"], h["h"], ref.ptr(rgb), ref.ptr(rot), C.c_float(np.float32(np.pi)), ref.ptr(f32(q["dir"])), ref.ptr(out))
            exp = f32(q["out"])
            assert np.allclose(out, exp, rtol=3e-7, atol=0), (q, out, exp)


def test_camera(ref, golden):
    L = ref.lib()
    for g in golden["camera"]:
        if "load_properties" in g:
            # PinholeCamera::loadProperties on cameras/cameraProps.json (the reference's own test: focus 450)
            assert g["focus"] == 450.0
            hh = np.float32(450.0) * np.float32(np.tan(np.float32(0.5) * np.float32(60.0) * np.float32(np.pi) / np.float32(180.0)))
            assert abs(g["half_h"] - hh) <= 1e-4 * hh
            # film normalisation: hit() sums per pixel, stop() divides by the count (PinholeCamera.cpp:70-108)
            hits = np.asarray(g["hits"], np.float32).reshape(-1, 4)
            film = np.zeros((6, 4), np.float32)
            for hrow in hits:
                film[int(hrow[0]), :3] += hrow[1:]
                film[int(hrow[0]), 3] += 1.0
            for px in range(6):
                sc = np.float32(1.0) if film[px, 3] < 0.5 else np.float32(1.0) / film[px, 3]
                film[px, :3] = sc * film[px, :3]
                film[px, 3] = 1.0
            assert np.array_equal(film, f32(g["film"]))
            continue
        d = ref.CameraDesc()
        d.width, d.height = g["w"], g["h"]
        for i in range(3):
            d.position[i] = g["pos"][i]
            d.rotation[i * 3 + 0], d.rotation[i * 3 + 1], d.rotation[i * 3 + 2] = g["left"][i], g["up"][i], g["fwd"][i]
        hw, hh, fo = C.c_float(), C.c_float(), C.c_float()
        L.ref_compute_camera(g["w"], g["h"], C.c_float(g["fov"]), C.byref(hw), C.byref(hh), C.byref(fo))
        assert np.float32(hw.value) == np.float32(g["half_w"]) and np.float32(hh.value) == np.float32(g["half_h"])
        assert np.float32(fo.value) == np.float32(g["focus"])
        d.half_w, d.half_h, d.focus = hw.value, hh.value, fo.value
        o, r = np.zeros(3, np.float32), np.zeros(3, np.float32)
        L.ref_camera_ray(C.byref(d), g["x"], g["y"], C.c_float(g["uniforms"][0]), C.c_float(g["uniforms"][1]), ref.ptr(o), ref.ptr(r))
        assert np.array_equal(o, f32(g["O"])) and np.array_equal(r, f32(g["R"])), (g, o, r)
        assert g["id"] == g["y"] * g["w"] + g["x"]


def test_light_point(ref, golden):
    L = ref.lib()
    for g in golden["light_point"]:
        inten = f32(g["intensity"])
        power = np.float32(np.float32(0.2126) * inten[0] + np.float32(0.7152) * inten[1]) + np.float32(0.0722) * inten[2]
        assert np.float32(power) == np.float32(g["power"])
        out = np.zeros(3, np.float32)
        used = L.ref_uniform_sphere(ref.ptr(f32(g["uniforms"] + [0.5] * 4)), len(g["uniforms"]), ref.ptr(out))
        assert used == g["used"]
        assert np.array_equal(out, f32(g["R"]))
        assert np.array_equal(f32(g["O"]), f32(g["pos"])) and np.array_equal(f32(g["rad"]), inten)
        assert np.float32(g["pdf"]) == np.float32(1.0) / (np.float32(4.0) * np.float32(np.pi))


def test_gamma_correction(ref, golden):
    L = ref.lib()
    for g in golden["gamma"]:
        img = f32(g["in"]).copy()
        L.ref_gamma_correction(ref.ptr(img), g["w"], g["h"], C.c_float(1.0), C.c_float(g["gamma"]))
        assert np.array_equal(img, f32(g["out"])), g
