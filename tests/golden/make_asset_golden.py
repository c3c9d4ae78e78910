#!/usr/bin/env python3
"""Regenerates the asset-loader fixtures in tests/golden/ from the reference
pathtracer's own loader (run in a container that has the reference checkout
and oracle/_ref/asset_harness built by oracle/build_ref.sh).

  assets/objects, scenes, materials, lights, cameras
      a subset of the reference's own sample assets (the files its
      AssetGeometry / PbrMaterial / PinholeCamera tests load), copied as data
  assets/synthetic
      glTF / GLB / HDR inputs written here to reach the loader's corners:
      TRS and matrix nodes, shared meshes, every index type, tangents, uv
      channel quirks, KHR_texture_transform, 8- and 16-bit / palette / grey
      PNGs, KHR_lights_punctual, reference links to the same file twice,
      material priority, error cases, flat and run-length .hdr files
  asset_golden.json
      what the reference's AssetGeometry loaded from each (counts, matrices,
      parameters and checksums of every buffer), and stbi_loadf's decode of
      each .hdr file

Usage: python3 tests/golden/make_asset_golden.py [REF]   (REF default /root/reference)
"""
import base64
import io
import json
import os
import shutil
import struct
import subprocess
import sys
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
ASSETS = os.path.join(HERE, "assets")
SYN = os.path.join(ASSETS, "synthetic")
HARNESS = os.path.join(ROOT, "oracle", "_ref", "asset_harness")

COPY = {
    "objects": ["cone_gold.gltf", "cube_aluminium.gltf", "cube_gold.gltf", "cube_pbr_gold.gltf", "cubes_pbr.gltf",
                "icospheres_gold_iron_aluminium.gltf", "icospheres_gold_iron_missing.gltf", "plane_pointlight.gltf",
                "sphere_gold.gltf", "torus_gold.gltf", "torus_aluminium.gltf", "torus_pbr.gltf",
                "data/cone.bin", "data/cube.bin", "data/cube_pbr_gold.bin", "data/cubes_pbr.bin", "data/icospheres.bin",
                "data/plane_pointlight.bin", "data/torus.bin", "textures/Cubes_PBR.png"],
    "scenes": ["multiple_instances.gltf", "multiple_objects.gltf"],
    "materials": ["gold.gltf", "aluminium.gltf", "iron.gltf", "data/gold_ior.gltf", "data/aluminium_ior.gltf",
                  "data/iron_ior.gltf"],
    "lights": ["pointLight.gltf"],
    "cameras": ["cameraProps.json"],
}
# loaded from the reference checkout only (large); the test skips them without it
REF_ONLY = ["objects/monkey_gold.gltf", "objects/teapot_gold.gltf", "objects/teapot_iron.gltf",
            "objects/truck_example.gltf", "objects/vehicle_example.gltf", "scenes/vehicle_pointlights.gltf"]


# ------------------------------------------------------------------ small PNG / HDR writers
def png(w, h, ctype, depth, rows, palette=None, trns=None):
    def chunk(t, b):
        return struct.pack(">I", len(b)) + t + b + struct.pack(">I", zlib.crc32(t + b) & 0xFFFFFFFF)
    raw = b"".join(bytes([f]) + r for f, r in rows)
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if palette is not None:
        out += chunk(b"PLTE", palette)
    if trns is not None:
        out += chunk(b"tRNS", trns)
    return out + chunk(b"IDAT", zlib.compress(raw, 9)) + chunk(b"IEND", b"")


def filtered_rows(pix, bpp, filters):
    """pix: (h, stride) uint8; returns PNG rows with the given filter types."""
    rows, prev = [], np.zeros(pix.shape[1], np.int32)
    for y, ft in zip(range(pix.shape[0]), filters):
        cur = pix[y].astype(np.int32)
        enc = np.zeros_like(cur)
        for i in range(len(cur)):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            p = [0, a, b, (a + b) >> 1, None][ft]
            if ft == 4:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            enc[i] = (cur[i] - p) & 255
        rows.append((ft, enc.astype(np.uint8).tobytes()))
        prev = cur
    return rows


def hdr_file(path, rgbe, rle, header=b"#?RADIANCE\n"):
    h, w, _ = rgbe.shape
    out = [header, b"# made for the asset loader tests\n", b"FORMAT=32-bit_rle_rgbe\n\n", f"-Y {h} +X {w}\n".encode()]
    for y in range(h):
        if not rle:
            out.append(rgbe[y].tobytes())
            continue
        out.append(bytes([2, 2, w >> 8, w & 255]))
        for k in range(4):
            v = rgbe[y, :, k]
            x = 0
            while x < w:
                run = 1
                while x + run < w and run < 127 and v[x + run] == v[x]:
                    run += 1
                if run >= 3:
                    out.append(bytes([128 + run, int(v[x])]))
                    x += run
                else:
                    n = 1
                    while x + n < w and n < 128 and not (x + n + 2 < w and v[x + n] == v[x + n + 1] == v[x + n + 2]):
                        n += 1
                    out.append(bytes([n]) + v[x:x + n].tobytes())
                    x += n
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


# ------------------------------------------------------------------ glTF builder
class Gltf:
    def __init__(self, title):
        self.bin = bytearray()
        self.j = {"asset": {"version": "2.0", "extensions": {"OpenMaterial_asset_info": {
            "title": title, "creator": "asset loader tests", "asset_type": "geometry"}}},
            "scene": 0, "scenes": [{"nodes": []}], "nodes": [], "meshes": [], "materials": [],
            "accessors": [], "bufferViews": [], "buffers": []}

    def view(self, data: bytes):
        while len(self.bin) % 4:
            self.bin.append(0)
        self.j["bufferViews"].append({"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)})
        self.bin += data
        return len(self.j["bufferViews"]) - 1

    def accessor(self, arr, ctype, typ, extra_offset=0):
        arr = np.ascontiguousarray(arr)
        data = b"\0" * extra_offset + arr.tobytes()
        count = arr.shape[0]
        a = {"bufferView": self.view(data), "componentType": ctype, "count": int(count), "type": typ}
        if extra_offset:
            a["byteOffset"] = extra_offset
        self.j["accessors"].append(a)
        return len(self.j["accessors"]) - 1

    def save_gltf(self, path, data_uri=False, bin_name=None):
        j = json.loads(json.dumps(self.j))
        if data_uri:
            j["buffers"] = [{"byteLength": len(self.bin),
                             "uri": "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()}]
        else:
            bin_name = bin_name or os.path.basename(path).rsplit(".", 1)[0] + ".bin"
            with open(os.path.join(os.path.dirname(path), bin_name), "wb") as fh:
                fh.write(bytes(self.bin))
            j["buffers"] = [{"byteLength": len(self.bin), "uri": bin_name}]
        with open(path, "w") as fh:
            json.dump(j, fh, indent=1)

    def save_glb(self, path):
        j = json.loads(json.dumps(self.j))
        j["buffers"] = [{"byteLength": len(self.bin)}]
        js = json.dumps(j).encode()
        js += b" " * (-len(js) % 4)
        bn = bytes(self.bin) + b"\0" * (-len(self.bin) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bn), 0x004E4942) + bn
        with open(path, "wb") as fh:
            fh.write(struct.pack("<4sII", b"glTF", 2, 12 + len(body)) + body)


F32, U8, I16, U16, U32 = 5126, 5121, 5122, 5123, 5125


def quad(g, rng, material=None, idx_type=U16, tangents=False, uv=(0,), extra_attrs=None):
    v = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32) + rng.normal(0, 0.05, (4, 3)).astype(np.float32)
    n = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    idx = np.array([0, 2, 1, 0, 3, 2], {U8: np.uint8, U16: np.uint16, U32: np.uint32, I16: np.int16}[idx_type])
    attrs = {"POSITION": g.accessor(v, F32, "VEC3"), "NORMAL": g.accessor(n, F32, "VEC3")}
    if tangents:
        t = np.concatenate([rng.normal(0, 1, (4, 3)), np.sign(rng.normal(0, 1, (4, 1)))], 1).astype(np.float32)
        attrs["TANGENT"] = g.accessor(t, F32, "VEC4")
    for ch in uv:
        attrs[f"TEXCOORD_{ch}"] = g.accessor(rng.uniform(-0.5, 1.5, (4, 2)).astype(np.float32), F32, "VEC2")
    for name in extra_attrs or ():
        attrs[name] = g.accessor(rng.uniform(0, 1, (4, 2)).astype(np.float32), F32, "VEC2")
    p = {"attributes": attrs, "indices": g.accessor(idx, idx_type, "SCALAR", extra_offset=4)}
    if material is not None:
        p["material"] = material
    return p


def quat(rng):
    q = rng.normal(0, 1, 4)
    return [float(x) for x in q / np.linalg.norm(q)]


def write_synthetic():
    os.makedirs(SYN, exist_ok=True)
    rng = np.random.default_rng(20240611)

    # images: RGB8 (external, all five filter types), grey+alpha (data URI),
    # 16-bit RGBA (external), palette with transparency (external), a missing file
    pix = rng.integers(0, 256, (8, 16 * 3), dtype=np.uint8)
    with open(os.path.join(SYN, "tex_rgb.png"), "wb") as fh:
        fh.write(png(16, 8, 2, 8, filtered_rows(pix, 3, [0, 1, 2, 3, 4, 1, 4, 2])))
    ga = rng.integers(0, 256, (8, 8 * 2), dtype=np.uint8)
    ga_png = png(8, 8, 4, 8, filtered_rows(ga, 2, [4] * 8))
    p16 = rng.integers(0, 256, (4, 4 * 8), dtype=np.uint8)
    with open(os.path.join(SYN, "tex_rgba16.png"), "wb") as fh:
        fh.write(png(4, 4, 6, 16, filtered_rows(p16, 8, [0, 1, 3, 4])))
    pal = rng.integers(0, 256, 3 * 6, dtype=np.uint8).tobytes()
    pidx = rng.integers(0, 6, (4, 8), dtype=np.uint8)
    with open(os.path.join(SYN, "tex_palette.png"), "wb") as fh:
        fh.write(png(8, 4, 3, 8, [(0, r.tobytes()) for r in pidx], palette=pal, trns=bytes([255, 0, 128])))

    # --- trs_textures.gltf (data-URI buffer)
    g = Gltf("synthetic TRS + textures")
    g.j["images"] = [{"uri": "tex_rgb.png"}, {"uri": "data:image/png;base64," + base64.b64encode(ga_png).decode()},
                     {"uri": "tex_rgba16.png"}, {"uri": "tex_palette.png"}, {"uri": "missing_texture.png"}]
    g.j["textures"] = [{"source": 0}, {"source": 1}, {"source": 2}, {"source": 3}]
    g.j["materials"] = [
        {"name": "painted", "doubleSided": True, "alphaMode": "MASK", "alphaCutoff": 1.7,
         "pbrMetallicRoughness": {"baseColorFactor": [1.5, 0.5, 0.25, 1.0], "metallicFactor": 1.5, "roughnessFactor": -0.2,
                                  "baseColorTexture": {"index": 0},
                                  "metallicRoughnessTexture": {"index": 2, "texCoord": 1, "extensions": {"KHR_texture_transform": {
                                      "offset": [0.25, -0.5], "scale": [2.0, 0.5], "rotation": 0.3}}}},
         "normalTexture": {"index": 1, "texCoord": 1, "scale": 0.7},
         "occlusionTexture": {"index": 3, "strength": 0.3},
         "emissiveFactor": [2.0, 0.0, 0.5], "emissiveTexture": {"index": 0, "extensions": {"KHR_texture_transform": {"scale": [3, 3]}}}},
        {"name": "nothing known"},
        {"name": "defaults", "pbrMetallicRoughness": {}},
    ]
    g.j["meshes"] = [
        {"name": "quad", "primitives": [quad(g, rng, 0, U16, tangents=True, uv=(0, 1)), quad(g, rng, None, U32, uv=())]},
        {"name": "tri", "primitives": [quad(g, rng, 1, U8, uv=(0,))]},
        {"name": "channels", "primitives": [quad(g, rng, 2, U16, uv=(0,), extra_attrs=("TEXCOORD_10", "TEXCOORD_8", "_TEXCOORD_3"))]},
    ]
    M = np.eye(4)
    M[:3, :3] = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]]) * np.array([1.5, 0.5, 2.0])
    M[:3, 3] = [0.3, -1.2, 2.5]
    g.j["nodes"] = [
        {"name": "A", "mesh": 0, "children": [1, 2], "rotation": quat(rng), "translation": [1.0, 2.0, -0.5], "scale": [1.0, 2.0, 0.5],
         "extras": {"category": "car", "id": "A56A4547-7648-4D9F-8B04-710506B7043F"}},
        {"name": "B", "mesh": 1, "children": [3], "matrix": [float(x) for x in M.T.reshape(-1)], "extras": {"id": "xyz"}},
        {"name": "C", "mesh": 0, "translation": [0.0, 0.0, 3.0]},
        {"name": "D", "mesh": 2, "scale": [0.25, 4.0, 1.0]},
        {"name": "E", "mesh": 2, "rotation": quat(rng)},
    ]
    g.j["scenes"][0]["nodes"] = [0, 4]
    g.save_gltf(os.path.join(SYN, "trs_textures.gltf"), data_uri=True)

    # --- lights.gltf (external .bin): a rig of KHR_lights_punctual lights over a floor
    g = Gltf("synthetic lights")
    g.j["extensionsUsed"] = ["KHR_lights_punctual"]
    g.j["extensions"] = {"KHR_lights_punctual": {"lights": [
        {"type": "point", "color": [1.0, 0.5, 0.25], "intensity": 20.0, "range": 10.0},
        {"type": "spot", "spot": {}},
        {"type": "directional", "color": [0.1, 0.2, 0.3]}]}}
    g.j["materials"] = [{"name": "floor", "pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.7, 0.6, 1.0], "metallicFactor": 0.0,
                                                                   "roughnessFactor": 0.6}}]
    g.j["meshes"] = [{"name": "floor", "primitives": [quad(g, rng, 0, U16, uv=())]}]
    g.j["nodes"] = [
        {"name": "rig", "children": [1, 2, 3, 4], "translation": [0.5, 3.0, -1.0], "rotation": quat(rng)},
        {"name": "L0", "translation": [1.0, 0.5, 0.0], "extensions": {"KHR_lights_punctual": {"light": 0}}},
        {"name": "L1", "translation": [-1.0, 0.25, 0.5], "extensions": {"KHR_lights_punctual": {"light": 1}}},
        {"name": "L2", "extensions": {"KHR_lights_punctual": {"light": 7}}},
        {"name": "L3", "translation": [0.0, 1.0, 1.0], "extensions": {"KHR_lights_punctual": {"light": 2}}},
        {"name": "floor", "mesh": 0, "scale": [4.0, 1.0, 4.0]},
    ]
    g.j["scenes"][0]["nodes"] = [0, 5]
    g.save_gltf(os.path.join(SYN, "lights.gltf"))

    # --- binary.glb: BIN chunk, image from a bufferView, OM + PBR material (priority)
    g = Gltf("synthetic GLB")
    img_view = g.view(png(4, 2, 2, 8, filtered_rows(rng.integers(0, 256, (2, 12), dtype=np.uint8), 3, [1, 2])))
    g.j["images"] = [{"bufferView": img_view, "mimeType": "image/png"}]
    g.j["textures"] = [{"source": 0}]
    g.j["materials"] = [{"name": "gold or paint",
                         "pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.1, 0.1, 1.0], "roughnessFactor": 0.3,
                                                  "baseColorTexture": {"index": 0}},
                         "extensions": {"OpenMaterial_reference_link": {"title": "gold", "uri": "../materials/gold.gltf"}}}]
    g.j["meshes"] = [{"name": "box", "primitives": [quad(g, rng, 0, U32, uv=(0,))]}]
    g.j["nodes"] = [{"name": "glb root", "mesh": 0, "translation": [0.0, -1.0, 0.0]}]
    g.j["scenes"][0]["nodes"] = [0]
    g.save_glb(os.path.join(SYN, "binary.glb"))

    # --- child.gltf: one object plus a nested reference into ../objects
    g = Gltf("synthetic child")
    g.j["asset"]["extensions"]["OpenMaterial_asset_info"]["category"] = "Traffic-Sign"
    g.j["materials"] = [{"name": "child paint", "pbrMetallicRoughness": {"metallicFactor": 0.2}}]
    g.j["meshes"] = [{"name": "plate", "primitives": [quad(g, rng, 0, U16, uv=(0,))]}]
    g.j["nodes"] = [
        {"name": "child root", "children": [1, 2], "translation": [0.0, 0.5, 0.0]},
        {"name": "plate", "mesh": 0, "rotation": quat(rng)},
        {"name": "cube link", "translation": [0.0, 2.0, 0.0], "scale": [0.5, 0.5, 0.5],
         "extensions": {"OpenMaterial_reference_link": {"title": "cube", "uri": "../objects/cube_gold.gltf"}}},
        {"name": "second root", "mesh": 0, "translation": [3.0, 0.0, 0.0]},
    ]
    g.j["scenes"][0]["nodes"] = [0, 3]
    g.save_gltf(os.path.join(SYN, "child.gltf"))

    # --- refs.gltf: a scene referencing child.gltf twice and trs_textures.gltf once
    g = Gltf("synthetic references")
    g.j["asset"]["extensions"]["OpenMaterial_asset_info"]["asset_type"] = "scene"
    del g.j["meshes"], g.j["materials"], g.j["accessors"], g.j["bufferViews"]
    g.j["nodes"] = [
        {"name": "R1", "translation": [-4.0, 0.0, 0.0], "extensions": {"OpenMaterial_reference_link": {"uri": "child.gltf"}}},
        {"name": "R2", "translation": [4.0, 0.0, 0.0], "rotation": quat(rng),
         "extensions": {"OpenMaterial_reference_link": {"uri": "child.gltf"}}, "extras": {"category": "bus"}},
        {"name": "R3", "scale": [0.5, 0.5, 0.5], "extensions": {"OpenMaterial_reference_link": {"uri": "trs_textures.gltf"}}},
    ]
    g.j["scenes"][0]["nodes"] = [0, 1, 2]
    j = dict(g.j)
    j.pop("buffers", None)
    with open(os.path.join(SYN, "refs.gltf"), "w") as fh:
        json.dump(j, fh, indent=1)

    # --- OpenMATERIAL material files with a wrong type / a negative temperature
    gold = json.load(open(os.path.join(ASSETS, "materials", "gold.gltf")))
    bad = json.loads(json.dumps(gold))
    bad["asset"]["extensions"]["OpenMaterial_asset_info"]["asset_type"] = "geometry"
    json.dump(bad, open(os.path.join(SYN, "material_wrong_type.gltf"), "w"), indent=1)
    neg = json.loads(json.dumps(gold))
    neg["materials"][0]["extensions"]["OpenMaterial_material_parameters"]["user_preferences"]["temperature"] = -1.0
    neg["materials"][0]["extensions"]["OpenMaterial_material_parameters"]["physical_properties"]["refractive_index_uri"] = "../materials/data/gold_ior.gltf"
    json.dump(neg, open(os.path.join(SYN, "material_negative_temperature.gltf"), "w"), indent=1)
    warm = json.loads(json.dumps(gold))
    warm["materials"][0]["extensions"]["OpenMaterial_material_parameters"]["user_preferences"]["temperature"] = 1234.5
    warm["materials"][0]["extensions"]["OpenMaterial_material_parameters"]["physical_properties"]["refractive_index_uri"] = "../materials/data/gold_ior.gltf"
    json.dump(warm, open(os.path.join(SYN, "material_warm_gold.gltf"), "w"), indent=1)

    # --- error cases (each one throws inside the reference's loader)
    def one_quad(name, mutate, material=None):
        g = Gltf(name)
        g.j["materials"] = [material or {"name": "m", "pbrMetallicRoughness": {}}]
        g.j["meshes"] = [{"name": "q", "primitives": [quad(g, rng, 0, U16, uv=(0,))]}]
        g.j["nodes"] = [{"name": "n", "mesh": 0}]
        g.j["scenes"][0]["nodes"] = [0]
        mutate(g.j)
        g.save_gltf(os.path.join(SYN, name + ".gltf"), data_uri=True)

    one_quad("err_mode", lambda j: j["meshes"][0]["primitives"][0].__setitem__("mode", 1))
    one_quad("err_position_vec2", lambda j: j["accessors"][j["meshes"][0]["primitives"][0]["attributes"]["POSITION"]].__setitem__("type", "VEC2"))
    one_quad("err_normal_type", lambda j: j["accessors"][j["meshes"][0]["primitives"][0]["attributes"]["NORMAL"]].__setitem__("componentType", 5123))
    one_quad("err_not_root", lambda j: (j["nodes"].append({"name": "parent", "children": [0]}), j["scenes"][0].__setitem__("nodes", [1, 0])))
    one_quad("err_alpha_mode", lambda j: None, {"name": "m", "alphaMode": "SOMETIMES", "pbrMetallicRoughness": {}})
    one_quad("err_texcoord_channel", lambda j: j["meshes"][0]["primitives"][0]["attributes"].__setitem__("TEXCOORD_x", 0))
    one_quad("err_default_scene", lambda j: j.__setitem__("scene", 3))
    one_quad("err_om_wrong_type", lambda j: None,
             {"name": "m", "extensions": {"OpenMaterial_reference_link": {"uri": "material_wrong_type.gltf"}}})
    one_quad("err_om_negative_temperature", lambda j: None,
             {"name": "m", "extensions": {"OpenMaterial_reference_link": {"uri": "material_negative_temperature.gltf"}}})
    one_quad("om_warm_gold", lambda j: None,
             {"name": "m", "extensions": {"OpenMaterial_reference_link": {"uri": "material_warm_gold.gltf"}}})
    one_quad("err_asset_type", lambda j: j["asset"]["extensions"]["OpenMaterial_asset_info"].__setitem__("asset_type", "spaceship"))
    one_quad("err_asset_uuid", lambda j: j["asset"]["extensions"]["OpenMaterial_asset_info"].__setitem__("id", "12345"))
    one_quad("err_light_without_type", lambda j: j.__setitem__("extensions", {"KHR_lights_punctual": {"lights": [{"color": [1, 1, 1]}]}}))
    one_quad("err_spot_without_cone", lambda j: j.__setitem__("extensions", {"KHR_lights_punctual": {"lights": [{"type": "spot"}]}}))
    shutil.copyfile(os.path.join(SYN, "err_mode.gltf"), os.path.join(SYN, "err_extension.gltf2"))

    # --- Radiance HDR files: flat (narrow), run-length (wide), RGBE header, zero exponents
    for name, (w, h, rle, header) in {"sky_flat.hdr": (7, 5, False, b"#?RADIANCE\n"), "sky_rle.hdr": (40, 6, True, b"#?RGBE\n"),
                                      "sky_rle_wide.hdr": (300, 3, True, b"#?RADIANCE\n")}.items():
        rgbe = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        rgbe[..., 3] = rng.integers(118, 140, (h, w))
        rgbe[0, :3, 3] = 0
        if rle:
            rgbe[1, 5:30] = rgbe[1, 5]          # runs
        hdr_file(os.path.join(SYN, name), rgbe, rle, header)
    with open(os.path.join(SYN, "ldr_rgb.hdr"), "wb") as fh:        # not Radiance: stb decodes it as 8-bit
        fh.write(png(5, 3, 2, 8, filtered_rows(rng.integers(0, 256, (3, 15), dtype=np.uint8), 3, [0, 1, 2])))
    with open(os.path.join(SYN, "ldr_rgba.hdr"), "wb") as fh:
        fh.write(png(3, 2, 6, 8, filtered_rows(rng.integers(0, 256, (2, 12), dtype=np.uint8), 4, [0, 4])))
    with open(os.path.join(SYN, "garbage.hdr"), "wb") as fh:
        fh.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 2\n" + bytes(5))


def copy_reference():
    for sub, files in COPY.items():
        for f in files:
            dst = os.path.join(ASSETS, sub, f)
            os.makedirs(os.path.dirname(dst), exist_ok=True)
            shutil.copyfile(os.path.join(REF, sub, f), dst)


def run_harness(cwd, files, lights, priority):
    out = []
    for f in files:
        with tempfile.NamedTemporaryFile(suffix=".json", delete=False) as tf:
            path = tf.name
        r = subprocess.run([HARNESS, path, str(int(lights)), priority, f], cwd=cwd, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
        if r.returncode != 0:
            raise SystemExit(f"asset_harness failed on {f} (rc {r.returncode})")
        out += json.load(open(path))
        os.unlink(path)
    return out


def main():
    if not os.path.exists(HARNESS):
        subprocess.run(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")], check=True, env=dict(os.environ, REF=REF))
    copy_reference()
    write_synthetic()
    fixtures = [f"{sub}/{f}" for sub, fs in COPY.items() for f in fs if f.endswith((".gltf", ".glb")) and not f.startswith("data/")
                and sub != "materials"]
    synth = sorted(f"synthetic/{f}" for f in os.listdir(SYN) if f.endswith((".gltf", ".glb", ".gltf2"))
                   and not f.startswith("material_") and f != "child.gltf")
    entries = []
    for lights, priority in ((0, "om"), (1, "om"), (1, "pbr")):
        for e in run_harness(ASSETS, fixtures + synth, lights, priority):
            e["root"] = "fixtures"
            entries.append(e)
    for e in run_harness(REF, REF_ONLY, 1, "om"):
        e["root"] = "reference"
        entries.append(e)
    with tempfile.NamedTemporaryFile(suffix=".json", delete=False) as tf:
        path = tf.name
    hdrs = sorted(f"synthetic/{f}" for f in os.listdir(SYN) if f.endswith(".hdr"))
    subprocess.run([HARNESS, path, "hdr"] + hdrs, cwd=ASSETS, check=True, stdout=subprocess.DEVNULL)
    hdr = json.load(open(path))
    os.unlink(path)
    with open(os.path.join(HERE, "asset_golden.json"), "w") as fh:
        json.dump({"assets": entries, "hdr": hdr}, fh, indent=None, separators=(",", ":"))
    print(f"{len(entries)} asset entries ({sum(1 for e in entries if e['error'])} errors), {len(hdr)} hdr files")


if __name__ == "__main__":
    main()
