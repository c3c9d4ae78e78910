"""Synthetic workloads for the render core (seeded, deterministic).

A SceneDesc holds plain numpy arrays and applies itself to any binding that
exposes the RenderScene entry points (the product's Scene, or the CPU
oracle's), so the same inputs reach both.

  plain_bdpt_icosphere   icosphere subdiv 4 (5,120 tris, r = 1), OpenMATERIAL
                         metal with per-RGB n ~ U[0.1,1.5], k ~ U[2,8] (seed 0),
                         point light (3,3,3) intensity 50; 256x256, 40 deg,
                         16 spp, depth 5
  hdri_torus_knot        (2,3) torus knot 1000 x 50 x 2 = 100,000 tris,
                         glTF metallic-roughness with 1024^2 roughness and
                         metallic maps from value noise in [0.05, 1],
                         2048x1024 HDRI: log-normal sky (sigma = 1) + 0.5 deg
                         sun at 1e4 x the mean sky; 512x512, 64 spp, depth 8
  large_instanced        200 instanced, randomly displaced subdivided
                         ellipsoids / boxes of ~10k tris each (2M tris),
                         8 materials (complex-IOR metals + PBR), 4 point
                         lights + the HDRI; 1920x1080, 16 spp, depth 8
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# z-up -> y-up background rotation and pi radiance scale (main.cpp:400-405)
HDR_ROTATION = np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]], np.float32)
HDR_SCALE = np.float32(math.pi)


@dataclass
class SceneDesc:
    meshes: list = field(default_factory=list)      # dict(vertices, faces, mat_id, normals, tangents, uvs)
    instances: list = field(default_factory=list)   # (4x4, mesh_id)
    materials: list = field(default_factory=list)   # ("om_nk", n, k) | ("om", tables, T) | ("pbr", kwargs)
    textures: list = field(default_factory=list)    # ("rgba8", HxWx4 u8) | ("f32", HxWx4 f32)
    lights: list = field(default_factory=list)      # (pos, intensity, range)
    background: tuple | None = None                 # ("const", rgb) | ("hdr", rgb HxWx3, rot, scale)
    fc_subject: str | None = None

    def triangles(self) -> int:
        return sum(len(self.meshes[m]["faces"]) for _, m in self.instances)

    def apply(self, scene, pbr_desc):
        """Drive RenderScene set-up in the reference main's order (main.cpp:314-416)."""
        scene.allocate(len(self.meshes), len(self.instances), len(self.materials), len(self.textures), len(self.lights))
        for i, tex in enumerate(self.textures):
            if tex is None:       # slot allocated, never set
                continue
            kind, img = tex
            if kind == "rgba8":
                scene.set_texture(i, img)
            else:
                scene.set_texture_f32(i, img)
        for i, m in enumerate(self.materials):
            if m[0] == "om_nk":
                scene.set_material_om_nk(i, m[1], m[2])
            elif m[0] == "om":
                scene.set_material_om(i, m[1], m[2])
            else:
                scene.set_material_pbr(i, pbr_desc(**m[1]))
        for i, (pos, inten, rng) in enumerate(self.lights):
            scene.set_light_point(i, pos, inten, rng)
        for i, m in enumerate(self.meshes):
            scene.set_mesh(i, m["vertices"], m["faces"], m["mat_id"], m.get("normals"), m.get("tangents"), m.get("uvs", ()),
                           m.get("tex_n"), m.get("tangent_stride", 4))
        for i, (tm, mesh) in enumerate(self.instances):
            scene.set_instance(i, tm, mesh)
        if self.background is not None:
            if self.background[0] == "const":
                scene.set_background_constant(self.background[1])
            else:
                scene.set_background_hdr(self.background[1], self.background[2], self.background[3])
        if self.fc_subject:
            scene.set_false_color_subject(self.fc_subject)
        return scene.commit()


@dataclass
class View:
    width: int
    height: int
    position: tuple
    forward: tuple
    up: tuple = (0.0, 1.0, 0.0)
    yfov: float = 60.0


# ----------------------------------------------------------------- geometry
def icosphere(subdiv: int, radius: float = 1.0):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
             (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return (np.array(verts) * radius).astype(np.float32), np.array(faces, np.uint32)


def torus_knot(p=2, q=3, segments=1000, tube=50, scale=0.5, tube_radius=0.2):
    """(p,q) torus knot tube with uv; (segments+1) x (tube+1) vertices (seams
    duplicated), 2 * segments * tube triangles."""
    t = np.linspace(0.0, 2.0 * np.pi, segments + 1)

    def curve(tt):
        r = np.cos(q * tt) + 2.0
        return np.stack([r * np.cos(p * tt), r * np.sin(p * tt), -np.sin(q * tt)], -1) * scale

    c = curve(t)
    dc = curve(t + 1e-4) - curve(t - 1e-4)
    T = dc / np.linalg.norm(dc, axis=1, keepdims=True)
    ref = np.array([0.0, 0.0, 1.0])
    N = np.cross(T, ref)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    B = np.cross(T, N)
    a = np.linspace(0.0, 2.0 * np.pi, tube + 1)
    ca, sa = np.cos(a), np.sin(a)
    P = c[:, None, :] + tube_radius * (ca[None, :, None] * N[:, None, :] + sa[None, :, None] * B[:, None, :])
    verts = P.reshape(-1, 3).astype(np.float32)
    uu, vv = np.meshgrid(np.linspace(0, 1, segments + 1), np.linspace(0, 1, tube + 1), indexing="ij")
    uvs = np.stack([uu.ravel() * 8.0, vv.ravel()], -1).astype(np.float32)
    i, j = np.meshgrid(np.arange(segments), np.arange(tube), indexing="ij")
    a0 = i * (tube + 1) + j
    a1 = (i + 1) * (tube + 1) + j
    a2 = (i + 1) * (tube + 1) + j + 1
    a3 = i * (tube + 1) + j + 1
    # counter-clockwise seen from outside: outward geometric normals
    faces = np.concatenate([np.stack([a0, a2, a1], -1).reshape(-1, 3), np.stack([a0, a3, a2], -1).reshape(-1, 3)]).astype(np.uint32)
    return verts, faces, uvs


def grid_sphere(stacks, slices, radii=(1.0, 1.0, 1.0), bump=None):
    th = np.linspace(0.0, np.pi, stacks + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, slices + 1)
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(Pp), np.cos(T), np.sin(T) * np.sin(Pp)], -1)
    r = 1.0 if bump is None else bump(T, Pp)
    P = d * np.asarray(radii)[None, None, :] * (r[..., None] if bump is not None else 1.0)
    verts = P.reshape(-1, 3).astype(np.float32)
    uvs = np.stack([(Pp / (2 * np.pi)).ravel(), (T / np.pi).ravel()], -1).astype(np.float32)
    i, j = np.meshgrid(np.arange(stacks), np.arange(slices), indexing="ij")
    a0 = i * (slices + 1) + j
    a1 = (i + 1) * (slices + 1) + j
    a2 = (i + 1) * (slices + 1) + j + 1
    a3 = i * (slices + 1) + j + 1
    faces = np.concatenate([np.stack([a0, a2, a1], -1).reshape(-1, 3), np.stack([a0, a3, a2], -1).reshape(-1, 3)]).astype(np.uint32)
    return verts, faces, uvs


def grid_box(n, size=(1.0, 1.0, 1.0), bump=None):
    verts, faces, uvs = [], [], []
    base = 0
    g = np.linspace(-1.0, 1.0, n + 1)
    U, V = np.meshgrid(g, g, indexing="ij")
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            P = np.zeros(U.shape + (3,))
            a1, a2 = (axis + 1) % 3, (axis + 2) % 3
            P[..., axis] = sgn
            P[..., a1] = U if sgn > 0 else V
            P[..., a2] = V if sgn > 0 else U
            if bump is not None:
                P = P * bump(P)[..., None]
            verts.append((P * np.asarray(size)).reshape(-1, 3))
            uvs.append(np.stack([(U.ravel() + 1) / 2, (V.ravel() + 1) / 2], -1))
            i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
            b0 = base + i * (n + 1) + j
            b1 = base + (i + 1) * (n + 1) + j
            b2 = base + (i + 1) * (n + 1) + j + 1
            b3 = base + i * (n + 1) + j + 1
            faces.append(np.stack([b0, b1, b2], -1).reshape(-1, 3))
            faces.append(np.stack([b0, b2, b3], -1).reshape(-1, 3))
            base += (n + 1) * (n + 1)
    return (np.concatenate(verts).astype(np.float32), np.concatenate(faces).astype(np.uint32), np.concatenate(uvs).astype(np.float32))


def plane(nx=1, nz=1, size=1.0, y=0.0):
    g1, g2 = np.linspace(-size, size, nx + 1), np.linspace(-size, size, nz + 1)
    X, Z = np.meshgrid(g1, g2, indexing="ij")
    verts = np.stack([X.ravel(), np.full(X.size, y), Z.ravel()], -1).astype(np.float32)
    uvs = np.stack([(X.ravel() + size) / (2 * size), (Z.ravel() + size) / (2 * size)], -1).astype(np.float32)
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    a0 = i * (nz + 1) + j
    a1 = (i + 1) * (nz + 1) + j
    a2 = (i + 1) * (nz + 1) + j + 1
    a3 = i * (nz + 1) + j + 1
    faces = np.concatenate([np.stack([a0, a3, a2], -1).reshape(-1, 3), np.stack([a0, a2, a1], -1).reshape(-1, 3)]).astype(np.uint32)
    return verts, faces, uvs


# ---------------------------------------------------------------- textures
def value_noise(h, w, seed, octaves=4, base=8):
    """Seeded multi-octave value noise in [0,1] (bilinear lattice, smoothstep)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((h, w), np.float64)
    amp, total = 1.0, 0.0
    y = (np.arange(h) + 0.5) / h
    x = (np.arange(w) + 0.5) / w
    for o in range(octaves):
        f = base * (2 ** o)
        lat = rng.random((f + 1, f + 1))
        fy, fx = y * f, x * f
        iy, ix = np.minimum(fy.astype(int), f - 1), np.minimum(fx.astype(int), f - 1)
        ty, tx = fy - iy, fx - ix
        ty, tx = ty * ty * (3 - 2 * ty), tx * tx * (3 - 2 * tx)
        a = lat[iy][:, ix]
        b = lat[iy][:, ix + 1]
        c = lat[iy + 1][:, ix]
        d = lat[iy + 1][:, ix + 1]
        top = a + (b - a) * tx[None, :]
        bot = c + (d - c) * tx[None, :]
        out += amp * (top + (bot - top) * ty[:, None])
        total += amp
        amp *= 0.5
    out /= total
    lo, hi = out.min(), out.max()
    return (out - lo) / max(hi - lo, 1e-12)


def hdri(w=2048, h=1024, seed=2, sigma=1.0, sun_deg=0.5, sun_factor=1e4, sun_dir=(0.35, 0.55, 0.3)):
    """Equirect RGB fp32: log-normal sky radiance (log value noise scaled to
    standard deviation sigma) plus a sun disk of angular diameter sun_deg at
    sun_factor x the mean sky radiance.  Directions follow the background's
    own (z-up) frame of HDRBackground::getRadiance."""
    n = value_noise(h, w, seed, octaves=5, base=4)
    z = (n - n.mean()) / max(n.std(), 1e-12)
    sky = np.exp(sigma * z)
    tint = np.stack([sky * 0.9, sky * 1.0, sky * 1.15], -1)
    mean = tint.mean()
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    yang = (1.0 - v) * np.pi - 0.5 * np.pi          # asin(z)
    xang = u * 2.0 * np.pi
    Z = np.sin(yang)[:, None]
    X = np.cos(yang)[:, None] * np.cos(xang)[None, :]
    Y = np.cos(yang)[:, None] * np.sin(xang)[None, :]
    s = np.asarray(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    cosang = X * s[0] + Y * s[1] + Z * s[2]
    disk = cosang >= math.cos(math.radians(sun_deg * 0.5))
    img = tint.copy()
    img[disk] = mean * sun_factor
    return img.astype(np.float32)


def rgba8_from_channels(r, g, b, a=None):
    a = np.ones_like(r) if a is None else a
    return np.clip(np.round(np.stack([r, g, b, a], -1) * 255.0), 0, 255).astype(np.uint8)


def rng_om(seed):
    rng = np.random.default_rng(seed)
    n = rng.uniform(0.1, 1.5, 3).astype(np.float32)
    k = rng.uniform(2.0, 8.0, 3).astype(np.float32)
    return n, k


def identity():
    return np.eye(4, dtype=np.float32)


def trs(t=(0, 0, 0), axis=(0, 1, 0), angle=0.0, scale=(1, 1, 1)):
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    c, s = math.cos(angle), math.sin(angle)
    x, y, z = a
    R = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(scale)
    M[:3, 3] = t
    return M.astype(np.float32)


# ----------------------------------------------------------------- scenes
def plain_bdpt_icosphere():
    """Plain BDPT, CPU-runnable configuration.  Under the reference estimator
    the OpenMATERIAL mirror's connection gate (OpenMaterial.cpp:105-109) makes
    this image all black; the OM rough-conductor extension renders it."""
    v, f = icosphere(4, 1.0)
    n, k = rng_om(0)
    d = SceneDesc()
    d.materials.append(("om_nk", n, k))
    d.meshes.append({"vertices": v, "faces": f, "mat_id": 0})
    d.instances.append((identity(), 0))
    d.lights.append(((3.0, 3.0, 3.0), (50.0, 50.0, 50.0), -1.0))
    view = View(256, 256, (0.0, 0.0, 4.0), (0.0, 0.0, -1.0), yfov=40.0)
    params = dict(camera_bounces=5, light_bounces=5, max_path_length=5, samples_per_pixel=16, seed=0)
    return d, view, params


def hdri_torus_knot(width=512, height=512, spp=64, depth=8, tex=1024, hdr_w=2048, hdr_h=1024, f32_maps=False, segments=1000, tube=50):
    """HDRI-lit scene.  f32_maps=False stores the maps as RGBA8 the way the
    reference loads images (ImageData.cpp:86-91); True uses fp32 textures."""
    v, f, uv = torus_knot(segments=segments, tube=tube)
    rough = (0.05 + 0.95 * value_noise(tex, tex, 11)).astype(np.float32)
    metal = (0.05 + 0.95 * value_noise(tex, tex, 12)).astype(np.float32)
    ones = np.ones_like(rough)
    d = SceneDesc()
    if f32_maps:
        d.textures.append(("f32", np.stack([ones, rough, metal, ones], -1).astype(np.float32)))
    else:
        d.textures.append(("rgba8", rgba8_from_channels(ones, rough, metal)))
    d.materials.append(("pbr", dict(base_color=(0.85, 0.62, 0.32, 1.0), metallic=1.0, roughness=1.0, textures={1: 0})))
    d.meshes.append({"vertices": v, "faces": f, "mat_id": 0, "uvs": [uv]})
    d.instances.append((identity(), 0))
    d.background = ("hdr", hdri(hdr_w, hdr_h), HDR_ROTATION, HDR_SCALE)
    view = View(width, height, (0.0, 0.4, 5.0), (0.0, -0.08, -1.0), yfov=40.0)
    params = dict(camera_bounces=depth, light_bounces=depth, max_path_length=depth, samples_per_pixel=spp, seed=0)
    return d, view, params


def large_instanced(width=1920, height=1080, spp=16, depth=8, count=200, tris_per=10000, ior_tables=None, hdr_w=2048, hdr_h=1024, seed=3):
    rng = np.random.default_rng(seed)
    d = SceneDesc()
    if ior_tables:
        for name in ("gold", "iron", "aluminium"):
            d.materials.append(("om", ior_tables[name], 300.0))
    else:
        for s in (100, 101, 102):
            d.materials.append(("om_nk",) + rng_om(s))
    d.materials.append(("om_nk",) + rng_om(103))
    d.materials.append(("om_nk",) + rng_om(104))
    tex = value_noise(256, 256, 31)
    d.textures.append(("rgba8", rgba8_from_channels(np.ones_like(tex), 0.1 + 0.9 * tex, 1.0 - tex)))
    d.materials.append(("pbr", dict(base_color=(0.8, 0.2, 0.15, 1.0), metallic=0.0, roughness=0.6)))
    d.materials.append(("pbr", dict(base_color=(0.3, 0.7, 0.9, 1.0), metallic=1.0, roughness=1.0, textures={1: 0})))
    d.materials.append(("pbr", dict(base_color=(0.9, 0.9, 0.85, 1.0), metallic=0.2, roughness=0.3)))
    stacks = int(round(math.sqrt(tris_per / 4)))
    nbox = int(round(math.sqrt(tris_per / 12)))
    for i in range(count):
        freq = rng.uniform(2, 7, 3)
        amp = rng.uniform(0.02, 0.12)
        ph = rng.uniform(0, 2 * np.pi, 3)
        if i % 2 == 0:
            bump = lambda T, P, f=freq, a=amp, p=ph: 1.0 + a * np.sin(f[0] * T + p[0]) * np.cos(f[1] * P + p[1])
            v, f_, uv = grid_sphere(stacks, 2 * stacks, radii=rng.uniform(0.5, 1.2, 3), bump=bump)
        else:
            bump = lambda P, f=freq, a=amp, p=ph: 1.0 + a * np.sin(f[0] * P[..., 0] + p[0]) * np.sin(f[1] * P[..., 1] + p[1]) * np.sin(f[2] * P[..., 2] + p[2])
            v, f_, uv = grid_box(nbox, size=rng.uniform(0.4, 1.0, 3), bump=bump)
        d.meshes.append({"vertices": v, "faces": f_, "mat_id": int(i % 8), "uvs": [uv]})
        gx, gz = i % 20, i // 20
        t = (gx * 2.6 - 24.7 + rng.uniform(-0.4, 0.4), rng.uniform(-0.5, 0.5), -gz * 2.6 - 2.0 + rng.uniform(-0.4, 0.4))
        d.instances.append((trs(t, rng.normal(size=3), rng.uniform(0, 2 * np.pi), np.full(3, rng.uniform(0.7, 1.1))), i))
    for p in ((-15.0, 8.0, -5.0), (0.0, 10.0, -12.0), (15.0, 8.0, -20.0), (0.0, 4.0, 2.0)):
        d.lights.append((p, (400.0, 380.0, 350.0), -1.0))
    d.background = ("hdr", hdri(hdr_w, hdr_h), HDR_ROTATION, HDR_SCALE)
    view = View(width, height, (0.0, 6.0, 12.0), (0.0, -0.35, -1.0), yfov=50.0)
    params = dict(camera_bounces=depth, light_bounces=depth, max_path_length=depth, samples_per_pixel=spp, seed=0)
    return d, view, params


def material_fit(width=256, height=256, spp=32, depth=8, tex=512, hdr_w=2048, hdr_h=1024, seed_target=5000, seed_init=5001,
                 segments=1000, tube=50):
    """Differentiable material fit: the HDRI torus-knot scene at width x
    height with fp32 base-colour (texture 0, RGBA) and metallic-roughness
    (texture 1: G roughness, B metallic) maps of tex^2 texels, plus an
    OpenMATERIAL mirror sphere (material 1) beside the knot, which the HDRI
    makes observable through escape paths.  Parameters come from value noise
    seeded by seed_target (hidden target) and seed_init (initial guess).
    Returns (target SceneDesc, initial SceneDesc, View, params, ids) with
    ids = {"base_color": 0, "metal_rough": 1, "om": 1}."""
    v, f, uv = torus_knot(segments=segments, tube=tube)
    sv, sf = icosphere(3, 0.45)
    sky = hdri(hdr_w, hdr_h)

    def desc(seed):
        rng = np.random.default_rng(seed)
        base = np.stack([0.05 + 0.9 * value_noise(tex, tex, seed * 10 + c) for c in range(3)] + [np.ones((tex, tex))], -1)
        rough = 0.05 + 0.95 * value_noise(tex, tex, seed * 10 + 5)
        metal = 0.05 + 0.95 * value_noise(tex, tex, seed * 10 + 6)
        mr = np.stack([np.ones_like(rough), rough, metal, np.ones_like(rough)], -1)
        d = SceneDesc()
        d.textures.append(("f32", base.astype(np.float32)))
        d.textures.append(("f32", mr.astype(np.float32)))
        d.materials.append(("pbr", dict(base_color=(1.0, 1.0, 1.0, 1.0), metallic=1.0, roughness=1.0, textures={0: 0, 1: 1})))
        n, k = rng_om(int(rng.integers(1 << 30)))
        d.materials.append(("om_nk", n, k))
        d.meshes.append({"vertices": v, "faces": f, "mat_id": 0, "uvs": [uv]})
        d.meshes.append({"vertices": sv, "faces": sf, "mat_id": 1})
        d.instances.append((identity(), 0))
        d.instances.append((trs((1.75, 0.55, 0.6)), 1))
        d.background = ("hdr", sky, HDR_ROTATION, HDR_SCALE)
        return d

    view = View(width, height, (0.3, 0.5, 5.2), (-0.05, -0.1, -1.0), yfov=42.0)
    params = dict(camera_bounces=depth, light_bounces=depth, max_path_length=depth, samples_per_pixel=spp, seed=0)
    return desc(seed_target), desc(seed_init), view, params, {"base_color": 0, "metal_rough": 1, "om": 1}


def lightbox_views(vertices, tm=None):
    """The eight lightbox viewpoints of main.cpp:434-457: for each corner p of
    the scene bounding box with centre C, the camera sits at 2p - C and looks
    at C (left = fwd x (0,1,0), up = left x fwd, not renormalised)."""
    v = np.asarray(vertices, np.float64)
    if tm is not None:
        v = v @ np.asarray(tm, np.float64)[:3, :3].T + np.asarray(tm, np.float64)[:3, 3]
    lo, hi = v.min(0).astype(np.float32), v.max(0).astype(np.float32)
    c = ((lo + hi) * np.float32(0.5)).astype(np.float32)
    views = []
    for i in range(8):
        p = np.array([hi[0] if i & 1 else lo[0], hi[1] if i & 2 else lo[1], hi[2] if i & 4 else lo[2]], np.float32)
        fwd = (c - p).astype(np.float32)
        pos = (np.float32(-2.0) * fwd + c).astype(np.float32)
        views.append((tuple(pos.tolist()), tuple(fwd.tolist())))
    return views


def batched_lightbox(jobs=256, views_per_asset=8, seed=4, tri_range=(1000, 200000), res_range=(128, 512), spp_range=(8, 64),
                     depth_range=(4, 8), hdr_w=512, hdr_h=256):
    """Batched validation: jobs // views_per_asset assets, each one seeded
    displaced mesh (sphere, box or torus knot) with a triangle count
    log-uniform in tri_range and a random material (OpenMATERIAL (n,k) or PBR
    base colour / metallic / roughness), lit by one point light and an HDRI,
    seen from its lightbox viewpoints; every view has its own resolution
    (square, uniform in res_range), spp and depth (uniform).  Returns
    (assets: [SceneDesc], jobs: [(asset index, View, params)])."""
    rng = np.random.default_rng(seed)
    n_assets = max(1, (jobs + views_per_asset - 1) // views_per_asset)
    sky = hdri(hdr_w, hdr_h, seed=seed + 100)
    assets, out = [], []
    for a in range(n_assets):
        tris = int(round(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))))
        kind = a % 3
        freq = rng.uniform(2, 7, 3)
        amp = rng.uniform(0.02, 0.12)
        ph = rng.uniform(0, 2 * np.pi, 3)
        if kind == 0:
            stacks = max(4, int(round(math.sqrt(tris / 4))))
            bump = lambda T, P, f=freq, am=amp, p=ph: 1.0 + am * np.sin(f[0] * T + p[0]) * np.cos(f[1] * P + p[1])
            v, f_, uv = grid_sphere(stacks, 2 * stacks, radii=rng.uniform(0.5, 1.2, 3), bump=bump)
        elif kind == 1:
            n = max(2, int(round(math.sqrt(tris / 12))))
            bump = lambda P, f=freq, am=amp, p=ph: 1.0 + am * np.sin(f[0] * P[..., 0] + p[0]) * np.sin(f[1] * P[..., 1] + p[1]) * np.sin(f[2] * P[..., 2] + p[2])
            v, f_, uv = grid_box(n, size=rng.uniform(0.4, 1.0, 3), bump=bump)
        else:
            tube = max(6, int(round(math.sqrt(tris / 40))))
            segs = max(12, int(round(tris / (2 * tube))))
            v, f_, uv = torus_knot(segments=segs, tube=tube)
        d = SceneDesc()
        if rng.uniform() < 0.5:
            n_, k_ = rng_om(int(rng.integers(1 << 30)))
            d.materials.append(("om_nk", n_, k_))
        else:
            d.materials.append(("pbr", dict(base_color=tuple(rng.uniform(0.05, 0.95, 3).tolist()) + (1.0,),
                                            metallic=float(rng.uniform(0.0, 1.0)), roughness=float(rng.uniform(0.05, 1.0)))))
        d.meshes.append({"vertices": v, "faces": f_, "mat_id": 0, "uvs": [uv]})
        d.instances.append((identity(), 0))
        ext = float(np.abs(v).max())
        d.lights.append(((2.5 * ext, 3.0 * ext, 2.0 * ext), tuple((rng.uniform(20, 60) * ext * np.ones(3)).tolist()), -1.0))
        d.background = ("hdr", sky, HDR_ROTATION, HDR_SCALE)
        assets.append(d)
        for (pos, fwd) in lightbox_views(v)[:views_per_asset]:
            if len(out) == jobs:
                break
            r = int(rng.integers(res_range[0], res_range[1] + 1))
            depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
            spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
            params = dict(camera_bounces=depth, light_bounces=depth, max_path_length=depth, samples_per_pixel=spp,
                          seed=int(len(out)))
            out.append((a, View(r, r, pos, fwd, yfov=60.0), params))
    return assets, out


def mixed_test_scene(hdr=True, lights=True, tex_f32=False, masked=False, missing=True):
    """Small scene exercising every material path: a textured PBR floor, an
    OpenMATERIAL sphere, a PBR cube with a normal map, a mesh with a missing
    material, a transformed second instance, point lights and a background."""
    d = SceneDesc()
    rough = value_noise(64, 64, 5)
    nrm = value_noise(32, 32, 6)
    ones = np.ones_like(rough)
    if tex_f32:
        d.textures.append(("f32", np.stack([ones, 0.05 + 0.9 * rough, 1.0 - rough, ones], -1).astype(np.float32)))
    else:
        d.textures.append(("rgba8", rgba8_from_channels(ones, 0.05 + 0.9 * rough, 1.0 - rough)))
    d.textures.append(("rgba8", rgba8_from_channels(0.5 + 0.3 * (nrm - 0.5), 0.5 - 0.3 * (nrm - 0.5), np.full_like(nrm, 0.9))))
    base = value_noise(32, 32, 7)
    d.textures.append(("rgba8", rgba8_from_channels(base, 1 - base, 0.5 * np.ones_like(base), np.where(base > 0.3, 1.0, 0.0))))
    d.materials.append(("pbr", dict(base_color=(0.7, 0.7, 0.7, 1.0), metallic=1.0, roughness=1.0, textures={1: 0})))
    d.materials.append(("om_nk",) + rng_om(1))
    d.materials.append(("pbr", dict(base_color=(0.9, 0.4, 0.2, 1.0), metallic=0.1, roughness=0.5, textures={2: 1})))
    if masked:
        d.materials.append(("pbr", dict(base_color=(1, 1, 1, 1), metallic=0.0, roughness=0.7, alpha_mode=1, alpha_cutoff=0.5, textures={0: 2})))
    pv, pf, puv = plane(4, 4, 3.0, -1.0)
    d.meshes.append({"vertices": pv, "faces": pf, "mat_id": 0, "uvs": [puv]})
    sv, sf = icosphere(2, 0.8)
    d.meshes.append({"vertices": sv, "faces": sf, "mat_id": 1})
    bv, bf, buv = grid_box(3, (0.5, 0.5, 0.5))
    d.meshes.append({"vertices": bv, "faces": bf, "mat_id": 2, "uvs": [buv]})
    mv, mf = icosphere(1, 0.4)
    # undefined material -> missing (pink) slot; otherwise the OM material
    d.meshes.append({"vertices": mv, "faces": mf, "mat_id": 7 if missing else 1})
    d.instances.append((identity(), 0))
    d.instances.append((trs((-0.9, -0.2, 0.0)), 1))
    d.instances.append((trs((1.1, -0.4, 0.3), (1, 1, 0), 0.6), 2))
    d.instances.append((trs((0.2, 0.6, 1.2), (0, 0, 1), 0.3, (1.5, 0.7, 1.0)), 3))
    d.instances.append((trs((1.4, 0.9, -1.2), (0, 1, 0), 0.9, (0.6, 0.6, 0.6)), 1))
    if masked:
        qv, qf, quv = plane(2, 2, 0.8, 0.0)
        d.meshes.append({"vertices": qv, "faces": qf, "mat_id": 3, "uvs": [quv]})
        d.instances.append((trs((0.0, 0.3, 1.6), (1, 0, 0), 1.3), len(d.meshes) - 1))
    if lights:
        d.lights.append(((2.0, 3.0, 2.5), (40.0, 40.0, 40.0), -1.0))
        d.lights.append(((-2.5, 2.0, 1.0), (10.0, 25.0, 30.0), 12.0))
    if hdr:
        d.background = ("hdr", hdri(64, 32, seed=9, sun_factor=50.0), HDR_ROTATION, HDR_SCALE)
    view = View(64, 48, (0.3, 0.9, 4.5), (-0.05, -0.25, -1.0), yfov=55.0)
    return d, view


# ------------------------------------------------- light-tracer splat checks
def _placed(t, flip=False):
    """Translation by t; flip = rotation by pi about x (a +y-facing plane() then faces -y)."""
    M = np.eye(4, dtype=np.float32)
    if flip:
        M[1, 1] = -1.0
        M[2, 2] = -1.0
    M[:3, 3] = t
    return M


def light_splat_scenes(width=32, height=32):
    """Configurations in which the light-tracer splat image (OMV_EXT_LIGHT_SPLAT)
    has a known expectation, all with inverse-square attenuation and every
    lit point inside the point light's emission octant (+x, +y, +z):

    * "direct": a point light under a diffuse PBR ceiling, the camera looking
      up at it.  Light vertex 1 splats estimate the reference estimator's own
      image of the same scene with B = 1, q = 2 (its only non-zero terms are
      the direct connections).
    * "mirror": the light reaches a diffuse floor only through a near-perfect
      OpenMATERIAL mirror (k = 1e4, F = 1 - O(1e-4)); light vertex 2 splats
      estimate the direct image of the mirror-image light (the "mirror_ref"
      scene: the floor and the light reflected in the mirror plane).

    Returns {name: (desc, view, params)}; params without the extension bit."""
    diffuse = ("pbr", dict(base_color=(0.8, 0.7, 0.6, 1.0), metallic=0.0, roughness=1.0))
    light = ((0.0, 0.0, 0.0), (50.0, 50.0, 50.0), -1.0)
    out = {}
    # direct: ceiling y = 2 (facing -y) over x, z in [0.5, 4.5]
    d = SceneDesc()
    d.materials.append(diffuse)
    pv, pf, puv = plane(4, 4, 2.0, 0.0)
    d.meshes.append({"vertices": pv, "faces": pf, "mat_id": 0, "uvs": [puv]})
    d.instances.append((_placed((2.5, 2.0, 2.5), flip=True), 0))
    d.lights.append(light)
    view = View(width, height, (2.5, 0.5, 2.5), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), 90.0)
    out["direct"] = (d, view, dict(camera_bounces=1, light_bounces=1, max_path_length=2, light_distance_attenuation=2))
    # mirror: mirror y = 2 (facing -y) over [0, 6]^2, floor y = 1 (facing +y) over [2.5, 6]^2
    floor_v, floor_f, floor_uv = plane(4, 4, 1.75, 0.0)
    d = SceneDesc()
    d.materials.append(diffuse)
    d.materials.append(("om_nk", np.ones(3, np.float32), np.full(3, 1e4, np.float32)))
    mv, mf, muv = plane(4, 4, 3.0, 0.0)
    d.meshes.append({"vertices": floor_v, "faces": floor_f, "mat_id": 0, "uvs": [floor_uv]})
    d.meshes.append({"vertices": mv, "faces": mf, "mat_id": 1, "uvs": [muv]})
    d.instances.append((_placed((4.25, 1.0, 4.25)), 0))
    d.instances.append((_placed((3.0, 2.0, 3.0), flip=True), 1))
    d.lights.append(light)
    view = View(width, height, (4.25, 1.8, 4.25), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), 90.0)
    out["mirror"] = (d, view, dict(camera_bounces=1, light_bounces=2, max_path_length=3, light_distance_attenuation=2))
    r = SceneDesc()
    r.materials.append(diffuse)
    r.meshes.append({"vertices": floor_v, "faces": floor_f, "mat_id": 0, "uvs": [floor_uv]})
    r.instances.append((_placed((4.25, 1.0, 4.25)), 0))
    r.lights.append(((0.0, 4.0, 0.0), (50.0, 50.0, 50.0), -1.0))
    out["mirror_ref"] = (r, view, dict(camera_bounces=1, light_bounces=1, max_path_length=2, light_distance_attenuation=2))
    return out
