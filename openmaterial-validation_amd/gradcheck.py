"""Gradient checks of the path-replay adjoint by central finite differences
on the device's own paths.

The adjoint (omv_renderer_backward) differentiates the detached functional:
directions and lobes sampled with the current parameters, bounce weights
re-evaluated with new ones.  omv_renderer_render_detached evaluates that
functional in FP64 on the device (same float paths as the forward), so a
central difference of an image loss over it resolves the parameter
perturbation without float rounding in the loss difference, and compares
with the adjoint on identical paths.

The loss is L = sum over pixels and rgb of (I - T)^2 with I the detached
film; d L / d I = 2 (I - T) is what the adjoint is given.
"""
from __future__ import annotations

import numpy as np


def loss(film: np.ndarray, target: np.ndarray) -> float:
    return float(((film[..., :3].astype(np.float64) - target[..., :3].astype(np.float64)) ** 2).sum())


def d_loss(film: np.ndarray, target: np.ndarray) -> np.ndarray:
    """d L / d film as float32 (H, W, 4), alpha 0."""
    d = np.zeros(film.shape, np.float32)
    d[..., :3] = (2.0 * (film[..., :3].astype(np.float64) - target[..., :3].astype(np.float64))).astype(np.float32)
    return d


def pick_texels(grads: dict, count: int, seed: int, channels, textures: dict, avoid=None, rel_floor=1e-3):
    """count texels (tex, y, x, channel), drawn with a seeded generator among
    those whose gradient magnitude is at least rel_floor of that channel's
    largest; avoid(tex, y, x, ch, textures) -> True skips a draw (e.g. a
    texel whose bilinear footprint reaches a clamp, where the functional has
    a kink within h)."""
    rng = np.random.default_rng(seed)
    picks, seen = [], set()
    tries = 0
    while len(picks) < count and tries < 100 * count:
        tries += 1
        t, c = channels[int(rng.integers(len(channels)))]
        g = grads[t][..., c]
        nz = np.argwhere(np.abs(g) > rel_floor * np.abs(g).max())
        if len(nz) == 0:
            continue
        y, x = (int(v) for v in nz[int(rng.integers(len(nz)))])
        key = (t, y, x, c)
        if key in seen or (avoid is not None and avoid(t, y, x, c, textures)):
            continue
        seen.add(key)
        picks.append(key)
    return picks


def central_differences(renderer, camera, eval_scene, target, textures: dict, picks, om=None, h=1e-3):
    """Central differences of loss(render_detached(camera, eval_scene)) for
    each picked texel and, with om = (mat_id, n, k), each of the 6 OM (n, k)
    components.  textures: {tex_id: (H, W, 4) float32 texels of eval_scene};
    each is restored after its differences.  The quotient divides by the
    float32 values actually set (x + h and x - h rounded).  Returns a list
    of (key, fd)."""
    out = []
    for (t, y, x, c) in picks:
        img = textures[t]
        x0 = np.float32(img[y, x, c])
        vals = []
        for sgn in (1.0, -1.0):
            xv = np.float32(x0 + np.float32(sgn * h))
            img[y, x, c] = xv
            eval_scene.update_texture(t, img)
            vals += [float(xv), loss(renderer.render_detached(camera, eval_scene), target)]
        img[y, x, c] = x0
        eval_scene.update_texture(t, img)
        out.append((("tex", t, y, x, c), (vals[1] - vals[3]) / (vals[0] - vals[2])))
    if om is not None:
        mat, n0, k0 = om
        n0, k0 = np.asarray(n0, np.float32), np.asarray(k0, np.float32)
        for j in range(6):
            vals = []
            for sgn in (1.0, -1.0):
                nn, kk = n0.copy(), k0.copy()
                arr = nn if j < 3 else kk
                arr[j % 3] = np.float32(arr[j % 3] + np.float32(sgn * h))
                eval_scene.update_om(mat, nn, kk)
                vals += [float(arr[j % 3]), loss(renderer.render_detached(camera, eval_scene), target)]
            eval_scene.update_om(mat, n0, k0)
            out.append((("om", mat, j), (vals[1] - vals[3]) / (vals[0] - vals[2])))
    return out


def relative_errors(checks, grads: dict, om_grads=None):
    """|adjoint - fd| / |fd| for each (key, fd) of central_differences;
    om_grads: the 6 adjoint (n, k) gradients of the OM material."""
    rel = []
    for key, fd in checks:
        if key[0] == "tex":
            _, t, y, x, c = key
            adj = float(grads[t][y, x, c])
        else:
            adj = float(om_grads[key[2]])
        rel.append((key, adj, fd, abs(adj - fd) / max(abs(fd), 1e-30)))
    return rel
