"""The oracle's path-replay adjoint against what the mathematics fixes.

The differentiated functional is the image of the reference estimator with
detached sampling: directions and lobes are drawn with the current material
parameters theta0, and each PBR bounce weight is w_ref(theta0) * f(theta) /
f(theta0) for the sampled lobe's integrand f (OM mirror bounces use F(n, k)
directly).  At theta = theta0 it is the plain render (checked bit for bit).

Pins:
  * the complex-IOR Fresnel derivative against central differences of the
    Fresnel formula itself (double precision);
  * the FP64 adjoint gradient of an L2 image loss against central finite
    differences (h = 1e-3, FP64, common random numbers) of the same loss on
    texels of the base-colour and metallic-roughness maps and on the OM
    sphere's (n, k): relative error below 1e-3 wherever the finite difference
    is not negligible (|fd| > 1e-3 * max |fd| of its group);
  * the checks have teeth: an adjoint with a planted mistake (the GGX
    D-term derivative dropped, or the base-colour term of F dropped) fails
    the same finite-difference comparison.
"""
import numpy as np
import pytest


def test_fresnel_gradient_closed_form(ref):
    rng = np.random.default_rng(3)
    for _ in range(40):
        n, k, c = rng.uniform(0.1, 2.5), rng.uniform(0.0, 8.0), rng.uniform(0.02, 1.0)
        g = ref.fresnel_grad(n, k, c)
        h = 1e-6
        rp_n = (np.array(ref.fresnel64(n + h, k, c)) - np.array(ref.fresnel64(n - h, k, c))) / (2 * h)
        rp_k = (np.array(ref.fresnel64(n, k + h, c)) - np.array(ref.fresnel64(n, k - h, c))) / (2 * h)
        np.testing.assert_allclose([g[0], g[2]], rp_n, rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose([g[1], g[3]], rp_k, rtol=1e-6, atol=1e-9)


def _setup(wl, ref, **kw):
    dt, di, view, params, ids = wl.material_fit(**kw)
    st, si = ref.Scene(), ref.Scene()
    dt.apply(st, ref.pbr_desc)
    di.apply(si, ref.pbr_desc)
    cam = ref.camera_desc(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    p = ref.default_params(**params)
    T, _ = st.render(p, cam, fp64=True, threads=8)
    return dt, di, si, cam, p, T.astype(np.float64), ids


def _loss(img, T):
    return float(((img[..., :3] - T[..., :3]) ** 2).sum())


@pytest.fixture(scope="module")
def fit(wl, ref):
    kw = dict(width=28, height=28, spp=4, depth=4, tex=12, hdr_w=64, hdr_h=32, segments=60, tube=8)
    dt, di, si, cam, p, T, ids = _setup(wl, ref, **kw)
    I = si.render_detached(si, p, cam, fp64=True, threads=8)
    dfilm = np.zeros(I.shape, np.float32)
    dfilm[..., :3] = 2.0 * (I[..., :3] - T[..., :3])
    shapes = {0: (12, 12, 4), 1: (12, 12, 4)}
    g, gom = si.adjoint(p, cam, dfilm, textures=(0, 1), om_materials=(1,), tex_shapes=shapes, fp64=True, threads=8)
    return dict(di=di, si=si, cam=cam, p=p, T=T, I=I, g=g, gom=gom, dfilm=dfilm)


def test_detached_at_theta0_is_the_render(fit, ref):
    plain, _ = fit["si"].render(fit["p"], fit["cam"], fp64=True, threads=8)
    np.testing.assert_array_equal(fit["I"].astype(np.float32), plain)


def _fd_texel(fit, ref, tex, y, x, ch, h=1e-3):
    import copy
    out = []
    for sgn in (1.0, -1.0):
        d = copy.deepcopy(fit["di"])
        img = d.textures[tex][1]
        t0 = np.float32(img[y, x, ch])
        img[y, x, ch] = np.float32(t0 + np.float32(sgn * h))
        out.append(float(img[y, x, ch]))
        e = ref.Scene()
        d.apply(e, ref.pbr_desc)
        out.append(_loss(fit["si"].render_detached(e, fit["p"], fit["cam"], fp64=True, threads=8), fit["T"]))
    tp, lp, tm, lm = out
    return (lp - lm) / (tp - tm)


def _fd_om(fit, ref, which, c, h=1e-3):
    import copy
    out = []
    for sgn in (1.0, -1.0):
        d = copy.deepcopy(fit["di"])
        m = list(d.materials[1])
        n, k = np.array(m[1], np.float32), np.array(m[2], np.float32)
        arr = n if which == 0 else k
        arr[c] = np.float32(arr[c] + np.float32(sgn * h))
        out.append(float(arr[c]))
        d.materials[1] = ("om_nk", n, k)
        e = ref.Scene()
        d.apply(e, ref.pbr_desc)
        out.append(_loss(fit["si"].render_detached(e, fit["p"], fit["cam"], fp64=True, threads=8), fit["T"]))
    tp, lp, tm, lm = out
    return (lp - lm) / (tp - tm)


@pytest.mark.parametrize("tex,ch", [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2)])
def test_texture_gradient_matches_fd(fit, ref, tex, ch):
    g = fit["g"][tex][..., ch]
    order = np.argsort(-np.abs(g).ravel())
    picks = list(order[:5]) + list(np.random.default_rng(tex * 10 + ch).choice(order[:40], 3, replace=False))
    fds, adj = [], []
    for flat in picks:
        y, x = np.unravel_index(flat, g.shape)
        fds.append(_fd_texel(fit, ref, tex, y, x, ch))
        adj.append(g[y, x])
    fds, adj = np.array(fds), np.array(adj)
    scale = np.abs(fds).max()
    assert scale > 0
    sig = np.abs(fds) > 1e-3 * scale
    rel = np.abs(adj - fds)[sig] / np.abs(fds)[sig]
    assert rel.max() < 1e-3, (tex, ch, list(zip(adj, fds)))


def test_om_gradient_matches_fd(fit, ref):
    dn, dk = fit["gom"][1]
    for which, arr in ((0, dn), (1, dk)):
        for c in range(3):
            fd = _fd_om(fit, ref, which, c)
            assert abs(arr[c] - fd) <= 1e-3 * abs(fd) + 1e-9, (which, c, arr[c], fd)


def test_float_adjoint_close_to_fp64(fit, ref):
    """Float and FP64 renders of the same sample take different branches on a
    few paths (fireflies at 4 spp), so the float adjoint is compared by its
    median texel error and by the OM gradients, not texel by texel."""
    shapes = {0: (12, 12, 4), 1: (12, 12, 4)}
    g32, gom32 = fit["si"].adjoint(fit["p"], fit["cam"], fit["dfilm"], textures=(0, 1), om_materials=(1,), tex_shapes=shapes,
                                   fp64=False, threads=8)
    for t in (0, 1):
        a, b = g32[t].ravel(), fit["g"][t].ravel()
        sig = np.abs(b) > 1e-3 * np.abs(b).max()
        assert np.median(np.abs(a - b)[sig] / np.abs(b)[sig]) < 1e-3
    np.testing.assert_allclose(np.concatenate(gom32[1]), np.concatenate(fit["gom"][1]), rtol=2e-2)


def test_adjoint_rejects_point_lights(wl, ref):
    desc, view = wl.mixed_test_scene(hdr=True, lights=True)
    s = ref.Scene()
    desc.apply(s, ref.pbr_desc)
    cam = ref.camera_desc(8, 8, view.position, view.forward, view.up, view.yfov)
    with pytest.raises(RuntimeError):
        s.adjoint(ref.default_params(samples_per_pixel=1), cam, np.ones((8, 8, 4), np.float32), textures=(0,),
                  tex_shapes={0: (64, 64, 4)})


@pytest.mark.parametrize("mask,tex,ch", [(1, 1, 1), (2, 0, 1)])
def test_planted_mistake_is_caught(fit, ref, mask, tex, ch):
    shapes = {0: (12, 12, 4), 1: (12, 12, 4)}
    ref.lib().ref_adjoint_plant(mask)
    try:
        g, _ = fit["si"].adjoint(fit["p"], fit["cam"], fit["dfilm"], textures=(tex,), tex_shapes=shapes, fp64=True, threads=8)
    finally:
        ref.lib().ref_adjoint_plant(0)
    good = fit["g"][tex][..., ch]
    flat = int(np.argmax(np.abs(good)))
    y, x = np.unravel_index(flat, good.shape)
    fd = _fd_texel(fit, ref, tex, y, x, ch)
    assert abs(good[y, x] - fd) < 1e-3 * abs(fd)
    assert abs(g[tex][y, x, ch] - fd) > 1e-2 * abs(fd)
