"""OMV_EXT_HDRI_NEE in the oracle: environment importance sampling + MIS.

Pins (the mathematics fixes all of them):
  * the alias tables reproduce the texel probabilities: for every table,
    prob/alias give each entry exactly n * p_k / n mass (computed exactly);
  * sampling and density agree: env_pdf(direction of a sample) equals the
    sample's own density, and directions land in the sampled texel;
  * the density integrates to 1 over the sphere (Monte Carlo with uniform
    directions);
  * the estimator is unbiased: with NEE the image mean equals the plain
    reference estimator's mean within 4 standard errors (the two estimate the
    same integral), and its error against a high-spp FP64 reference is lower.
"""
import numpy as np
import pytest


def _env_scene(wl, ref, w=64, h=32, sun=200.0):
    desc, view = wl.mixed_test_scene(hdr=True, lights=False, missing=False)
    sky = wl.hdri(w, h, seed=9, sun_factor=sun)
    desc.background = ("hdr", sky, wl.HDR_ROTATION, wl.HDR_SCALE)
    s = ref.Scene()
    desc.apply(s, ref.pbr_desc)
    return desc, view, s, sky


def test_alias_tables_reproduce_probabilities(wl, ref):
    _, _, s, sky = _env_scene(wl, ref)
    h, w = sky.shape[:2]
    rp, ra, cp, ca, tp = s.env_tables(w, h)
    assert abs(tp.sum() - 1.0) < 1e-5

    def mass(prob, alias):
        n = len(prob)
        m = np.zeros(n)
        for k in range(n):
            m[k] += prob[k] / n
            m[alias[k]] += (1.0 - prob[k]) / n
        return m

    row = tp.sum(axis=1)
    np.testing.assert_allclose(mass(rp.astype(np.float64), ra), row / row.sum(), rtol=1e-4, atol=1e-7)
    for i in range(0, h, 5):
        if row[i] > 0:
            np.testing.assert_allclose(mass(cp[i].astype(np.float64), ca[i]), tp[i] / row[i], rtol=1e-4, atol=1e-7)


def test_sample_density_consistent(wl, ref):
    _, _, s, sky = _env_scene(wl, ref)
    h, w = sky.shape[:2]
    rng = np.random.default_rng(1)
    for _ in range(400):
        u = rng.random(6).astype(np.float32)
        d, pdf, row, col = s.env_sample(u)
        assert abs(np.linalg.norm(d) - 1.0) < 1e-9
        if pdf > 0:
            np.testing.assert_allclose(s.env_pdf(d), pdf, rtol=1e-6)


def test_density_integrates_to_one(wl, ref):
    _, _, s, _ = _env_scene(wl, ref, sun=5.0)
    rng = np.random.default_rng(2)
    v = rng.normal(size=(60000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    vals = np.array([s.env_pdf(x) for x in v])
    est = 4 * np.pi * vals.mean()
    se = 4 * np.pi * vals.std() / np.sqrt(len(vals))
    assert abs(est - 1.0) < 4 * se + 1e-3, (est, se)


def test_nee_is_unbiased_and_reduces_error(wl, ref):
    desc, view, s, _ = _env_scene(wl, ref, sun=200.0)
    cam = ref.camera_desc(24, 18, view.position, view.forward, view.up, view.yfov)
    base = dict(camera_bounces=4, light_bounces=2, max_path_length=5, seed=3)
    spp = 64
    plain = s.render_samples(ref.default_params(samples_per_pixel=spp, **base), cam, fp64=True, threads=8)
    nee = s.render_samples(ref.default_params(samples_per_pixel=spp, extensions=2, **base), cam, fp64=True, threads=8)
    # per-sample radiance -> image mean and its standard error
    def stats(x):
        per = x.reshape(-1, 3).astype(np.float64)
        return per.mean(axis=0), per.std(axis=0) / np.sqrt(per.shape[0])
    mp, sp_ = stats(plain)
    mn, sn = stats(nee)
    assert np.all(np.abs(mp - mn) < 4 * np.sqrt(sp_ ** 2 + sn ** 2)), (mp, mn, sp_, sn)
    hi, _ = s.render(ref.default_params(samples_per_pixel=1024, **dict(base, seed=777)), cam, fp64=True, threads=8)
    ip, inee = plain.mean(axis=2), nee.mean(axis=2)
    rp = np.sqrt(np.mean((ip - hi[..., :3]) ** 2))
    rn = np.sqrt(np.mean((inee - hi[..., :3]) ** 2))
    assert rn < rp, (rn, rp)
