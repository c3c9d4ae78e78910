"""Closed-form pins of the oracle's shading-frame steps, whose reference
translation units (SurfacePoint.cpp, OpenMaterial.cpp,
PhysicallyBasedMaterial.cpp) include Embree headers and do not build here:

* SurfacePoint::applyTextureNormal (SurfacePoint.cpp:19-34): the new normal
  is the tangent-space vector in the (T, B, N) frame; the tangent and
  binormal are rebuilt by cross products with the handedness sign;
* the normal-map texel step of PhysicallyBasedMaterial::modifyFrame
  (PhysicallyBasedMaterial.cpp:158-164), the glTF 2.0 normalTexture rule
  normalize((2 * rgb - 1) * (scale, scale, 1)) with the reference's blue
  floor of 0.5 + eps;
* OpenMaterial::defineNextDirection (OpenMaterial.cpp:72-93): the mirror
  direction about the shading normal, whose component along the geometric
  normal of the face (through the instance's normal matrix, in the face's
  winding) is raised to 0.1 when below it, the tangential part kept.

Expected values come from vector identities for chosen frames and angles,
so a swapped cross product, a lost handedness sign, a missing normal matrix
or a flipped winding fails them.
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def L(ref):
    return ref.lib()


def _rotations(n, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        q, r = np.linalg.qr(rng.normal(size=(3, 3)))
        q = q * np.sign(np.diag(r))
        if np.linalg.det(q) < 0:
            q[:, 2] *= -1.0
        out.append(q)
    return out


def _apply(ref, L, T, B, N, tw, tn, calls=1):
    f = ref.f32(np.concatenate([T, B, N, [tw]]))
    out = np.zeros(9, np.float32)
    L.ref_apply_texture_normal(ref.ptr(f), ref.ptr(ref.f32(tn)), calls, ref.ptr(out))
    return out.reshape(3, 3).astype(np.float64)


def _frame(q):
    # right-handed: T x B = N
    return q[:, 0].astype(np.float32).astype(np.float64), q[:, 1].astype(np.float32).astype(np.float64), \
        q[:, 2].astype(np.float32).astype(np.float64)


def test_texture_normal_up_keeps_a_right_handed_frame(ref, L):
    for q in _rotations(20):
        T, B, N = _frame(q)
        out = _apply(ref, L, T, B, N, 1.0, (0.0, 0.0, 1.0))
        np.testing.assert_allclose(out, np.stack([T, B, N]), atol=2e-6)


def test_texture_normal_up_left_handed(ref, L):
    """Handedness -1 with B = -(N x T): T' = B x N = -T and B' = -(N x T') =
    N x T, so tangent and binormal both turn over and N stays."""
    for q in _rotations(20, seed=6):
        T, B, N = _frame(q)
        out = _apply(ref, L, T, -B, N, -1.0, (0.0, 0.0, 1.0))
        np.testing.assert_allclose(out, np.stack([-T, B, N]), atol=2e-6)


@pytest.mark.parametrize("theta", [0.1, 0.5, 1.0, 1.4])
def test_texture_normal_tilt_about_binormal(ref, L, theta):
    """tn = (sin, 0, cos): N' = sin T + cos N, T' = B x N' = cos T - sin N,
    B' = N' x T' = B."""
    s, c = math.sin(theta), math.cos(theta)
    for q in _rotations(8, seed=7):
        T, B, N = _frame(q)
        out = _apply(ref, L, T, B, N, 1.0, (s, 0.0, c))
        np.testing.assert_allclose(out[2], s * T + c * N, atol=3e-6)
        np.testing.assert_allclose(out[0], c * T - s * N, atol=3e-6)
        np.testing.assert_allclose(out[1], B, atol=3e-6)


@pytest.mark.parametrize("theta", [0.2, 0.9])
def test_texture_normal_tilt_about_tangent(ref, L, theta):
    """tn = (0, sin, cos): N' = sin B + cos N, T' = B x N' = cos T (unit:
    T), B' = N' x T = cos B - sin N."""
    s, c = math.sin(theta), math.cos(theta)
    for q in _rotations(8, seed=8):
        T, B, N = _frame(q)
        out = _apply(ref, L, T, B, N, 1.0, (0.0, s, c))
        np.testing.assert_allclose(out[2], s * B + c * N, atol=3e-6)
        np.testing.assert_allclose(out[0], T, atol=3e-6)
        np.testing.assert_allclose(out[1], c * B - s * N, atol=3e-6)


def test_texture_normal_scale_and_once_only(ref, L):
    T, B, N = _frame(_rotations(1, seed=9)[0])
    base = _apply(ref, L, T, B, N, 1.0, (0.0, 0.0, 1.0))
    np.testing.assert_allclose(_apply(ref, L, T, B, N, 1.0, (0.0, 0.0, 3.0)), base, atol=2e-6)   # length does not matter
    tn = (0.3, -0.2, math.sqrt(1 - 0.13))
    once = _apply(ref, L, T, B, N, 1.0, tn)
    assert np.array_equal(_apply(ref, L, T, B, N, 1.0, tn, calls=3), once)   # the frame is modified once per surface point
    assert not np.allclose(once, base)


def test_texture_normal_frame_stays_orthonormal(ref, L):
    rng = np.random.default_rng(10)
    for q in _rotations(30, seed=11):
        T, B, N = _frame(q)
        tw = float(rng.choice([-1.0, 1.0]))
        v = rng.normal(size=3)
        v[2] = abs(v[2]) + 0.2
        tn = v / np.linalg.norm(v)
        Tp, Bp, Np = _apply(ref, L, T, tw * B, N, tw, tn)
        for x in (Tp, Bp, Np):
            assert abs(np.linalg.norm(x) - 1.0) < 3e-6
        assert abs(Tp @ Np) < 3e-6 and abs(Bp @ Np) < 3e-6 and abs(Tp @ Bp) < 3e-6
        np.testing.assert_allclose(Bp, tw * np.cross(Np, Tp), atol=3e-6)


def _texel(ref, L, rgba, scale):
    out = np.zeros(3, np.float32)
    L.ref_texel_normal(ref.ptr(ref.f32(rgba)), scale, ref.ptr(out))
    return out.astype(np.float64)


def test_texel_normal_gltf_rule(ref, L):
    assert np.array_equal(_texel(ref, L, (0.5, 0.5, 1.0, 1.0), 1.0), [0.0, 0.0, 1.0])   # the flat texel
    np.testing.assert_allclose(_texel(ref, L, (0.75, 0.5, 1.0, 1.0), 1.0), np.array([0.5, 0.0, 1.0]) / math.sqrt(1.25), atol=2e-7)
    np.testing.assert_allclose(_texel(ref, L, (0.75, 0.5, 1.0, 1.0), 2.0), np.array([1.0, 0.0, 1.0]) / math.sqrt(2.0), atol=2e-7)
    np.testing.assert_allclose(_texel(ref, L, (0.25, 0.75, 0.75, 0.0), 0.5), np.array([-0.25, 0.25, 0.5]) / math.sqrt(0.375), atol=2e-7)
    # blue at or below 0.5 is lifted to 0.5 + eps: the normal stays just above the surface
    n = _texel(ref, L, (0.5, 0.25, 0.3, 1.0), 1.0)
    eps = float(np.finfo(np.float32).eps)
    np.testing.assert_allclose(n, [0.0, -1.0, 4.0 * eps], atol=1e-7)
    assert n[2] > 0.0


def _next_dir(ref, L, inV, normal, tri, nm=np.eye(3)):
    out = np.zeros(3, np.float32)
    L.ref_om_next_direction(ref.ptr(ref.f32(inV)), ref.ptr(ref.f32(normal)), ref.ptr(ref.f32(np.asarray(tri).ravel())),
                            ref.ptr(ref.f32(np.asarray(nm).ravel())), ref.ptr(out))
    return out.astype(np.float64)


UP_TRI = [(0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0)]   # (P1 - P0) x (P2 - P0) = +y


def _lifted(out_r, ng):
    """out_r with its ng component raised to 0.1, normalized."""
    v = out_r - (out_r @ ng) * ng + 0.1 * ng
    return v / np.linalg.norm(v)


def test_om_direction_mirror_law(ref, L):
    d = np.array([1.0, -1.0, 0.0]) / math.sqrt(2.0)
    np.testing.assert_allclose(_next_dir(ref, L, d, (0.0, 1.0, 0.0), UP_TRI), np.array([1.0, 1.0, 0.0]) / math.sqrt(2.0), atol=1e-7)
    d = np.array([0.3, -0.8, 0.52])
    d /= np.linalg.norm(d)
    np.testing.assert_allclose(_next_dir(ref, L, d, (0.0, 1.0, 0.0), UP_TRI), d * [1, -1, 1], atol=1e-7)


def test_om_direction_grazing_lift(ref, L):
    c = math.sqrt(1.0 - 0.05 ** 2)
    got = _next_dir(ref, L, (c, -0.05, 0.0), (0.0, 1.0, 0.0), UP_TRI)
    np.testing.assert_allclose(got, np.array([c, 0.1, 0.0]) / math.hypot(c, 0.1), atol=2e-7)


def test_om_direction_tilted_shading_normal(ref, L):
    ns = np.array([0.4, 1.0, 0.0]) / math.hypot(0.4, 1.0)
    d = np.array([1.0, -0.3, 0.2])
    d /= np.linalg.norm(d)
    out_r = d - 2.0 * (d @ ns) * ns
    ng = np.array([0.0, 1.0, 0.0])
    assert out_r @ ng < 0.1                                 # the mirror direction dips below the lift
    np.testing.assert_allclose(_next_dir(ref, L, d, ns, UP_TRI), _lifted(out_r, ng), atol=3e-7)


def test_om_direction_normal_matrix(ref, L):
    """Object plane x + y = 0 under the scale diag(2, 1, 1): the world plane
    X / 2 + Y = 0 has normal (0.5, 1, 0); the normal matrix is diag(0.5, 1, 1)."""
    tri = [(0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, -1.0, 0.0)]
    ng = np.array([0.5, 1.0, 0.0]) / math.hypot(0.5, 1.0)
    ns = np.array([0.0, 1.0, 0.0])
    d = np.array([-0.9, -0.3, 0.3])
    d /= np.linalg.norm(d)
    out_r = d - 2.0 * (d @ ns) * ns
    assert out_r @ ng < 0.1
    np.testing.assert_allclose(_next_dir(ref, L, d, ns, tri, np.diag([0.5, 1.0, 1.0])), _lifted(out_r, ng), atol=3e-7)
    # without the normal matrix the object normal (1, 1, 0) / sqrt 2 would be used
    assert not np.allclose(_next_dir(ref, L, d, ns, tri), _lifted(out_r, ng), atol=1e-3)


def test_om_direction_uses_the_face_winding(ref, L):
    """Swapping P1 and P2 turns the face normal to -y: the mirror direction
    (1, 1, 0) / sqrt 2 lies 0.707 below 0.1 along -y and is pushed down."""
    tri = [UP_TRI[0], UP_TRI[2], UP_TRI[1]]
    d = np.array([1.0, -1.0, 0.0]) / math.sqrt(2.0)
    got = _next_dir(ref, L, d, (0.0, 1.0, 0.0), tri)
    np.testing.assert_allclose(got, np.array([1.0 / math.sqrt(2.0), -0.1, 0.0]) / math.hypot(1.0 / math.sqrt(2.0), 0.1), atol=2e-7)
