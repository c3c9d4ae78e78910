/*
 * refcpu.cpp -- CPU oracle: a from-scratch restatement of the reference
 * pathtracer's BDPT render loop.  TEST INFRASTRUCTURE ONLY: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it as the
 * checker; the product never links or calls it.
 *
 * What each part restates (reference file:line, paths under pathtracer/src/),
 * and what pins it (tests/):
 *
 *   Philox4x32-10 .............. replaces the shared std::mt19937 of
 *                                RenderScene.cpp:20 / RandomSampler.h:38-55.
 *                                Pinned: Random123 known-answer vectors and
 *                                rocRAND's independent header implementation.
 *   warps (uniformHemisphere, uniformSphere, GGXHemisphere,
 *          cosineSqrtHemisphere) RandomSampler.h:66-103.  Pinned: golden
 *                                vectors from the reference header fed the
 *                                same mt19937 stream; closed-form moments.
 *   DistanceRandom ............. DistanceRandom.h:85-121.  Pinned: golden.
 *   Interpolation / getIor ..... Interpolation.h:44-78,170-182,
 *                                AssetMaterialIor.cpp:115-178.  Pinned: golden
 *                                (reference AssetMaterialIor on the Au/Al/Fe
 *                                tables) + table knots.
 *   fresnelReflection .......... OpenMaterial.cpp:20-38.  Pinned: closed forms
 *                                (normal incidence ((n-1)^2+k^2)/((n+1)^2+k^2),
 *                                grazing R=1, k=0 dielectric Fresnel) and the
 *                                reference test OpenMaterial.cpp:42-52.
 *   BSDFModel evaluate / evaluateDirect / sampleIndirect, GGX, SmithG
 *                                BRDFModel.h:25-337.  Pinned: golden vectors
 *                                from the reference header.
 *   BitmapTexture::texture ..... BitmapTexture.cpp:181-234.  Pinned: golden.
 *   HDRBackground::getRadiance . HdrBackground.cpp:55-97.  Pinned: golden.
 *   PinholeCamera getRay/setYFoV/adjustToResolution/stop
 *                                PinholeCamera.cpp:65-174.  Pinned: golden.
 *   gammaCorrection ............ ToneMapping.cpp:24-54.  Pinned: golden.
 *   RenderLightPoint ........... RenderLightPoint.cpp:20-45.  Pinned: golden
 *                                (power) + +octant invariant.
 *   RenderMesh commit / computeNormals / computeTangents / randomTangent /
 *   computeSurfacePoint ........ RenderMesh.cpp:49-432.  The reference TU needs
 *                                Embree headers (unbuildable here).  Pinned:
 *                                hand-computed small meshes and invariants
 *                                (unit normals, tangent orthogonality,
 *                                handedness sign) in tests.
 *   SurfacePoint::applyTextureNormal  SurfacePoint.cpp:19-34.  Pinned:
 *                                closed forms for rotated frames of both
 *                                handednesses and tilts about T and B;
 *                                orthonormality; applied once per point.
 *   OpenMaterial ............... OpenMaterial.cpp:66-232.  Unbuildable TU.
 *                                Pinned: mirror closed form (throughput = F),
 *                                gate; defineNextDirection by the mirror law,
 *                                the 0.1 lift along the face normal (grazing,
 *                                tilted shading normal, normal matrix of a
 *                                non-uniform scale, reversed winding).
 *   normal-map texel step ...... PhysicallyBasedMaterial.cpp:158-164.
 *                                Pinned: the glTF 2.0 normalTexture rule and
 *                                the blue floor.
 *   PhysicallyBasedMaterial .... PhysicallyBasedMaterial.cpp:27-316.
 *                                Unbuildable TU.  Pinned: evaluateDirect/
 *                                sampleIndirect golden vectors (the material
 *                                only feeds them) + FC values of the
 *                                reference FC tests (FCSubject.cpp:189-315).
 *   DiffuseColorMaterial ....... DiffuseColorMaterial.cpp:35-108.
 *   BDPTRenderer ............... BDPTRenderer.cpp:35-393.  Unbuildable TU.
 *                                Pinned: closed-form scenes (constant
 *                                background -> every pixel = background;
 *                                mirror sphere -> 100*F(cos); one lit PBR
 *                                plane -> evaluateDirect * geometry by hand).
 *                                isConnected (:208-235): an occluder between
 *                                the light and a lit floor -> 0; a second
 *                                floor just beyond the lit point, inside
 *                                tfar -> film unchanged.
 *   FC paint rules ............. FCSubject.h:153-482.  Pinned: golden
 *                                vectors from the reference's own subject
 *                                classes (FC_VALIDATION build of the header)
 *                                and its unit tests FCSubject.cpp:189-315.
 *   FC density / gradient ...... RenderMesh.cpp:197-212, BDPTRenderer.cpp:
 *                                69-71.  Pinned: closed forms on triangles of
 *                                known area and known vertex-normal tilt.
 *   Embree rtcIntersect1 ....... replaced by the watertight ray/triangle test
 *                                of Woop, Benthin & Wald (JCGT 2013) with a
 *                                deterministic (t, instance, primitive)
 *                                tie-break.  Embree's own edge behaviour is
 *                                parity unpinned (no reference test exists).
 *                                Pinned: brute force == BVH closest hit;
 *                                (t, u, v) against a float64 intersection;
 *                                no gaps on shared edges and vertices.
 *
 * Two shortcuts are exact (they change no result; see the comments where
 * they sit): light subpaths stop after max_path_length-1 bounces and camera
 * subpaths stop once their throughput is exactly zero or, in scenes without
 * emissive materials, once they have more than max_path_length vertices.
 * Camera and light subpaths draw from separate Philox sub-streams so these
 * shortcuts cannot shift any random number.
 */
#include "refcpu.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <mutex>
#include <thread>
#include <set>
#include <vector>

namespace refcpu {

/* ------------------------------------------------------------------ RNG */
static const uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
static const uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

static void philox(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += kPhiloxW0; k1 += kPhiloxW1; }
        uint64_t p0 = (uint64_t)kPhiloxM0 * c0, p1 = (uint64_t)kPhiloxM1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform float in [0,1) with 24 random bits */
static inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

enum { STREAM_CAMERA = 0, STREAM_LIGHT = 1, STREAM_NEE = 2 };  /* NEE: 2 + 16 * bounce */

/* One sample's sub-stream: key (seed, 0), counter (call, stream, sample lo, hi). */
struct Stream {
    uint32_t key[2], ctr[4], buf[4];
    int idx;
    Stream(uint32_t seed, uint32_t stream, uint64_t sample) {
        key[0] = seed; key[1] = 0;
        ctr[0] = 0; ctr[1] = stream; ctr[2] = (uint32_t)sample; ctr[3] = (uint32_t)(sample >> 32);
        idx = 4;
    }
    float next() {
        if (idx == 4) { philox(ctr, key, buf); ++ctr[0]; idx = 0; }
        return u01(buf[idx++]);
    }
};

/* Uniform source for the warps: either a Philox stream or an explicit list. */
struct Uniforms {
    Stream* s = nullptr;
    const float* list = nullptr;
    int n = 0, used = 0;
    float next() {
        if (s) { ++used; return s->next(); }
        float v = used < n ? list[used] : 0.5f;
        ++used;
        return v;
    }
};

/* ----------------------------------------------------------------- math */
template <class R> struct V3 { R x, y, z; };
template <class R> static inline V3<R> mk(R a, R b, R c) { return V3<R>{a, b, c}; }
template <class R> static inline R at(const V3<R>& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
template <class R> static inline R& at(V3<R>& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

/* UtilityFunctions.h:128-247 with the same operation order */
template <class R> static inline R dot3(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> static inline V3<R> cross3(const V3<R>& a, const V3<R>& b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> static inline R length3(const V3<R>& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
template <class R> static inline V3<R> sub(const V3<R>& a, const V3<R>& b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> static inline V3<R> neg(const V3<R>& a) { return mk(-a.x, -a.y, -a.z); }
template <class R> static inline R distance3(const V3<R>& a, const V3<R>& b) { return length3(sub(a, b)); }
template <class R> static inline V3<R> normalize3(const V3<R>& a) {
    R L = length3(a);
    if (L != R(0)) { L = R(1) / L; return mk(a.x * L, a.y * L, a.z * L); }
    return mk(R(1), R(0), R(0));
}
template <class R> static inline V3<R> normalizeIfNeeded3(const V3<R>& a) {
    const R eps = R(10) * std::numeric_limits<R>::epsilon();
    R sL = dot3(a, a);
    if (sL < R(1) + eps && sL > R(1) - eps) return a;
    R L = std::sqrt(sL);
    if (L != R(0)) { L = R(1) / L; return mk(a.x * L, a.y * L, a.z * L); }
    return mk(R(1), R(0), R(0));
}
template <class R> static inline R mn(R a, R b) { return a < b ? a : b; }
template <class R> static inline R mx(R a, R b) { return a > b ? a : b; }
template <class R> static inline R clampv(R v, R lo, R hi) { return mx(lo, mn(v, hi)); }
template <class R> static inline R clamp01(R v) { return clampv(v, R(0), R(1)); }
template <class R> static inline R lerpv(R lo, R hi, R t) { return (hi - lo) * t + lo; }
/* reflect3 / refract3 (UtilityFunctions.h:472-495) */
template <class R> static inline V3<R> reflect3(const V3<R>& I, const V3<R>& N) {
    const R w = R(2) * dot3(I, N);
    return mk(I.x - w * N.x, I.y - w * N.y, I.z - w * N.z);
}
template <class R> static inline bool refract3(const V3<R>& I, const V3<R>& N, R eta, V3<R>& out) {
    R NI = dot3(N, I);
    R k = R(1) - eta * eta * (R(1) - NI * NI);
    if (k < R(0)) { out = neg(I); return false; }
    k = eta * NI + std::sqrt(k);
    out = mk(eta * I.x - k * N.x, eta * I.y - k * N.y, eta * I.z - k * N.z);
    return true;
}
/* multVector3 variants (UtilityFunctions.h:285-307) */
template <class R> static inline V3<R> mul33(const float* A, const V3<R>& v) {  // A row-major 3x3
    return mk(R(A[0]) * v.x + R(A[1]) * v.y + R(A[2]) * v.z, R(A[3]) * v.x + R(A[4]) * v.y + R(A[5]) * v.z,
              R(A[6]) * v.x + R(A[7]) * v.y + R(A[8]) * v.z);
}
template <class R> static inline V3<R> mul34(const float* A, const V3<R>& v) {  // A row-major 3x4 (+translation)
    return mk(R(A[0]) * v.x + R(A[1]) * v.y + R(A[2]) * v.z + R(A[3]),
              R(A[4]) * v.x + R(A[5]) * v.y + R(A[6]) * v.z + R(A[7]),
              R(A[8]) * v.x + R(A[9]) * v.y + R(A[10]) * v.z + R(A[11]));
}
template <class R> static inline V3<R> mul33tm(const float* A, const V3<R>& v) {  // the 3x3 part of a row-major 3x4
    return mk(R(A[0]) * v.x + R(A[1]) * v.y + R(A[2]) * v.z, R(A[4]) * v.x + R(A[5]) * v.y + R(A[6]) * v.z,
              R(A[8]) * v.x + R(A[9]) * v.y + R(A[10]) * v.z);
}
template <class R> static inline V3<R> mulT33(const V3<R>& v, const float* A) {  // v * A (row vector)
    return mk(R(A[0]) * v.x + R(A[3]) * v.y + R(A[6]) * v.z, R(A[1]) * v.x + R(A[4]) * v.y + R(A[7]) * v.z,
              R(A[2]) * v.x + R(A[5]) * v.y + R(A[8]) * v.z);
}
template <class R> static inline V3<R> ld3(const float* p) { return mk(R(p[0]), R(p[1]), R(p[2])); }
template <class R> static inline R kPi();
template <> inline float kPi<float>() { return 3.14159265358979323846f; }
template <> inline double kPi<double>() { return 3.14159265358979323846; }
template <class R> static inline R feps() { return std::numeric_limits<R>::epsilon(); }
/* Rec.601 luma used for the cut thresholds (BDPTRenderer.cpp:52-55) */
template <class R> static inline R intensity(const V3<R>& r) { return R(0.299f) * r.x + R(0.587f) * r.y + R(0.114f) * r.z; }
/* maxAbsInVector (BDPTRenderer.cpp:35-40) */
template <class R> static inline R maxAbs(const V3<R>& v, R th = R(1)) {
    R a = std::abs(v.x), b = std::abs(v.y), c = std::abs(v.z);
    R m = (a > b && a > c) ? a : (b > c ? b : c);
    return m < th ? th : m;
}

/* ---------------------------------------------------------------- scene */
enum MatType { MAT_NONE = 0, MAT_OM = 1, MAT_PBR = 2, MAT_DIFFUSE = 3 };

struct Mesh {
    int id = -1, vn = 0, fn = 0, texChN = 0, stride = 0, filledTexChN = 0, matId = -1;
    std::vector<uint32_t> F;
    std::vector<float> V;
    int texMap[8], texOffs[8];
    bool texDef[8], vertexDef = false, normalDef = false, tangentDef = false, valid = false;
    Mesh() { for (int i = 0; i < 8; ++i) { texMap[i] = -1; texOffs[i] = 0; texDef[i] = false; } }
    const float* vtx(uint32_t i) const { return V.data() + (size_t)i * stride; }
};

struct Instance {
    int instId = -1, meshId = -1;
    float tm[16];       // row-major 4x4 (top 3x4 used)
    float nm[9];        // normal matrix: inverse of the transposed 3x3 (RenderMesh.h:36-43)
    float inv[12];      // world -> object 3x4, computed in double and rounded
    double invd[12];
    float wbmin[3], wbmax[3];
};

struct Texture {
    int w = 0, h = 0;
    bool f32 = false;
    std::vector<uint8_t> rgba;
    std::vector<float> rgbaf;
};

struct Material {
    int type = MAT_NONE;
    bool omHasBase = false;
    float n[3] = {0, 0, 0}, k[3] = {0, 0, 0};
    omv_pbr_desc pbr;
    float diffuse[4] = {1000.0f, 0.0f, 1000.0f, 1.0f};
    bool pbrEmissive = false;
    // raw IOR tables kept until commit
    float omTemperature = 300.0f;
    std::vector<float> tTemp;
    std::vector<std::vector<float>> tnw, tnv, tkw, tkv;
    Material() { std::memset(&pbr, 0, sizeof(pbr)); }
};

struct Light { float pos[3], intensity[3], range, power; bool set = false; };

struct BNode { float bmin[3], bmax[3]; int left, right, start, count; };  // count>0: leaf

struct Bvh {
    std::vector<BNode> nodes;
    std::vector<int> prims;
};

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Instance> instances;
    std::vector<Material> materials;  // size materials+1, last = missing
    std::vector<Texture> textures;
    std::vector<Light> lights;
    std::vector<float> lightCdf;
    float lightMax = 0.0f;
    int bgType = 0;  // 0 none, 1 constant, 2 hdr
    float bgConst[3] = {100, 100, 100};
    int bgW = 0, bgH = 0;
    std::vector<float> bgRgb;
    float bgRot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float bgScale = 1.0f;
    int fcSubject = OMV_FC_NONE;
    float minDensity = std::numeric_limits<float>::infinity(), maxDensity = 0.0f;
    bool committed = false, anyEmissive = false;
    std::vector<Bvh> blas;
    Bvh tlas;
    std::vector<int> faceBase;
    /* environment importance sampling (OMV_EXT_HDRI_NEE): a row table and one
     * column table per row (Vose alias method), and the texel probabilities */
    std::vector<float> envRowProb, envColProb, envTexPdf;
    std::vector<int> envRowAlias, envColAlias;
};

/* ------------------------------------------ environment alias tables */
/* Texel weight: Rec.709 luminance times sin(polar angle) at the row centre,
 * the sine by Bhaskara I's rational form 16x(pi-x) / (5pi^2 - 4x(pi-x)) so
 * that host and device build identical tables from + - * / alone. */
static float envRowSine(int i, int H) {
    const float pi = 3.14159265358979323846f;
    const float x = ((float)i + 0.5f) * pi / (float)H;
    const float a = x * (pi - x);
    return 16.0f * a / (5.0f * pi * pi - 4.0f * a);
}
/* Vose's alias method over n weights, sequential, in the order the device
 * kernel uses (sum in index order, LIFO work lists) */
static void voseBuild(const float* w, int n, float* prob, int* alias, std::vector<float>& p, std::vector<int>& small,
                      std::vector<int>& large) {
    float S = 0.0f;
    for (int k = 0; k < n; ++k) S = S + w[k];
    if (!(S > 0.0f)) { for (int k = 0; k < n; ++k) { prob[k] = 1.0f; alias[k] = k; } return; }
    const float scale = (float)n / S;
    p.resize(n); small.resize(n); large.resize(n);
    int ns = 0, nl = 0;
    for (int k = 0; k < n; ++k) {
        p[k] = w[k] * scale;
        if (p[k] < 1.0f) small[ns++] = k; else large[nl++] = k;
    }
    while (ns > 0 && nl > 0) {
        const int l = small[--ns], g = large[--nl];
        prob[l] = p[l]; alias[l] = g;
        p[g] = (p[g] + p[l]) - 1.0f;
        if (p[g] < 1.0f) small[ns++] = g; else large[nl++] = g;
    }
    while (nl > 0) { const int g = large[--nl]; prob[g] = 1.0f; alias[g] = g; }
    while (ns > 0) { const int l = small[--ns]; prob[l] = 1.0f; alias[l] = l; }
}
static void buildEnvTables(Scene& S) {
    S.envRowProb.clear(); S.envColProb.clear(); S.envTexPdf.clear(); S.envRowAlias.clear(); S.envColAlias.clear();
    if (S.bgType != 2 || S.bgW <= 0 || S.bgH <= 0) return;
    const int W = S.bgW, H = S.bgH;
    std::vector<float> w((size_t)W * H), rowW(H);
    for (int i = 0; i < H; ++i) {
        const float sn = envRowSine(i, H);
        float rs = 0.0f;
        for (int j = 0; j < W; ++j) {
            const float* t = &S.bgRgb[((size_t)i * W + j) * 3];
            const float lum = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
            const float v = std::max(lum, 0.0f) * sn;
            w[(size_t)i * W + j] = v;
            rs = rs + v;
        }
        rowW[i] = rs;
    }
    float total = 0.0f;
    for (int i = 0; i < H; ++i) total = total + rowW[i];
    S.envRowProb.resize(H); S.envRowAlias.resize(H);
    S.envColProb.resize((size_t)W * H); S.envColAlias.resize((size_t)W * H); S.envTexPdf.resize((size_t)W * H);
    std::vector<float> p; std::vector<int> sm, lg;
    voseBuild(rowW.data(), H, S.envRowProb.data(), S.envRowAlias.data(), p, sm, lg);
    for (int i = 0; i < H; ++i)
        voseBuild(&w[(size_t)i * W], W, &S.envColProb[(size_t)i * W], &S.envColAlias[(size_t)i * W], p, sm, lg);
    for (size_t k = 0; k < w.size(); ++k) S.envTexPdf[k] = total > 0.0f ? w[k] / total : 0.0f;
}

/* ------------------------------------------------------- mesh commit */
/* RenderMesh::computeNormals (RenderMesh.cpp:300-330) */
static void computeNormals(Mesh& m) {
    const float e = std::numeric_limits<float>::epsilon();
    for (int i = 0; i < m.vn; ++i) { float* cn = &m.V[(size_t)i * m.stride + 3]; cn[0] = 0.0f; cn[1] = 0.0f; cn[2] = e; }
    for (int i = 0; i < m.fn; ++i) {
        const float* p0 = m.vtx(m.F[3 * i]); const float* p1 = m.vtx(m.F[3 * i + 1]); const float* p2 = m.vtx(m.F[3 * i + 2]);
        V3<float> v0 = mk(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]);
        V3<float> v1 = mk(p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]);
        V3<float> N = cross3(v0, v1);
        if (N.x * N.x + N.y * N.y + N.z * N.z > e) N = normalize3(N);
        for (int j = 0; j < 3; ++j) {
            float* cn = &m.V[(size_t)m.F[3 * i + j] * m.stride + 3];
            cn[0] += N.x; cn[1] += N.y; cn[2] += N.z;
        }
    }
    for (int i = 0; i < m.vn; ++i) {
        float* cn = &m.V[(size_t)i * m.stride + 3];
        V3<float> n = normalize3(mk(cn[0], cn[1], cn[2]));
        cn[0] = n.x; cn[1] = n.y; cn[2] = n.z;
    }
    m.normalDef = true;
}

/* RenderMesh::randomTangent (RenderMesh.cpp:418-432) */
static void randomTangent(const float* nv, float* tv) {
    V3<float> n = mk(nv[0], nv[1], nv[2]), t;
    if (nv[2] < 0.95f && nv[2] > -0.95f) t = cross3(mk(0.0f, 0.0f, 1.0f), n);
    else if (nv[1] < 0.95f && nv[1] > -0.95f) t = cross3(mk(0.0f, 1.0f, 0.0f), n);
    else t = cross3(mk(1.0f, 0.0f, 0.0f), n);
    t = normalize3(t);
    tv[0] = t.x * 1.0f; tv[1] = t.y * 1.0f; tv[2] = t.z * 1.0f; tv[3] = 1.0f;
}

/* RenderMesh::computeTangents(int) (RenderMesh.cpp:332-409) */
static void computeTangentsUV(Mesh& m, int ch) {
    const float e = std::numeric_limits<float>::epsilon();
    std::vector<float> tv((size_t)m.vn * 3, 0.0f), bv((size_t)m.vn * 3, 0.0f);
    const int k = m.texOffs[ch];
    for (int i = 0; i < m.fn; ++i) {
        const uint32_t f0 = m.F[3 * i], f1 = m.F[3 * i + 1], f2 = m.F[3 * i + 2];
        const float *X0 = m.vtx(f0), *X1 = m.vtx(f1), *X2 = m.vtx(f2);
        const float *U0 = X0 + k, *U1 = X1 + k, *U2 = X2 + k;
        float v0[3] = {X1[0] - X0[0], X1[1] - X0[1], X1[2] - X0[2]};
        float v1[3] = {X2[0] - X0[0], X2[1] - X0[1], X2[2] - X0[2]};
        float t0[2] = {U1[0] - U0[0], U1[1] - U0[1]};
        float t1[2] = {U2[0] - U0[0], U2[1] - U0[1]};
        float det = t0[0] * t1[1] - t1[0] * t0[1];
        if (det < e) continue;
        float r = 1.0f / det;
        float S00 = r * t1[1], S01 = -r * t1[0], S10 = -r * t0[1], S11 = r * t0[0];
        float ud[3], vd[3];
        for (int j = 0; j < 3; ++j) { ud[j] = v0[j] * S00 + v1[j] * S10; vd[j] = v0[j] * S01 + v1[j] * S11; }
        for (int j = 0; j < 3; ++j) {
            tv[f0 * 3 + j] += ud[j]; tv[f1 * 3 + j] += ud[j]; tv[f2 * 3 + j] += ud[j];
            bv[f0 * 3 + j] += vd[j]; bv[f1 * 3 + j] += vd[j]; bv[f2 * 3 + j] += vd[j];
        }
    }
    for (int i = 0; i < m.vn; ++i) {
        float* tan = &m.V[(size_t)i * m.stride + 6];
        const float* nv = &m.V[(size_t)i * m.stride + 3];
        V3<float> n = mk(nv[0], nv[1], nv[2]), t = mk(tv[3 * i], tv[3 * i + 1], tv[3 * i + 2]);
        float ntv = dot3(n, t);
        tan[0] = t.x - nv[0] * ntv; tan[1] = t.y - nv[1] * ntv; tan[2] = t.z - nv[2] * ntv;
        float len = length3(mk(tan[0], tan[1], tan[2]));
        if (len <= e) randomTangent(nv, tan);
        else { tan[0] /= len; tan[1] /= len; tan[2] /= len; }
        V3<float> nt = cross3(n, mk(tan[0], tan[1], tan[2]));
        tan[3] = dot3(nt, mk(bv[3 * i], bv[3 * i + 1], bv[3 * i + 2])) < 0.0f ? -1.0f : 1.0f;
    }
    m.tangentDef = true;
}

/* ------------------------------------------------------- IOR lookup */
/* Interpolation::_get (Interpolation.h:44-65) for a linear table */
static float interpLinear(const std::vector<float>& xs, const std::vector<float>& ys, float x) {
    int L = 0, Rr = (int)xs.size() - 1;
    while (Rr - L > 1) {
        int M = (L + Rr) / 2;
        if (x < xs[M]) Rr = M; else L = M;
    }
    float dx = xs[Rr] - xs[L];
    float dy = ys[Rr] - ys[L];
    return ys[L] + (x - xs[L]) * dy / dx;
}

struct IorTable { float T; std::vector<float> nw, nv, kw, kv; float mn, mx; };

static std::vector<IorTable> makeTables(int n, const omv_ior_table* t) {
    std::vector<IorTable> out;
    for (int i = 0; i < n; ++i) {
        IorTable it;
        it.T = t[i].temperature;
        std::vector<std::pair<float, float>> pn, pk;
        for (int j = 0; j < t[i].n_count; ++j) pn.emplace_back(t[i].n_wavelength[j], t[i].n_value[j]);
        for (int j = 0; j < t[i].k_count; ++j) pk.emplace_back(t[i].k_wavelength[j], t[i].k_value[j]);
        std::sort(pn.begin(), pn.end());  // Interpolation::sort (Interpolation.h:143-146)
        std::sort(pk.begin(), pk.end());
        for (auto& p : pn) { it.nw.push_back(p.first); it.nv.push_back(p.second); }
        for (auto& p : pk) { it.kw.push_back(p.first); it.kv.push_back(p.second); }
        if (it.nw.empty() || it.kw.empty()) { it.mn = 1.0f; it.mx = 0.0f; }
        else { it.mn = std::max(it.nw.front(), it.kw.front()); it.mx = std::min(it.nw.back(), it.kw.back()); }
        out.push_back(it);
    }
    std::stable_sort(out.begin(), out.end(), [](const IorTable& a, const IorTable& b) { return a.T < b.T; });
    return out;
}

/* AssetMaterialIor::findClosest + getIor (AssetMaterialIor.cpp:115-178).
 * The Lorentz branch is never populated by the loader, so a wavelength
 * outside the table is the reference's throw; reported as -1. */
static int getIor(const std::vector<IorTable>& tabs, float T, float wl, float& n, float& k) {
    if (tabs.empty()) return -1;
    size_t index = 0;
    float delta = std::fabs(tabs[0].T - T);
    for (size_t i = 1; i < tabs.size(); ++i) {
        float di = std::fabs(tabs[i].T - T);
        if (di < delta) { index = i; delta = di; }
    }
    const IorTable& d = tabs[index];
    if (d.mn <= wl && wl <= d.mx) {
        if (d.nw.size() < 2 || d.kw.size() < 2) return -1;
        n = interpLinear(d.nw, d.nv, wl);
        k = interpLinear(d.kw, d.kv, wl);
        return 0;
    }
    return -1;
}

/* physicalconstants::rgb_wl (Constants.h:53) */
static const float kRgbWl[3] = {6.50000004e-07f, 5.10000007e-07f, 4.40000008e-07f};

/* ------------------------------------------------------------ textures */
/* BitmapTexture::getPixCoords (BitmapTexture.cpp:181-208): the four texels
 * and bilinear weights of (u, v), texel-centre wrap addressing.  Returns
 * false for a missing image (texture() then returns 0.5 grey). */
static bool texAddr(const Texture& t, float u, float v, int id[4], float w[4]) {
    if (t.w <= 0 || t.h <= 0 || (!t.f32 && t.rgba.empty()) || (t.f32 && t.rgbaf.empty())) return false;
    if (!std::isfinite(u)) u = 0.0f;   // guard: the reference indexes out of bounds here
    if (!std::isfinite(v)) v = 0.0f;
    const float W = (float)t.w, H = (float)t.h;
    float upage = std::floor(u - 0.5f / W);
    float vpage = std::floor(v - 0.5f / H);
    u -= upage; v -= vpage;
    u *= W; v *= H;
    int iu = (int)u, iv = (int)v;
    float kui = u - (float)iu, kvi = v - (float)iv;
    float ku = 1.0f - kui, kv = 1.0f - kvi;
    int l = iu % t.w, r = (iu + 1) % t.w, b = iv % t.h, tt = (iv + 1) % t.h;
    id[0] = b * t.w + l; id[1] = b * t.w + r; id[2] = tt * t.w + l; id[3] = tt * t.w + r;
    w[0] = ku * kv; w[1] = kui * kv; w[2] = ku * kvi; w[3] = kui * kvi;
    for (int i = 0; i < 4; ++i) id[i] = std::min(std::max(id[i], 0), t.w * t.h - 1);
    return true;
}

/* BitmapTexture::texture (BitmapTexture.cpp:210-234).  The weighted sum is
 * formed in R (float for the float oracle, double for the FP64 oracle). */
template <class R>
static void texLookup(const Texture& t, float u, float v, R out[4]) {
    int id[4];
    float w[4];
    if (!texAddr(t, u, v, id, w)) {
        out[0] = R(0.5f); out[1] = R(0.5f); out[2] = R(0.5f); out[3] = R(1.0f);
        return;
    }
    if (!t.f32) {
        const R s = R(1.0f / 255.0f);
        for (int i = 0; i < 4; ++i)
            out[i] = R(t.rgba[(size_t)id[0] * 4 + i]) * R(w[0]) * s + R(t.rgba[(size_t)id[1] * 4 + i]) * R(w[1]) * s +
                     R(t.rgba[(size_t)id[2] * 4 + i]) * R(w[2]) * s + R(t.rgba[(size_t)id[3] * 4 + i]) * R(w[3]) * s;
    } else {
        for (int i = 0; i < 4; ++i)
            out[i] = R(t.rgbaf[(size_t)id[0] * 4 + i]) * R(w[0]) + R(t.rgbaf[(size_t)id[1] * 4 + i]) * R(w[1]) +
                     R(t.rgbaf[(size_t)id[2] * 4 + i]) * R(w[2]) + R(t.rgbaf[(size_t)id[3] * 4 + i]) * R(w[3]);
    }
}

/* ------------------------------------------------------- surface point */
template <class R>
struct SurfacePoint {
    int inst = -1, mesh = -1, face = -1, globalFace = -1;
    V3<R> position, normal, tangent, binormal, bary;
    R tw = R(1);          // tangent handedness (m_tangent[3])
    bool isTexNorm = false;
    R density = R(0), gradient = R(0);
};

/* ------------------------------------------------------------ BVH (oracle) */
struct Box { float mn[3], mx[3]; };
static Box emptyBox() { Box b; for (int i = 0; i < 3; ++i) { b.mn[i] = INFINITY; b.mx[i] = -INFINITY; } return b; }
static void grow(Box& b, const Box& o) { for (int i = 0; i < 3; ++i) { b.mn[i] = std::min(b.mn[i], o.mn[i]); b.mx[i] = std::max(b.mx[i], o.mx[i]); } }
static float area(const Box& b) {
    float d[3];
    for (int i = 0; i < 3; ++i) d[i] = std::max(0.0f, b.mx[i] - b.mn[i]);
    return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

/* Binned-SAH binary BVH, leaf <= 4 primitives (the canonical tree the
 * algorithmic-bytes model counts against). */
static void buildBvh(Bvh& bvh, const std::vector<Box>& boxes) {
    const int n = (int)boxes.size();
    bvh.nodes.clear(); bvh.prims.resize(n);
    for (int i = 0; i < n; ++i) bvh.prims[i] = i;
    if (n == 0) return;
    std::vector<float> cen((size_t)n * 3);
    for (int i = 0; i < n; ++i) for (int a = 0; a < 3; ++a) cen[(size_t)i * 3 + a] = 0.5f * (boxes[i].mn[a] + boxes[i].mx[a]);
    struct Job { int node, start, count; };
    std::vector<Job> stack;
    bvh.nodes.push_back(BNode());
    stack.push_back({0, 0, n});
    const int NB = 16;
    while (!stack.empty()) {
        Job j = stack.back(); stack.pop_back();
        Box bb = emptyBox(), cb = emptyBox();
        for (int i = j.start; i < j.start + j.count; ++i) {
            int p = bvh.prims[i];
            grow(bb, boxes[p]);
            Box c; for (int a = 0; a < 3; ++a) { c.mn[a] = c.mx[a] = cen[(size_t)p * 3 + a]; }
            grow(cb, c);
        }
        BNode& nd = bvh.nodes[j.node];
        for (int a = 0; a < 3; ++a) { nd.bmin[a] = bb.mn[a]; nd.bmax[a] = bb.mx[a]; }
        nd.left = nd.right = -1; nd.start = j.start; nd.count = j.count;
        if (j.count <= 4) continue;
        int bestAxis = -1, bestSplit = -1; float bestCost = INFINITY;
        for (int a = 0; a < 3; ++a) {
            float lo = cb.mn[a], hi = cb.mx[a];
            if (!(hi > lo)) continue;
            Box bins[NB]; int cnt[NB];
            for (int b = 0; b < NB; ++b) { bins[b] = emptyBox(); cnt[b] = 0; }
            float sc = NB / (hi - lo);
            for (int i = j.start; i < j.start + j.count; ++i) {
                int p = bvh.prims[i];
                int b = (int)((cen[(size_t)p * 3 + a] - lo) * sc); b = std::min(std::max(b, 0), NB - 1);
                grow(bins[b], boxes[p]); cnt[b]++;
            }
            float ra[NB]; int rc[NB]; Box acc = emptyBox(); int c = 0;
            for (int b = NB - 1; b > 0; --b) { grow(acc, bins[b]); c += cnt[b]; ra[b] = area(acc); rc[b] = c; }
            acc = emptyBox(); c = 0;
            for (int b = 0; b < NB - 1; ++b) {
                grow(acc, bins[b]); c += cnt[b];
                if (c == 0 || rc[b + 1] == 0) continue;
                float cost = area(acc) * c + ra[b + 1] * rc[b + 1];
                if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = b; }
            }
        }
        int mid;
        if (bestAxis < 0) {
            mid = j.start + j.count / 2;
        } else {
            float lo = cb.mn[bestAxis], sc = NB / (cb.mx[bestAxis] - lo);
            auto it = std::stable_partition(bvh.prims.begin() + j.start, bvh.prims.begin() + j.start + j.count, [&](int p) {
                int b = (int)((cen[(size_t)p * 3 + bestAxis] - lo) * sc); b = std::min(std::max(b, 0), NB - 1);
                return b <= bestSplit;
            });
            mid = (int)(it - bvh.prims.begin());
            if (mid == j.start || mid == j.start + j.count) mid = j.start + j.count / 2;
        }
        int l = (int)bvh.nodes.size();
        bvh.nodes.push_back(BNode()); bvh.nodes.push_back(BNode());
        bvh.nodes[j.node].left = l; bvh.nodes[j.node].right = l + 1; bvh.nodes[j.node].count = 0;
        stack.push_back({l + 1, mid, j.start + j.count - mid});
        stack.push_back({l, j.start, mid - j.start});
    }
}

/* World -> object affine inverse, computed in double (cofactors), as the
 * canonical transform both the oracle and the GPU path use. */
static void affineInverse(const float* M, double* inv) {
    double a = M[0], b = M[1], c = M[2], d = M[4], e = M[5], f = M[6], g = M[8], h = M[9], i = M[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    double r = 1.0 / det;
    double m[9] = {A * r, -(b * i - c * h) * r, (b * f - c * e) * r,
                   B * r, (a * i - c * g) * r, -(a * f - c * d) * r,
                   C * r, -(a * h - b * g) * r, (a * e - b * d) * r};
    double t[3] = {M[3], M[7], M[11]};
    for (int row = 0; row < 3; ++row) {
        inv[row * 4 + 0] = m[row * 3 + 0]; inv[row * 4 + 1] = m[row * 3 + 1]; inv[row * 4 + 2] = m[row * 3 + 2];
        inv[row * 4 + 3] = -(m[row * 3 + 0] * t[0] + m[row * 3 + 1] * t[1] + m[row * 3 + 2] * t[2]);
    }
}

/* Planted mistakes for the pin tests (test infrastructure; 0 = none).  Each
 * bit puts one plausible slip into the oracle so that a test can show the
 * check it runs fails on it:
 *   1  rmp / mmp paint toward blue / green (the targets swapped)
 *   2  md: linear density normalisation d / max instead of the log one
 *   4  watertight test returns (v, u) instead of (u, v)
 *   8  watertight test returns t * (1 + 2^-10)
 *  16  isConnected: visible whatever the closest hit
 *  32  isConnected: any hit beyond B (up to tfar) occludes
 *  64  sg: vertex normals by the position matrix instead of the normal one */
static int gPinPlant = 0;

/* False-colour paint rules (FCSubject.h:153-482).  The argument is what a
 * material's FC branch hands to the subject: a clamped float for the
 * map / density / gradient subjects (GenericFCArgument clamps it to [0, 1]),
 * an id into the colour table for the enumerated ones, the material kind for
 * mn, a flag for in. */
template <class R> static V3<R> fcPaintClamped(int sub, R a) {
    a = clampv(a, R(0), R(1));
    V3<R> hi;
    switch (sub) {
        case OMV_FC_RMP: hi = (gPinPlant & 1) ? mk(R(0), R(0), R(1)) : mk(R(0), R(1), R(0)); break;   // RoughnessCftFC: green
        case OMV_FC_MMP: hi = (gPinPlant & 1) ? mk(R(0), R(1), R(0)) : mk(R(0), R(0), R(1)); break;   // MetallicCftFC: blue
        case OMV_FC_MD: hi = mk(R(1), R(0), R(0)); break;                                             // MeshDensityFC: red
        case OMV_FC_SG: return mk(R(1) - a, R(1) - a, R(1) - a);                                        // SurfaceGradientFC
        default: return mk(R(0), R(0), R(0));
    }
    /* lerp(lower, upper, t) = (upper - lower) * t + lower from white (UtilityFunctions.h:459-469) */
    return mk(lerpv(R(1), hi.x, a), lerpv(R(1), hi.y, a), lerpv(R(1), hi.z, a));
}
/* EnumeratedFCSubject::paint (FCSubject.h:172-177): the id's table colour,
 * black past the table; MaterialIDFC::paint (:234-245) drops red. */
template <class R> static V3<R> fcPaintEnumerated(uint32_t id, const uint32_t* table, uint32_t size, bool midStyle) {
    uint32_t c = id < size ? table[id] : 0u;
    R r = R((float)((c >> 16) & 0xFF) / 255.0f), g = R((float)((c >> 8) & 0xFF) / 255.0f), b = R((float)(c & 0xFF) / 255.0f);
    return midStyle ? mk(R(0), g, b) : mk(r, g, b);
}
/* MaterialNameFC::paint (FCSubject.h:412-418) with the dictionary of
 * RenderScene::setFalseColorSubject (RenderScene.cpp:142): "om" green,
 * "pbr" blue; anything else (the missing material passes "mm") pink. */
template <class R> static V3<R> fcPaintMaterialName(int matType) {
    if (matType == MAT_OM) return mk(R(0), R(1), R(0));
    if (matType == MAT_PBR) return mk(R(0), R(0), R(1));
    return mk(R(1000), R(0), R(1000));
}
/* InvertedNormalFC::paint (FCSubject.h:472-476) */
template <class R> static V3<R> fcPaintInverted(bool inv) {
    R f = inv ? R(1) : R(0);
    return mk(f, R(1) - f, R(0));
}

/* ---------------------------------------------- watertight ray/triangle */
template <class R> struct RayPre { int kx, ky, kz; R Sx, Sy, Sz; V3<R> o, d; };

template <class R>
static bool rayPrep(const V3<R>& o, const V3<R>& d, RayPre<R>& p) {
    R ax = std::abs(d.x), ay = std::abs(d.y), az = std::abs(d.z);
    if (!(ax > R(0) || ay > R(0) || az > R(0))) return false;   // zero / NaN direction: no hit
    int kz = 0; R m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (at(d, kz) < R(0)) std::swap(kx, ky);
    p.kx = kx; p.ky = ky; p.kz = kz;
    p.Sx = at(d, kx) / at(d, kz);
    p.Sy = at(d, ky) / at(d, kz);
    p.Sz = R(1) / at(d, kz);
    p.o = o; p.d = d;
    return true;
}

/* Woop, Benthin & Wald 2013, "Watertight Ray/Triangle Intersection".
 * Returns t and the weights (u of v1, v of v2) on hit in (tnear, tfar]. */
template <class R>
static bool triIntersect(const RayPre<R>& p, const float* v0, const float* v1, const float* v2, R tnear, R tfar,
                         R& t, R& u, R& v) {
    V3<R> A = sub(ld3<R>(v0), p.o), B = sub(ld3<R>(v1), p.o), C = sub(ld3<R>(v2), p.o);
    R Ax = at(A, p.kx) - p.Sx * at(A, p.kz), Ay = at(A, p.ky) - p.Sy * at(A, p.kz);
    R Bx = at(B, p.kx) - p.Sx * at(B, p.kz), By = at(B, p.ky) - p.Sy * at(B, p.kz);
    R Cx = at(C, p.kx) - p.Sx * at(C, p.kz), Cy = at(C, p.ky) - p.Sy * at(C, p.kz);
    R U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (sizeof(R) == 4 && (U == R(0) || V == R(0) || W == R(0))) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (R)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (R)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (R)(BxAy - ByAx);
    }
    if ((U < R(0) || V < R(0) || W < R(0)) && (U > R(0) || V > R(0) || W > R(0))) return false;
    R det = U + V + W;
    if (det == R(0)) return false;
    R Az = p.Sz * at(A, p.kz), Bz = p.Sz * at(B, p.kz), Cz = p.Sz * at(C, p.kz);
    R T = U * Az + V * Bz + W * Cz;
    R rcp = R(1) / det;
    R tt = T * rcp;
    if (!(tt > tnear && tt <= tfar)) return false;
    t = tt; u = V * rcp; v = W * rcp;
    if (gPinPlant & 4) std::swap(u, v);
    if (gPinPlant & 8) t = t * (R(1) + R(1.0 / 1024.0));
    return true;
}

template <class R> struct Hit { R t, u, v; int inst, prim; bool tie; };

/* candidate acceptance with the (t, instance, primitive) tie-break */
template <class R>
static inline void offer(Hit<R>& h, bool& found, R t, R u, R v, int inst, int prim) {
    if (!found || t < h.t) { h.t = t; h.u = u; h.v = v; h.inst = inst; h.prim = prim; h.tie = false; found = true; return; }
    if (t == h.t) {
        if (inst != h.inst || prim != h.prim) h.tie = true;
        if (inst < h.inst || (inst == h.inst && prim < h.prim)) { h.u = u; h.v = v; h.inst = inst; h.prim = prim; }
    }
}

template <class R>
static void objectRay(const Instance& in, const V3<R>& o, const V3<R>& d, V3<R>& oo, V3<R>& od) {
    if (sizeof(R) == 4) {
        const float* I = in.inv;
        oo = mk(R(I[0]) * o.x + R(I[1]) * o.y + R(I[2]) * o.z + R(I[3]), R(I[4]) * o.x + R(I[5]) * o.y + R(I[6]) * o.z + R(I[7]),
                R(I[8]) * o.x + R(I[9]) * o.y + R(I[10]) * o.z + R(I[11]));
        od = mk(R(I[0]) * d.x + R(I[1]) * d.y + R(I[2]) * d.z, R(I[4]) * d.x + R(I[5]) * d.y + R(I[6]) * d.z,
                R(I[8]) * d.x + R(I[9]) * d.y + R(I[10]) * d.z);
    } else {
        const double* I = in.invd;
        oo = mk(R(I[0] * o.x + I[1] * o.y + I[2] * o.z + I[3]), R(I[4] * o.x + I[5] * o.y + I[6] * o.z + I[7]),
                R(I[8] * o.x + I[9] * o.y + I[10] * o.z + I[11]));
        od = mk(R(I[0] * d.x + I[1] * d.y + I[2] * d.z), R(I[4] * d.x + I[5] * d.y + I[6] * d.z),
                R(I[8] * d.x + I[9] * d.y + I[10] * d.z));
    }
}

/* conservative slab test in double with a relative pad */
static inline bool boxHit(const float* bmin, const float* bmax, const double o[3], const double id[3], double tnear,
                          double tfar, double& tenter) {
    double t0 = tnear, t1 = tfar;
    for (int a = 0; a < 3; ++a) {
        double pad = 1e-6 * (std::fabs((double)bmin[a]) + std::fabs((double)bmax[a])) + 1e-30;
        double lo = ((double)bmin[a] - pad - o[a]) * id[a], hi = ((double)bmax[a] + pad - o[a]) * id[a];
        if (std::isnan(lo) || std::isnan(hi)) continue;   // origin on the slab plane with zero direction
        if (lo > hi) std::swap(lo, hi);
        t0 = lo > t0 ? lo : t0; t1 = hi < t1 ? hi : t1;
    }
    tenter = t0;
    return t0 <= t1 * (1.0 + 1e-9) + 1e-30;
}

struct Counters { uint64_t nodes = 0, tris = 0; };

template <class R>
static void traverseMesh(const Scene& S, const Instance& in, const V3<R>& wo, const V3<R>& wd, R tnear, R tfar, bool useBvh,
                         Hit<R>& h, bool& found, Counters& cnt) {
    const Mesh& m = S.meshes[in.meshId];
    V3<R> oo, od;
    objectRay(in, wo, wd, oo, od);
    RayPre<R> rp;
    if (!rayPrep(oo, od, rp)) return;
    auto test = [&](int prim) {
        ++cnt.tris;
        const uint32_t* f = &m.F[(size_t)3 * prim];
        R t, u, v;
        R lim = found ? h.t : tfar;
        if (triIntersect(rp, m.vtx(f[0]), m.vtx(f[1]), m.vtx(f[2]), tnear, lim, t, u, v)) offer(h, found, t, u, v, in.instId, prim);
    };
    if (!useBvh) { for (int p = 0; p < m.fn; ++p) test(p); return; }
    const Bvh& b = S.blas[in.meshId];
    if (b.nodes.empty()) return;
    double o[3] = {(double)oo.x, (double)oo.y, (double)oo.z};
    double idr[3] = {1.0 / (double)od.x, 1.0 / (double)od.y, 1.0 / (double)od.z};
    int st[128]; int sp = 0; st[sp++] = 0;
    while (sp) {
        const BNode& nd = b.nodes[st[--sp]];
        if (nd.count == 0) ++cnt.nodes;   // interior nodes visited (the byte model's unit)
        double te;
        if (!boxHit(nd.bmin, nd.bmax, o, idr, (double)tnear, found ? (double)h.t : (double)tfar, te)) continue;
        if (nd.count > 0) { for (int i = nd.start; i < nd.start + nd.count; ++i) test(b.prims[i]); continue; }
        if (sp + 2 > 128) { std::fprintf(stderr, "refcpu: BVH stack overflow\n"); std::abort(); }
        st[sp++] = nd.right; st[sp++] = nd.left;
    }
}

template <class R>
static bool closestHit(const Scene& S, const V3<R>& o, const V3<R>& d, R tnear, R tfar, bool useBvh, Hit<R>& h, Counters& cnt) {
    bool found = false;
    h.tie = false;
    if (!useBvh || S.tlas.nodes.empty()) {
        for (const Instance& in : S.instances) if (in.meshId >= 0 && S.meshes[in.meshId].valid) traverseMesh(S, in, o, d, tnear, tfar, useBvh, h, found, cnt);
        return found;
    }
    double od[3] = {(double)o.x, (double)o.y, (double)o.z};
    double idr[3] = {1.0 / (double)d.x, 1.0 / (double)d.y, 1.0 / (double)d.z};
    int st[128]; int sp = 0; st[sp++] = 0;
    while (sp) {
        const BNode& nd = S.tlas.nodes[st[--sp]];
        if (nd.count == 0) ++cnt.nodes;
        double te;
        if (!boxHit(nd.bmin, nd.bmax, od, idr, (double)tnear, found ? (double)h.t : (double)tfar, te)) continue;
        if (nd.count > 0) {
            for (int i = nd.start; i < nd.start + nd.count; ++i) {
                const Instance& in = S.instances[S.tlas.prims[i]];
                traverseMesh(S, in, o, d, tnear, tfar, true, h, found, cnt);
            }
            continue;
        }
        st[sp++] = nd.right; st[sp++] = nd.left;
    }
    return found;
}

/* ------------------------------------------------ surface point (A9, A16) */
/* BDPTRenderer::computeSurfacePoint (BDPTRenderer.cpp:57-73) + RenderMesh::computeSurfacePoint (RenderMesh.cpp:159-234) */
template <class R>
static void computeSurfacePoint(const Scene& S, const Hit<R>& h, SurfacePoint<R>& sp) {
    const Instance& in = S.instances[h.inst];
    const Mesh& m = S.meshes[in.meshId];
    sp.inst = h.inst; sp.mesh = in.meshId; sp.face = h.prim;
    sp.globalFace = h.prim + S.faceBase[in.meshId];
    sp.bary = mk(R(1) - h.u - h.v, h.u, h.v);
    sp.isTexNorm = false;
    const uint32_t* f = &m.F[(size_t)3 * h.prim];
    const float* v[3] = {m.vtx(f[0]), m.vtx(f[1]), m.vtx(f[2])};
    const R b0 = sp.bary.x, b1 = sp.bary.y, b2 = sp.bary.z;
    V3<R> pp, pn, pt;
    for (int i = 0; i < 3; ++i) {
        at(pp, i) = b0 * R(v[0][i]) + b1 * R(v[1][i]) + b2 * R(v[2][i]);
        at(pn, i) = b0 * R(v[0][3 + i]) + b1 * R(v[1][3 + i]) + b2 * R(v[2][3 + i]);
        at(pt, i) = b0 * R(v[0][6 + i]) + b1 * R(v[1][6 + i]) + b2 * R(v[2][6 + i]);
    }
    R pw = b0 * R(v[0][9]) + b1 * R(v[1][9]) + b2 * R(v[2][9]);
    pw = pw >= R(0) ? R(1) : R(-1);
    pn = normalizeIfNeeded3(pn);
    R tn = dot3(pn, pt);
    pt = mk(pt.x - tn * pn.x, pt.y - tn * pn.y, pt.z - tn * pn.z);
    pt = normalizeIfNeeded3(pt);
    V3<R> pbn = cross3(pn, pt);
    pbn = mk(pbn.x * pw, pbn.y * pw, pbn.z * pw);
    sp.position = mul34(in.tm, pp);
    sp.normal = normalizeIfNeeded3(mul33(in.nm, pn));
    sp.tangent = normalizeIfNeeded3(mul33(in.nm, pt));
    sp.binormal = normalizeIfNeeded3(mul33(in.nm, pbn));
    sp.tw = pw;
    if (S.fcSubject == OMV_FC_MD || S.fcSubject == OMV_FC_SG) {
        V3<R> V0 = mul34(in.tm, ld3<R>(v[0])), V1 = mul34(in.tm, ld3<R>(v[1])), V2 = mul34(in.tm, ld3<R>(v[2]));
        V3<R> tfn = cross3(sub(V1, V0), sub(V2, V0));
        sp.density = R(1) / (R(0) + length3(tfn) / R(2));
        tfn = normalize3(tfn);
        R ag = R(0);
        for (int k = 0; k < 3; ++k) {
            V3<R> tvn = normalizeIfNeeded3((gPinPlant & 64) ? mul33tm(in.tm, ld3<R>(v[k] + 3)) : mul33(in.nm, ld3<R>(v[k] + 3)));
            ag = ag + std::sqrt((R(1) - clampv(dot3(tvn, tfn), R(-1), R(1))) / R(2));
        }
        sp.gradient = std::pow(clampv(ag / R(3), R(1e-6f), R(1)), R(0.286f));
        const R mnD = R(S.minDensity), mxD = R(S.maxDensity);
        if (gPinPlant & 2) sp.density = clamp01(clampv(sp.density, mnD, mxD) / mxD);
        else sp.density = clamp01(R(1) - std::log(mxD / clampv(sp.density, mnD, mxD)) / std::log(mxD / mnD));
    }
}

/* SurfacePoint::applyTextureNormal (SurfacePoint.cpp:19-34) */
template <class R>
static void applyTextureNormal(SurfacePoint<R>& sp, const V3<R>& tn) {
    if (sp.isTexNorm) return;
    V3<R> nn = mk(tn.x * sp.tangent.x + tn.y * sp.binormal.x + tn.z * sp.normal.x,
                  tn.x * sp.tangent.y + tn.y * sp.binormal.y + tn.z * sp.normal.y,
                  tn.x * sp.tangent.z + tn.y * sp.binormal.z + tn.z * sp.normal.z);
    sp.normal = normalizeIfNeeded3(nn);
    sp.tangent = normalizeIfNeeded3(cross3(sp.binormal, sp.normal));
    V3<R> b = cross3(sp.normal, sp.tangent);
    sp.binormal = mk(sp.tw * b.x, sp.tw * b.y, sp.tw * b.z);
    sp.isTexNorm = true;
}

/* RenderMesh::getTexCoord (RenderMesh.cpp:145-157) */
template <class R>
static bool getTexCoord(const Scene& S, const SurfacePoint<R>& sp, int ch, R& u, R& v) {
    u = R(0); v = R(0);
    const Mesh& m = S.meshes[sp.mesh];
    if (sp.face < 0 || sp.face >= m.fn || ch < 0 || ch >= 8 || m.texMap[ch] < 0) return false;
    const uint32_t* f = &m.F[(size_t)3 * sp.face];
    const int off = m.texOffs[m.texMap[ch]];
    const float* a = m.vtx(f[0]) + off; const float* b = m.vtx(f[1]) + off; const float* c = m.vtx(f[2]) + off;
    u = R(a[0]) * sp.bary.x + R(b[0]) * sp.bary.y + R(c[0]) * sp.bary.z;
    v = R(a[1]) * sp.bary.x + R(b[1]) * sp.bary.y + R(c[1]) * sp.bary.z;
    return true;
}

/* ----------------------------------------------------------- warps (A37) */
template <class R> static V3<R> uniformHemisphere(Uniforms& U) {
    R phi = lerpv(R(0), R(2) * kPi<R>(), R(U.next()));
    R cth = R(U.next());
    R sth = std::sqrt(std::max(R(1) - cth * cth, R(0)));
    return mk(sth * std::cos(phi), sth * std::sin(phi), cth);
}
/* rejection sampling in [0,1)^3 (positive octant only, RandomSampler.h:75-84);
 * bounded at 256 rounds, then the fixed diagonal (reached with p ~ 0.48^256) */
template <class R> static V3<R> uniformSphere(Uniforms& U) {
    V3<R> v; R sl = R(0);
    int tries = 0;
    do {
        if (++tries > 256) { v = mk(R(1), R(1), R(1)); sl = R(3); break; }
        v.x = R(U.next()); v.y = R(U.next()); v.z = R(U.next());
        sl = dot3(v, v);
    } while (sl > R(1) || sl < feps<R>());
    R s = R(1) / std::sqrt(sl);
    return mk(s * v.x, s * v.y, s * v.z);
}
template <class R> static V3<R> ggxHemisphere(Uniforms& U, R alpha2) {
    R phi = lerpv(R(0), R(2) * kPi<R>(), R(U.next()));
    R rnd = R(U.next());
    R cth = std::sqrt((R(1) - rnd) / (R(1) - (R(1) - alpha2) * rnd));
    R sth = std::sqrt(std::max(R(1) - cth * cth, R(0)));
    return mk(sth * std::cos(phi), sth * std::sin(phi), cth);
}
template <class R> static V3<R> cosineSqrtHemisphere(Uniforms& U) {
    R phi = lerpv(R(0), R(2) * kPi<R>(), R(U.next()));
    R rnd = R(U.next());
    R sth = std::sqrt(rnd);
    return mk(sth * std::cos(phi), sth * std::sin(phi), std::sqrt(R(1) - rnd));
}
/* transformToNormalSpace (BRDFModel.h:25-34): R = ((b0*c0 + 0) + b1*c1) + b2*c2 */
template <class R> static V3<R> toBasis(const V3<R>& b0, const V3<R>& b1, const V3<R>& b2, const V3<R>& c) {
    V3<R> r = mk(b0.x * c.x + R(0), b0.y * c.x + R(0), b0.z * c.x + R(0));
    r = mk(b1.x * c.y + r.x, b1.y * c.y + r.y, b1.z * c.y + r.z);
    r = mk(b2.x * c.z + r.x, b2.y * c.z + r.y, b2.z * c.z + r.z);
    return r;
}

/* ------------------------------------------------------- BSDF (A31-A33) */
template <class R> static R ggxPdf(R c, R alpha) {
    R c2 = c * c, a2 = alpha * alpha;
    return a2 / (kPi<R>() * std::pow(lerpv(R(1), a2, c2), R(2)));
}
template <class R> static R ggxLambda(R a) { return R(0.5f) * (std::sqrt(R(1) + std::pow(a, R(-2))) - R(1)); }
template <class R> static R smithG1(R c, R alpha) {
    R a = c / (alpha * std::sqrt(R(1) - c * c));
    return (alpha > R(0) || c < R(0.99f)) ? R(1) / (R(1) + ggxLambda(a)) : R(1);
}
template <class R> static R smithG2(R s, R m) { R p = s * m; return p / (s + m - p); }

template <class R> struct Bsdf {
    V3<R> I, O, N, H, T;
    R color[4];
    bool transmissive;
    R metal, rough, alpha, eta;
    int lobe;   /* lobe sampleIndirect took: 0 none, 1 diffuse, 2 reflection, 3 transmission */
    Bsdf() { I = O = N = H = T = mk(R(0), R(0), R(0)); color[0] = color[1] = color[2] = color[3] = R(1);
             transmissive = false; metal = R(0.5f); rough = R(0); alpha = R(0.25f); eta = R(1); lobe = 0; }
};

/* BSDFModel::evaluate (BRDFModel.h:139-207) */
template <class R>
static V3<R> bsdfEvaluate(const V3<R>& vIn, const V3<R>& vOut, const V3<R>& N, const R* color, R metal, R alpha,
                          bool specular, bool transmitted, const V3<R>& Hin, R P) {
    V3<R> H;
    if (!transmitted) H = normalize3(sub(vOut, vIn)); else H = Hin;
    R VdotN = clampv(std::abs(dot3(vIn, N)), R(0.0001f), R(1));
    R LdotN = clampv(std::abs(dot3(vOut, N)), R(0.0001f), R(1));
    R NdotH = std::abs(dot3(N, H));
    R VdotH = std::abs(dot3(vIn, H));
    R LdotH = std::abs(dot3(vOut, H));
    R cd[3], F0[3], F[3], fd[3], fs[3];
    R p5 = std::pow(clamp01(R(1) - VdotH), R(5));
    for (int i = 0; i < 3; ++i) {
        cd[i] = (R(1) - metal) * color[i];
        F0[i] = lerpv(R(0.04f), color[i], metal);
        F[i] = lerpv(p5, R(1), F0[i]);
        fd[i] = (R(1) - F[i]) * cd[i] / kPi<R>();
    }
    R G = smithG2(smithG1(VdotN, alpha), smithG1(LdotN, alpha));
    for (int i = 0; i < 3; ++i)
        fs[i] = clampv(R(0.25f) * (transmitted ? R(1) - F[i] : F[i]) * G / (VdotN * LdotN), R(0), std::numeric_limits<R>::max());
    R k;
    V3<R> out;
    if (specular) { k = R(4) * LdotH * VdotN / NdotH / LdotN / (R(1) - P); out = mk(k * fs[0], k * fs[1], k * fs[2]); }
    else if (transmitted) { k = R(4) * LdotH * VdotN / NdotH / LdotN / P; out = mk(k * fs[0], k * fs[1], k * fs[2]); }
    else { k = kPi<R>() / LdotN / P; out = mk(k * fd[0], k * fd[1], k * fd[2]); }
    return out;
}

/* BSDFModel::evaluateDirect (BRDFModel.h:211-261) */
template <class R>
static V3<R> evaluateDirect(const Bsdf<R>& d) {
    R VdotN = -dot3(d.I, d.N), LdotN = dot3(d.O, d.N);
    V3<R> H = normalize3(sub(d.O, d.I));
    R VdotH = std::abs(dot3(d.I, H)), NdotH = std::abs(dot3(d.N, H));
    R p5 = std::pow(clamp01(R(1) - VdotH), R(5));
    R F[3], cd[3];
    for (int i = 0; i < 3; ++i) { R F0 = lerpv(R(0.04f), d.color[i], d.metal); F[i] = lerpv(p5, R(1), F0); cd[i] = (R(1) - d.metal) * d.color[i]; }
    R fd[3] = {0, 0, 0}, fs[3] = {0, 0, 0};
    if (VdotN > R(0) && LdotN > R(0)) {
        for (int i = 0; i < 3; ++i) fd[i] = (R(1) - F[i]) * cd[i] / kPi<R>();
        R D = ggxPdf(clamp01(NdotH), d.alpha);
        R G = smithG2(smithG1(VdotN, d.alpha), smithG1(LdotN, d.alpha));
        for (int i = 0; i < 3; ++i) fs[i] = clampv(R(0.25f) * F[i] * D * G / (VdotN * LdotN), R(0), std::numeric_limits<R>::max());
    }
    return mk(fd[0] + fs[0], fd[1] + fs[1], fd[2] + fs[2]);
}

/* BSDFModel::sampleIndirect (BRDFModel.h:265-337); O keeps the reference's
 * value on the zero-weight early returns. */
template <class R>
static V3<R> sampleIndirect(Bsdf<R>& d, Uniforms& U) {
    R VdotN = -dot3(d.I, d.N);
    V3<R> binorm = cross3(d.N, d.T);
    V3<R> b0 = d.T, b1 = binorm;
    if (VdotN < R(0)) {
        if (d.transmissive) { std::swap(b0, b1); d.N = neg(d.N); }
        else return mk(R(0), R(0), R(0));
    }
    const R P = R(0.8f) * std::abs(VdotN) + R(0.1f);
    R rnd = R(U.next());
    V3<R> z = mk(R(0), R(0), R(0));
    if (rnd < P) {
        if (rnd < P * (R(1) - d.color[3])) {
            if (VdotN < R(0)) d.eta = R(1) / d.eta;
            V3<R> rd = ggxHemisphere(U, std::pow(d.alpha, R(2)));
            d.H = normalize3(toBasis(b0, b1, d.N, rd));
            if (refract3(d.I, d.H, d.eta, d.O) && dot3(d.N, d.O) < R(0)) {
                d.O = normalize3(d.O);
                d.lobe = 3;
                return bsdfEvaluate(d.I, d.O, d.N, d.color, d.metal, d.alpha, false, true, d.H, P);
            }
            return z;
        }
        V3<R> rd = cosineSqrtHemisphere<R>(U);
        d.O = normalize3(toBasis(b0, b1, d.N, rd));
        d.lobe = 1;
        return bsdfEvaluate(d.I, d.O, d.N, d.color, d.metal, d.alpha, false, false, d.H, P);
    }
    V3<R> rd = ggxHemisphere(U, std::pow(d.alpha, R(2)));
    d.H = normalize3(toBasis(b0, b1, d.N, rd));
    d.O = reflect3(d.I, d.H);
    if (dot3(d.N, d.O) > R(0)) {
        d.O = normalize3(d.O);
        d.lobe = 2;
        return bsdfEvaluate(d.I, d.O, d.N, d.color, d.metal, d.alpha, true, false, d.H, P);
    }
    return z;
}

/* ------------------------------------------- detached lobe values (adjoint)
 * With the sampled lobe and direction held fixed, the estimator's weight of
 * a PBR bounce is w_ref(theta0) * f(theta) / f(theta0), where f is the
 * sampled lobe's integrand up to factors that do not depend on the material
 * (the reference's weight times its sampling density): diffuse
 * (1-F)(1-metal)color, reflection F*D*G, transmission (1-F)*D*G, with the
 * geometry bsdfEvaluate uses.  The gradient of the bounce is the weight
 * times d ln f / d(color_c, metal, alpha). */
template <class R> struct LobeGeo { int lobe; R p5, NH, VN, LN; };

template <class R>
static LobeGeo<R> lobeGeo(const Bsdf<R>& d) {
    LobeGeo<R> g;
    g.lobe = d.lobe;
    V3<R> H = d.lobe == 3 ? d.H : normalize3(sub(d.O, d.I));
    g.VN = clampv(std::abs(dot3(d.I, d.N)), R(0.0001f), R(1));
    g.LN = clampv(std::abs(dot3(d.O, d.N)), R(0.0001f), R(1));
    g.NH = std::abs(dot3(d.N, H));
    R VH = std::abs(dot3(d.I, H));
    g.p5 = std::pow(clamp01(R(1) - VH), R(5));
    return g;
}

template <class R>
static void lobeValue(const LobeGeo<R>& g, const R color[3], R metal, R alpha, R f[3]) {
    R D = R(1), G = R(1);
    if (g.lobe >= 2) { D = ggxPdf(g.NH, alpha); G = smithG2(smithG1(g.VN, alpha), smithG1(g.LN, alpha)); }
    for (int c = 0; c < 3; ++c) {
        R F0 = lerpv(R(0.04f), color[c], metal);
        R F = lerpv(g.p5, R(1), F0);
        if (g.lobe == 1) f[c] = (R(1) - F) * (R(1) - metal) * color[c];
        else if (g.lobe == 2) f[c] = F * D * G;
        else if (g.lobe == 3) f[c] = (R(1) - F) * D * G;
        else f[c] = R(0);
    }
}

/* d ln G1 / d alpha of smithG1 */
template <class R>
static R dlnSmithG1(R c, R alpha) {
    if (!(alpha > R(0) || c < R(0.99f))) return R(0);
    R a = c / (alpha * std::sqrt(R(1) - c * c));
    R ai2 = std::pow(a, R(-2));
    R lam = R(0.5f) * (std::sqrt(R(1) + ai2) - R(1));
    R dlam = ai2 / (R(2) * alpha * std::sqrt(R(1) + ai2));
    return -dlam / (R(1) + lam);
}

/* Planted mistakes for the tests that show the adjoint checks have teeth:
 * bit 0 drops d ln D / d alpha, bit 1 drops the base-colour term of F. */
static int gAdjPlant = 0;
static int gSplatPlant = 0;   /* test aid: 1 drops a cos(theta) of the pinhole Jacobian, 2 the 1/p of the emission octant */

/* d ln f_c / d(color_c, metal, alpha) for the three lobes; 0 where f_c == 0 */
template <class R>
static void lobeDlog(const LobeGeo<R>& g, const R color[3], R metal, R alpha, R dColor[3], R dMetal[3], R dAlpha[3]) {
    R dlnDG = R(0);
    if (g.lobe >= 2) {
        R c2 = g.NH * g.NH;
        R L = lerpv(R(1), alpha * alpha, c2);
        R dlnD = (gAdjPlant & 1) ? R(0) : R(2) / alpha - R(4) * alpha * c2 / L;
        R x = smithG1(g.VN, alpha), y = smithG1(g.LN, alpha);
        R dx = x * dlnSmithG1(g.VN, alpha), dy = y * dlnSmithG1(g.LN, alpha);
        R den = x + y - x * y;
        R dlnG = dx / x + dy / y - (dx + dy - dx * y - x * dy) / den;
        dlnDG = dlnD + dlnG;
    }
    R f[3];
    lobeValue(g, color, metal, alpha, f);
    for (int c = 0; c < 3; ++c) {
        dColor[c] = dMetal[c] = dAlpha[c] = R(0);
        if (!(f[c] != R(0))) continue;
        R F0 = lerpv(R(0.04f), color[c], metal);
        R F = lerpv(g.p5, R(1), F0);
        R dFdC = (gAdjPlant & 2) ? R(0) : (R(1) - g.p5) * metal, dFdM = (R(1) - g.p5) * (color[c] - R(0.04f));
        if (g.lobe == 1) {
            dColor[c] = -dFdC / (R(1) - F) + R(1) / color[c];
            dMetal[c] = -dFdM / (R(1) - F) - R(1) / (R(1) - metal);
        } else if (g.lobe == 2) {
            dColor[c] = dFdC / F; dMetal[c] = dFdM / F; dAlpha[c] = dlnDG;
        } else {
            dColor[c] = -dFdC / (R(1) - F); dMetal[c] = -dFdM / (R(1) - F); dAlpha[c] = dlnDG;
        }
    }
}

/* fresnelReflection (OpenMaterial.cpp:20-38), std::complex as the reference */
template <class R>
static void fresnel(R n, R k, R cosT, R& Rp, R& Rs) {
    typedef std::complex<R> C;
    C N(n, k);
    C N2 = N * N;
    cosT = std::abs(cosT);
    const R c2 = cosT * cosT, s2 = R(1) - c2;
    C np = N2 * cosT - std::sqrt(N2 - s2);
    C dp = N2 * cosT + std::sqrt(N2 - s2);
    C rp = np / dp;
    Rp = std::real(rp * std::conj(rp));
    C ns = cosT - std::sqrt(N2 - s2);
    C ds = cosT + std::sqrt(N2 - s2);
    C rs = ns / ds;
    Rs = std::real(rs * std::conj(rs));
}

/* d(Rp, Rs)/d(n, k) of fresnel() at fixed cos: with N2 = (n + ik)^2 and
 * u = sqrt(N2 - s^2), d rp/d N2 = c (2u^2 - N2) / (u (N2 c + u)^2),
 * d rs/d N2 = -c / (u (c + u)^2), d N2/dn = 2N, d N2/dk = 2iN, and
 * d|r|^2 = 2 Re(conj(r) dr). */
template <class R>
static void fresnelGrad(R n, R k, R cosT, R& dRpdn, R& dRpdk, R& dRsdn, R& dRsdk) {
    typedef std::complex<R> C;
    C N(n, k);
    C N2 = N * N;
    cosT = std::abs(cosT);
    const R c2 = cosT * cosT, s2 = R(1) - c2;
    C u = std::sqrt(N2 - s2);
    C a = N2 * cosT;
    C rp = (a - u) / (a + u), rs = (C(cosT) - u) / (C(cosT) + u);
    C drp = cosT * (R(2) * u * u - N2) / (u * (a + u) * (a + u));
    C drs = -cosT / (u * (C(cosT) + u) * (C(cosT) + u));
    C dn = R(2) * N, dk = C(0, 2) * N;
    dRpdn = R(2) * std::real(std::conj(rp) * drp * dn);
    dRpdk = R(2) * std::real(std::conj(rp) * drp * dk);
    dRsdn = R(2) * std::real(std::conj(rs) * drs * dn);
    dRsdk = R(2) * std::real(std::conj(rs) * drs * dk);
}

/* ---------------------------------------------------------- materials */
template <class R> struct Ctx { const Scene* S; const omv_renderer_params* P; };

/* 3x3 KHR_texture_transform applied to (u, v, 1) (PhysicallyBasedMaterial.cpp:39-45) */
template <class R>
static bool pbrTex(const Scene& S, const SurfacePoint<R>& sp, const omv_pbr_desc& d, int slot, R out[4]) {
    int id = d.texture_map[slot];
    if (!(0 <= id && id < (int)S.textures.size())) return false;
    R u, v;
    if (!getTexCoord(S, sp, d.texture_channel[slot], u, v)) return false;
    V3<R> uvT = mul33(d.texture_transform[slot], mk(u, v, R(1)));
    texLookup(S.textures[id], (float)uvT.x, (float)uvT.y, out);
    return true;
}

/* Helper::getTextureRGBA / getTextureOccMetRoughValues / occluseColor (PhysicallyBasedMaterial.cpp:32-100) */
template <class R>
static void pbrInputs(const Scene& S, const SurfacePoint<R>& sp, const omv_pbr_desc& d, R color[4], R omr[3]) {
    R tex[4];
    if (!pbrTex(S, sp, d, OMV_PBR_BASE_COLOR, color)) { color[0] = color[1] = color[2] = color[3] = R(1); }
    for (int i = 0; i < 4; ++i) color[i] = color[i] * R(d.base_color_factor[i]);
    omr[0] = R(1); omr[1] = R(d.metallic_factor); omr[2] = R(d.roughness_factor);
    if (pbrTex(S, sp, d, OMV_PBR_METALLIC_ROUGHNESS, tex)) { omr[0] *= tex[0]; omr[1] *= tex[2]; omr[2] *= tex[1]; }
    bool sep = d.texture_map[OMV_PBR_OCCLUSION] != d.texture_map[OMV_PBR_METALLIC_ROUGHNESS];
    if (sep && pbrTex(S, sp, d, OMV_PBR_OCCLUSION, tex)) omr[0] = tex[0];
}
template <class R>
static void pbrPrepare(const Scene& S, const SurfacePoint<R>& sp, const omv_pbr_desc& d, Bsdf<R>& b, R omr[3]) {
    pbrInputs(S, sp, d, b.color, omr);
    if (d.alpha_mode != OMV_ALPHA_BLEND) b.color[3] = R(1);
    b.transmissive = d.double_sided || b.color[3] < R(1);
}
template <class R>
static void pbrFinish(const omv_pbr_desc& d, Bsdf<R>& b, const R omr[3]) {
    for (int i = 0; i < 3; ++i) b.color[i] = lerpv(b.color[i], omr[0] * b.color[i], R(d.occlusion_strength));
    b.metal = omr[1];
    b.rough = clampv(omr[2], R(0.00001f), R(0.99999f));
    b.alpha = std::pow(b.rough, R(2));
}

/* tangent-space normal of a normal-map texel (PhysicallyBasedMaterial.cpp:158-164):
 * blue lifted above 0.5, 2 * rgb - 1, x and y scaled, normalized */
template <class R>
static V3<R> texelNormal(const R texel[4], R scale) {
    R tn[3] = {texel[0], texel[1], texel[2]};
    if (tn[2] <= R(0.5f)) tn[2] = R(0.5f) + feps<R>();
    for (int i = 0; i < 3; ++i) tn[i] = R(2) * tn[i] + R(-1);
    for (int i = 0; i < 2; ++i) tn[i] = scale * tn[i];
    return normalize3(mk(tn[0], tn[1], tn[2]));
}

/* PhysicallyBasedMaterial::modifyFrame (PhysicallyBasedMaterial.cpp:147-167) */
template <class R>
static void pbrModifyFrame(const Scene& S, SurfacePoint<R>& sp, const omv_pbr_desc& d) {
    R tn[4] = {R(0), R(0), R(1), R(0)};
    if (pbrTex(S, sp, d, OMV_PBR_NORMAL, tn)) {
        applyTextureNormal(sp, texelNormal(tn, R(d.normal_scale)));
        return;
    }
    applyTextureNormal(sp, mk(tn[0], tn[1], tn[2]));
}

/* emission of PhysicallyBasedMaterial::getRayAndBrdf (:282-287, :293-299, :102-123) */
template <class R>
static V3<R> pbrEmission(const Scene& S, const SurfacePoint<R>& sp, const omv_pbr_desc& d) {
    R u, v;
    getTexCoord(S, sp, d.texture_channel[OMV_PBR_EMISSIVITY], u, v);
    V3<float> ef = mk(d.emissive_factor[0], d.emissive_factor[1], d.emissive_factor[2]);
    if (!(length3(ef) > std::numeric_limits<float>::epsilon())) return mk(R(0), R(0), R(0));
    int id = d.texture_map[OMV_PBR_EMISSIVITY];
    if (!(0 <= id && id < (int)S.textures.size())) return mk(R(0), R(0), R(0));
    V3<R> uvT = mul33(d.texture_transform[OMV_PBR_EMISSIVITY], mk(u, v, R(1)));
    R tex[4];
    texLookup(S.textures[id], (float)uvT.x, (float)uvT.y, tex);
    return mk(tex[0] * R(ef.x), tex[1] * R(ef.y), tex[2] * R(ef.z));
}

/* PhysicallyBasedMaterial::isMasked (:301-316) */
template <class R>
static bool isMasked(const Scene& S, const SurfacePoint<R>& sp) {
    const Mesh& m = S.meshes[sp.mesh];
    const Material& mat = S.materials[m.matId];
    if (mat.type != MAT_PBR || mat.pbr.alpha_mode != OMV_ALPHA_MASK) return false;
    float alpha = mat.pbr.base_color_factor[3];
    if (alpha < mat.pbr.alpha_cutoff) return true;
    int id = mat.pbr.texture_map[OMV_PBR_BASE_COLOR];
    if (0 <= id && id < (int)S.textures.size()) {
        R u, v, tc[4];
        if (!getTexCoord(S, sp, mat.pbr.texture_channel[OMV_PBR_BASE_COLOR], u, v)) { tc[3] = R(1); }
        else texLookup(S.textures[id], (float)u, (float)v, tc);
        alpha *= (float)tc[3];
    }
    return alpha < mat.pbr.alpha_cutoff;
}

/* geometric normal of the hit face through the normal matrix (OpenMaterial.cpp:76-86) */
template <class R>
static V3<R> faceNormalOf(const V3<R>& P0, const V3<R>& P1, const V3<R>& P2, const float* nm) {
    V3<R> fn = cross3(sub(P1, P0), sub(P2, P0));
    return normalize3(mul33(nm, fn));
}
template <class R>
static V3<R> faceNormal(const Scene& S, const SurfacePoint<R>& sp) {
    const Mesh& m = S.meshes[sp.mesh];
    const uint32_t* f = &m.F[(size_t)3 * sp.face];
    return faceNormalOf(ld3<R>(m.vtx(f[0])), ld3<R>(m.vtx(f[1])), ld3<R>(m.vtx(f[2])), S.instances[sp.inst].nm);
}

/* OpenMaterial::defineNextDirection (OpenMaterial.cpp:72-93): mirror
 * direction about the shading normal, lifted to 0.1 above the geometric face
 * normal tfn */
template <class R>
static V3<R> omNextDirectionOf(const V3<R>& inV, const V3<R>& normal, const V3<R>& tfn) {
    V3<R> out = reflect3(inV, normal);
    R corr = dot3(out, tfn);
    if (corr < R(0.1f)) {
        R k = R(0.1f) - corr;
        out = normalize3(mk(tfn.x * k + out.x, tfn.y * k + out.y, tfn.z * k + out.z));
    }
    return out;
}
template <class R>
static V3<R> omNextDirection(const Scene& S, const V3<R>& inV, const SurfacePoint<R>& sp) {
    return omNextDirectionOf(inV, sp.normal, faceNormal(S, sp));
}
template <class R>
static V3<R> omFresnelBrdf(const Material& m, R fCos) {
    V3<R> b;
    for (int c = 0; c < 3; ++c) {
        R rp, rs;
        fresnel(R(m.n[c]), R(m.k[c]), fCos, rp, rs);
        at(b, c) = lerpv(rp, rs, R(0.5f)) / std::abs(fCos);
    }
    return b;
}

/* Estimator extensions selected by omv_renderer_params.extensions. */
struct Ext {
    bool omRough = false;   // OMV_EXT_OM_ROUGH
    float omAlpha = 0.09f;  // GGX alpha = roughness^2
    bool nee = false;       // OMV_EXT_HDRI_NEE
    bool splat = false;     // OMV_EXT_LIGHT_SPLAT
    explicit Ext(const omv_renderer_params& p) {
        omRough = (p.extensions & OMV_EXT_OM_ROUGH) != 0;
        omAlpha = p.om_roughness * p.om_roughness;
        nee = (p.extensions & OMV_EXT_HDRI_NEE) != 0;
        splat = (p.extensions & OMV_EXT_LIGHT_SPLAT) != 0;
    }
};

/* OMV_EXT_OM_ROUGH, connection: GGX rough conductor whose F is the complex-IOR
 * Fresnel of fresnelReflection at |I.H|; f = F D G2 / (4 V.N L.N), zero unless
 * both cosines are positive (the evaluateDirect convention). */
template <class R>
static V3<R> omRoughEval(const Material& m, R alpha, const V3<R>& I, const V3<R>& O, const V3<R>& N) {
    R VdotN = -dot3(I, N), LdotN = dot3(O, N);
    if (!(VdotN > R(0) && LdotN > R(0))) return mk(R(0), R(0), R(0));
    V3<R> H = normalize3(sub(O, I));
    R VdotH = std::abs(dot3(I, H)), NdotH = std::abs(dot3(N, H));
    R D = ggxPdf(clamp01(NdotH), alpha);
    R G = smithG2(smithG1(VdotN, alpha), smithG1(LdotN, alpha));
    R k = D * G / (R(4) * VdotN * LdotN);
    V3<R> out;
    for (int c = 0; c < 3; ++c) {
        R rp, rs;
        fresnel(R(m.n[c]), R(m.k[c]), VdotH, rp, rs);
        at(out, c) = lerpv(rp, rs, R(0.5f)) * k;
    }
    return out;
}

/* OMV_EXT_OM_ROUGH, sampling: half vector from GGXHemisphere (2 uniforms),
 * O = reflect(I, H); the returned brdf times the path's cosine (|O.N| on
 * camera paths, |I.N| on light paths) is F G |I.H| / (|I.N| |N.H|). */
template <class R>
static V3<R> omRoughSample(const Material& m, R alpha, const V3<R>& I, const SurfacePoint<R>& sp, bool inv, V3<R>& O,
                           Uniforms& U) {
    const V3<R>& N = sp.normal;
    const V3<R> zero = mk(R(0), R(0), R(0));
    O = reflect3(I, N);
    R VdotN = -dot3(I, N);
    if (!(VdotN > R(0))) return zero;
    V3<R> B = cross3(N, sp.tangent);
    V3<R> rd = ggxHemisphere(U, alpha * alpha);
    V3<R> H = normalize3(toBasis(sp.tangent, B, N, rd));
    V3<R> o = reflect3(I, H);
    if (!(dot3(N, o) > R(0))) return zero;
    O = normalize3(o);
    R LdotN = dot3(O, N), VdotH = std::abs(dot3(I, H)), NdotH = std::abs(dot3(N, H));
    R G = smithG2(smithG1(VdotN, alpha), smithG1(LdotN, alpha));
    R w = G * VdotH / (VdotN * NdotH);
    R ct = inv ? LdotN : VdotN;
    V3<R> out;
    for (int c = 0; c < 3; ++c) {
        R rp, rs;
        fresnel(R(m.n[c]), R(m.k[c]), VdotH, rp, rs);
        at(out, c) = lerpv(rp, rs, R(0.5f)) * w / ct;
    }
    return out;
}

/* RenderMaterial::getBrdf dispatch for a connection (BDPTRenderer.cpp:287-289) */
template <class R>
static V3<R> matGetBrdf(const Scene& S, const Ext& X, const SurfacePoint<R>& sp, const V3<R>& inV, const V3<R>& outV) {
    const Material& m = S.materials[S.meshes[sp.mesh].matId];
    if (m.type == MAT_OM) {
        if (!m.omHasBase) return mk(R(0), R(0), R(0));
        if (X.omRough) return omRoughEval<R>(m, R(X.omAlpha), inV, outV, sp.normal);
        R fCos = -dot3(inV, sp.normal);
        if (std::abs(dot3(outV, sp.normal) - fCos) > R(1e-6f)) return mk(R(0), R(0), R(0));
        return omFresnelBrdf<R>(m, fCos);
    }
    if (m.type == MAT_PBR) {
        Bsdf<R> b; R omr[3];
        b.I = inV; b.O = outV; b.N = sp.normal;
        pbrPrepare(S, sp, m.pbr, b, omr);
        if (dot3(inV, sp.normal) > R(0) && b.transmissive) b.N = neg(b.N);
        pbrFinish(m.pbr, b, omr);
        return evaluateDirect(b);
    }
    /* DiffuseColorMaterial::getBrdf (DiffuseColorMaterial.cpp:89-104) */
    const R pdf = R(1) / kPi<R>();
    const R brdfPdf = R(1) / (pdf * kPi<R>());
    return mk(brdfPdf * R(m.diffuse[0]), brdfPdf * R(m.diffuse[1]), brdfPdf * R(m.diffuse[2]));
}

template <class R>
static void modifyFrame(const Scene& S, SurfacePoint<R>& sp) {
    const Material& m = S.materials[S.meshes[sp.mesh].matId];
    if (m.type == MAT_PBR) pbrModifyFrame(S, sp, m.pbr);
    else applyTextureNormal(sp, mk(R(0), R(0), R(1)));
}

/* What the adjoint needs from one PBR bounce: the sampled Bsdf state and the
 * material inputs before occlusion. */
template <class R> struct BounceInfo {
    int kind = 0;            /* 0 other, 1 PBR, 2 OM mirror */
    Bsdf<R> b;
    R colorPre[3] = {R(0), R(0), R(0)}, omr[3] = {R(0), R(0), R(0)};
    R fCos = R(0);
};

/* RenderMaterial::getRayAndBrdf dispatch (computePath, BDPTRenderer.cpp:158-159).
 * E (optional): the scene whose material parameters the weight is evaluated
 * with, sampling staying with S ("detached" sampling, used for finite
 * differences of the adjoint): PBR weights are scaled by f(E)/f(S) of the
 * sampled lobe, OM mirror weights use E's (n, k). */
template <class R>
static void matGetRayAndBrdf(const Scene& S, const Ext& X, const SurfacePoint<R>& sp, const V3<R>& inV, V3<R>& outV,
                             V3<R>& brdf, V3<R>& erad, bool inv, Uniforms& U, const Scene* E = nullptr,
                             BounceInfo<R>* info = nullptr) {
    const Material& m = S.materials[S.meshes[sp.mesh].matId];
    erad = mk(R(0), R(0), R(0));
    if (m.type == MAT_OM) {
        if (X.omRough && m.omHasBase) { brdf = omRoughSample<R>(m, R(X.omAlpha), inV, sp, inv, outV, U); return; }
        outV = omNextDirection(S, inV, sp);
        if (!m.omHasBase) { brdf = mk(R(0), R(0), R(0)); return; }
        const Material& me = E ? E->materials[E->meshes[sp.mesh].matId] : m;
        brdf = omFresnelBrdf<R>(me, dot3(outV, sp.normal));
        if (info) { info->kind = 2; info->fCos = dot3(outV, sp.normal); }
        return;
    }
    if (m.type == MAT_PBR) {
        Bsdf<R> b; R omr[3];
        b.I = inV; b.N = sp.normal; b.T = sp.tangent;
        pbrPrepare(S, sp, m.pbr, b, omr);
        R pre[3] = {b.color[0], b.color[1], b.color[2]};
        pbrFinish(m.pbr, b, omr);
        b.eta = R(1) / (m.pbr.double_sided ? R(1) : R(m.pbr.ior));
        brdf = sampleIndirect(b, U);
        outV = b.O;
        if (m.pbrEmissive) erad = pbrEmission(S, sp, m.pbr);
        if (E && b.lobe != 0) {
            const Material& me = E->materials[E->meshes[sp.mesh].matId];
            Bsdf<R> be; R omre[3];
            pbrPrepare(*E, sp, me.pbr, be, omre);
            pbrFinish(me.pbr, be, omre);
            LobeGeo<R> g = lobeGeo(b);
            R f0[3], f1[3];
            lobeValue(g, b.color, b.metal, b.alpha, f0);
            lobeValue(g, be.color, be.metal, be.alpha, f1);
            for (int c = 0; c < 3; ++c) at(brdf, c) = f0[c] != R(0) ? at(brdf, c) * (f1[c] / f0[c]) : R(0);
        }
        if (info) {
            info->kind = 1; info->b = b;
            for (int c = 0; c < 3; ++c) { info->colorPre[c] = pre[c]; info->omr[c] = omr[c]; }
        }
        return;
    }
    /* DiffuseColorMaterial via RenderMaterial::getRayAndBrdf (RenderMaterial.cpp:51-72) */
    V3<R> v = uniformHemisphere<R>(U);
    outV = toBasis(sp.tangent, sp.binormal, sp.normal, v);
    if (!inv) brdf = matGetBrdf(S, X, sp, inV, outV);
    else brdf = matGetBrdf(S, X, sp, neg(outV), neg(inV));
}

/* ------------------------------------------------------ renderer (A4-A12) */
template <class R>
struct Vtx {
    V3<R> pos, outRay, outRad, outFactor;
    V3<R> w;   /* sampled weight of the bounce (getRayAndBrdf's brdf) */
    SurfacePoint<R> sp;
};

template <class R>
struct Tracer {
    const Scene& S;
    const omv_renderer_params& P;
    const Ext X;
    const Scene* E = nullptr;   /* detached evaluation scene (see matGetRayAndBrdf) */
    uint64_t rays = 0, camRays = 0, lightRays = 0, shadowRays = 0;
    Counters cnt;
    Tracer(const Scene& s, const omv_renderer_params& p) : S(s), P(p), X(p) {}
    /* false-colour table of an enumerated subject, first instance-face index per instance (ipid) */
    std::vector<uint32_t> fcTable, fcInstBase;

    /* sceneIntersect (BDPTRenderer.cpp:75-102); tnear persists in the caller */
    bool sceneIntersect(const V3<R>& org, const V3<R>& dir, R& tnear, R tfar, SurfacePoint<R>& sp) {
        const R t1 = tfar;
        int itr = 10;
        while (itr > 0) {
            Hit<R> h;
            ++rays;
            if (!closestHit(S, org, dir, tnear, tfar, true, h, cnt)) return false;
            computeSurfacePoint(S, h, sp);
            if (!isMasked(S, sp)) return true;
            V3<R> o2 = mk(org.x + h.t * dir.x, org.y + h.t * dir.y, org.z + h.t * dir.z);
            R v = maxAbs(o2);
            v = v > h.t ? v : h.t;
            R delta = R(3) * feps<R>() * v;
            tnear = h.t + delta;
            tfar = t1;
            if (tfar <= tnear) return false;
            --itr;
        }
        return false;
    }

    /* computeLightAttenuation (BDPTRenderer.cpp:104-126) */
    R attenuation(R d, R attD) {
        if (attD >= R(std::numeric_limits<float>::max())) {
            d = d < R(P.light_min_distance) ? R(P.light_min_distance) : d;
            switch (P.light_distance_attenuation) { case 1: return R(1) / d; case 2: return R(1) / (d * d); default: return R(1); }
        }
        R rD = mx(R(1) - d / attD, R(0));
        switch (P.light_distance_attenuation) { case 1: return rD; case 2: return std::sqrt(rD); default: return d < attD ? R(1) : R(0); }
    }

    /* ---- OMV_EXT_HDRI_NEE ------------------------------------------- */
    R lastEscW = R(1);   /* MIS weight of an escape after the last camera vertex */

    /* inverse of HDRBackground::getRadiance's mapping for a point (ux, uy) of
     * texel (row, col); sinT = sine of the polar angle */
    V3<R> envDirection(int row, int col, float ux, float uy, R& sinT) const {
        const R pi = kPi<R>();
        const R u = (R(col) + R(ux)) / R(S.bgW), v = (R(row) + R(uy)) / R(S.bgH);
        const R yang = (R(1) - v) * pi;
        const R rz = -std::cos(yang);
        const R xang = u * R(2) * pi;
        const R r = std::sqrt(mx(R(1) - rz * rz, R(0)));
        sinT = r;
        return mul33<R>(S.bgRot, mk(r * std::cos(xang), r * std::sin(xang), rz));
    }
    /* solid-angle density of the environment sampler at dir */
    R envPdf(const V3<R>& dir) const {
        if (S.envTexPdf.empty()) return R(0);
        const R pi = kPi<R>();
        V3<R> rdir = mulT33(dir, S.bgRot);
        V3<R> rxy = normalize3(mk(rdir.x, rdir.y, R(0)));
        R xang = std::acos(clampv(rxy.x, R(-1), R(1)));
        xang = rxy.y < R(0) ? R(2) * pi - xang : xang;
        const R z = clampv(rdir.z, R(-1), R(1));
        const R yang = std::asin(z) + R(0.5f) * pi;
        const R u = xang / (R(2) * pi), v = R(1) - yang / pi;
        const int W = S.bgW, H = S.bgH;
        const int col = std::min(std::max((int)(u * R(W)), 0), W - 1), row = std::min(std::max((int)(v * R(H)), 0), H - 1);
        const R sinT = std::sqrt(mx(R(1) - z * z, R(0)));
        if (!(sinT > R(0))) return R(0);
        return R(S.envTexPdf[(size_t)row * W + col]) * R(W) * R(H) / (R(2) * pi * pi * sinT);
    }
    /* density of the reference's BSDF sampler (sampleIndirect's lobe mixture)
     * of direction O at an opaque PBR vertex */
    R bsdfPdf(const Bsdf<R>& b, const V3<R>& O) const {
        const R VN = -dot3(b.I, b.N);
        if (VN < R(0)) return R(0);
        const R P = R(0.8f) * std::abs(VN) + R(0.1f);
        const R LN = dot3(O, b.N);
        if (!(LN > R(0))) return R(0);
        R p = P * LN / kPi<R>();
        const V3<R> H = normalize3(sub(O, b.I));
        const R NH = dot3(b.N, H), LH = dot3(O, H);
        if (NH > R(0) && LH > R(0)) p += (R(1) - P) * ggxPdf(NH, b.alpha) * NH / (R(4) * LH);
        return p;
    }
    /* the reference estimator's integrand per solid angle (weight x density x
     * cosine of sampleIndirect, clamps included) of direction O */
    V3<R> bsdfEffCos(const Bsdf<R>& b, const V3<R>& O) const {
        const V3<R> H = normalize3(sub(O, b.I));
        const R VNc = clampv(std::abs(dot3(b.I, b.N)), R(0.0001f), R(1));
        const R LNc = clampv(std::abs(dot3(O, b.N)), R(0.0001f), R(1));
        const R LN = std::abs(dot3(O, b.N));
        const R NH = std::abs(dot3(b.N, H)), VH = std::abs(dot3(b.I, H));
        const R p5 = std::pow(clamp01(R(1) - VH), R(5));
        const R D = ggxPdf(NH, b.alpha);
        const R G = smithG2(smithG1(VNc, b.alpha), smithG1(LNc, b.alpha));
        V3<R> out;
        for (int c = 0; c < 3; ++c) {
            const R F0 = lerpv(R(0.04f), b.color[c], b.metal);
            const R F = lerpv(p5, R(1), F0);
            const R fd = (R(1) - F) * ((R(1) - b.metal) * b.color[c]) / kPi<R>();
            at(out, c) = fd * LN * LN / LNc + F * G * D * LN / (R(4) * LNc * LNc);
        }
        return out;
    }
    /* next-event estimate toward the environment at camera vertex i, and the
     * MIS weight of the escape along the sampled direction outV */
    V3<R> neeEstimate(const SurfacePoint<R>& sp, const BounceInfo<R>& bi, const V3<R>& ofPrev, const V3<R>& outV, int i,
                      uint64_t sample) {
        V3<R> z = mk(R(0), R(0), R(0));
        lastEscW = R(1);
        if (!X.nee || S.envTexPdf.empty() || bi.kind != 1) return z;
        const Bsdf<R>& b = bi.b;
        /* opaque, front-facing (sampleIndirect returns zero for V.N < 0), and
         * an escape from this vertex must still be counted (n <= q) */
        if (b.transmissive || -dot3(b.I, b.N) < R(0) || i >= P.camera_bounces || i + 1 > P.max_path_length) return z;
        /* MIS weight of the escape along the BSDF-sampled direction */
        const R pbS = bsdfPdf(b, outV), peS = envPdf(outV);
        if (pbS + peS > R(0)) lastEscW = pbS / (pbS + peS);
        /* environment sample: row, column (alias tables), point in the texel */
        Stream ns(P.seed, STREAM_NEE + 16u * (uint32_t)i, sample);
        const float u0 = ns.next(), u1 = ns.next(), u2 = ns.next(), u3 = ns.next(), u4 = ns.next(), u5 = ns.next();
        const int W = S.bgW, H = S.bgH;
        int k = std::min((int)(u0 * (float)H), H - 1);
        const int row = u1 < S.envRowProb[k] ? k : S.envRowAlias[k];
        k = std::min((int)(u2 * (float)W), W - 1);
        const size_t rk = (size_t)row * W + k;
        const int col = u3 < S.envColProb[rk] ? k : S.envColAlias[rk];
        R sinT;
        const V3<R> w = envDirection(row, col, u4, u5, sinT);
        if (!(sinT > R(0))) return z;
        const R pe = R(S.envTexPdf[(size_t)row * W + col]) * R(W) * R(H) / (R(2) * kPi<R>() * kPi<R>() * sinT);
        const R pb = bsdfPdf(b, w);
        if (!(pe > R(0)) || !(dot3(w, b.N) > R(0))) return z;
        const V3<R> f = bsdfEffCos(b, w);
        const V3<R> bg = background(w);
        const R inv = R(1) / (pe + pb);
        V3<R> c = mk(ofPrev.x * f.x * bg.x * inv, ofPrev.y * f.y * bg.y * inv, ofPrev.z * f.z * bg.z * inv);
        if (c.x == R(0) && c.y == R(0) && c.z == R(0)) return z;
        /* visibility toward the environment (sceneIntersect from the shifted vertex) */
        const R delta = R(3) * feps<R>() * maxAbs(sp.position);
        const V3<R> org = mk(sp.position.x + delta * w.x, sp.position.y + delta * w.y, sp.position.z + delta * w.z);
        R tn = R(0);
        SurfacePoint<R> hs;
        ++shadowRays;
        if (sceneIntersect(org, w, tn, std::numeric_limits<R>::max(), hs)) return z;
        return c;
    }

    /* ---- OMV_EXT_LIGHT_SPLAT ---------------------------------------- */
    const omv_camera_desc* cam = nullptr;   /* the pinhole the splats go to */
    double* splat = nullptr;                /* W*H*3 splat sums (unnormalised) */

    const Light* splatLight = nullptr;      /* the light of the subpath being walked */
    V3<R> splatBeta;                        /* particle weight at the current light vertex */

    /* Light-tracer (t = 1) splat of light-subpath vertex i (every hit of the
     * walk, before the hard cut decides whether the walk goes on).  Particle
     * weight: beta_1 = I * (lightScale * pi/2 * att(r1) * r1^2) at the first
     * hit, pi/2 being 1 / the density of uniformSphere's + octant (so att =
     * 1/d^2 gives the physical particle weight and any attenuation mode
     * reproduces the reference's direct connection in expectation);
     * beta_{i+1} = beta_i * w_i * |w_out . n_i| (the camera subpath's
     * throughput update).  Vertex i projects through the pinhole to pixel p;
     * its contribution
     *   beta_i * getBrdf(w_in, w_c) * |w_c . n_i| * f^2 / (d^2 cos^3(theta))
     * (f: focus in pixels, theta: angle to the optical axis, d: distance to
     * the pinhole) is added to p if the segment to the pinhole is unoccluded;
     * the film adds sum / (W*H*spp). */
    void splatVertex(const Vtx<R>* Pv, int i) {
        const omv_camera_desc& c = *cam;
        if (i == 1) {
            const R r1 = distance3(Pv[1].pos, Pv[0].pos);
            const R k1 = R(P.light_scale_value) * ((gSplatPlant & 2) ? R(1) : kPi<R>() * R(0.5f)) * attenuation(r1, R(splatLight->range)) * r1 * r1;
            splatBeta = mk(R(splatLight->intensity[0]) * k1, R(splatLight->intensity[1]) * k1, R(splatLight->intensity[2]) * k1);
        } else {
            const R ct = std::abs(dot3(Pv[i - 1].outRay, Pv[i - 1].sp.normal));
            const V3<R>& w = Pv[i - 1].w;
            splatBeta = mk(splatBeta.x * w.x * ct, splatBeta.y * w.y * ct, splatBeta.z * w.z * ct);
        }
        const V3<R>& beta = splatBeta;
        const V3<R> C = ld3<R>(c.position);
        const V3<R>& y = Pv[i].pos;
        const V3<R> v = sub(C, y);
        const R d = length3(v);
        if (!(d > R(0))) return;
        const V3<R> wc = mk(v.x / d, v.y / d, v.z / d);
        const V3<R> l = mulT33(sub(y, C), c.rotation);
        if (!(l.z > R(0))) return;
        const R px = R(c.focus) * l.x / l.z + R(c.half_w), py = R(c.focus) * l.y / l.z + R(c.half_h);
        if (!(px >= R(0) && px < R(c.width) && py >= R(0) && py < R(c.height))) return;
        const int ix = std::min((int)px, c.width - 1), iy = std::min((int)py, c.height - 1);
        const R cosT = l.z / length3(l);
        const V3<R> f = matGetBrdf(S, X, Pv[i].sp, Pv[i - 1].outRay, wc);
        const R k = R(c.focus) * R(c.focus) * std::abs(dot3(wc, Pv[i].sp.normal)) / (d * d * cosT * cosT * ((gSplatPlant & 1) ? R(1) : cosT));
        const V3<R> ctb = mk(beta.x * f.x * k, beta.y * f.y * k, beta.z * f.z * k);
        if (ctb.x == R(0) && ctb.y == R(0) && ctb.z == R(0)) return;
        if (!(std::isfinite((double)ctb.x) && std::isfinite((double)ctb.y) && std::isfinite((double)ctb.z))) return;
        /* visibility of the pinhole: ray from the vertex, tnear = the
         * isConnected offset, tfar = d; visible iff nothing is hit */
        R tn = R(2) * R(3) * feps<R>() * maxAbs(y);
        SurfacePoint<R> hs;
        ++shadowRays;
        if (sceneIntersect(y, wc, tn, d, hs)) return;
        double* o = splat + ((size_t)iy * c.width + ix) * 3;
        o[0] += (double)ctb.x; o[1] += (double)ctb.y; o[2] += (double)ctb.z;
    }

    /* computePath (BDPTRenderer.cpp:128-192).  maxN vertices incl. the start. */
    uint64_t neeSample = 0;   /* sample index of the camera path being walked */
    bool computePath(Vtx<R>* Pv, int maxN, const V3<R>& O, const V3<R>& Rd, V3<R>& Rad, int& n, bool inv, R attD,
                     Uniforms& U) {
        if (inv) lastEscW = R(1);
        Pv[0].sp = SurfacePoint<R>();
        Pv[0].pos = O; Pv[0].outRay = Rd; Pv[0].outRad = Rad; Pv[0].outFactor = mk(R(1), R(1), R(1));
        R rDterm = R(1);
        R tnear = R(0);
        n = 1;
        bool exitScene = false;
        for (int i = 1; i < maxN; ++i) {
            /* setRay + shiftRayAlongDirection (BDPTRenderer.cpp:42-50, 309-321) */
            const V3<R>& o0 = Pv[i - 1].pos; const V3<R>& d0 = Pv[i - 1].outRay;
            R delta = R(3) * feps<R>() * maxAbs(o0);
            V3<R> org = mk(o0.x + delta * d0.x, o0.y + delta * d0.y, o0.z + delta * d0.z);
            if (inv) ++camRays; else ++lightRays;
            if (!sceneIntersect(org, d0, tnear, std::numeric_limits<R>::max(), Pv[i].sp)) { exitScene = true; break; }
            Pv[i].pos = Pv[i].sp.position;
            modifyFrame(S, Pv[i].sp);
            V3<R> brdf, erad;
            BounceInfo<R> bi;
            matGetRayAndBrdf(S, X, Pv[i].sp, d0, Pv[i].outRay, brdf, erad, inv, U, E, &bi);
            Pv[i].w = brdf;
            R ct = std::abs(inv ? dot3(Pv[i].outRay, Pv[i].sp.normal) : dot3(d0, Pv[i].sp.normal));
            const V3<R>& of = Pv[i - 1].outFactor;
            Pv[i].outFactor = mk(of.x * brdf.x * ct, of.y * brdf.y * ct, of.z * brdf.z * ct);
            if (inv) {
                Rad = mk(Rad.x + Pv[i].outFactor.x * erad.x, Rad.y + Pv[i].outFactor.y * erad.y, Rad.z + Pv[i].outFactor.z * erad.z);
                Pv[i].outRad = mk(R(0), R(0), R(0));
                if (X.nee) {   /* OMV_EXT_HDRI_NEE: after the vertex's emission */
                    V3<R> c = neeEstimate(Pv[i].sp, bi, of, Pv[i].outRay, i, neeSample);
                    Rad = mk(Rad.x + c.x, Rad.y + c.y, Rad.z + c.z);
                }
            } else {
                if (i == 1) rDterm = attenuation(distance3(Pv[1].pos, Pv[0].pos), attD);
                if (X.splat && splat && cam) splatVertex(Pv, i);   /* OMV_EXT_LIGHT_SPLAT, before the hard cut */
                Pv[i].outRad = mk(Pv[i].outFactor.x * Rad.x * rDterm, Pv[i].outFactor.y * Rad.y * rDterm, Pv[i].outFactor.z * Rad.z * rDterm);
                if (intensity(Pv[i].outRad) < R(P.ray_cut_pix_value)) break;
            }
            n++;
            /* exact shortcut: a camera subpath whose throughput is exactly zero adds
             * nothing more (emission, background and connections all carry it) */
            if (inv && Pv[i].outFactor.x == R(0) && Pv[i].outFactor.y == R(0) && Pv[i].outFactor.z == R(0)) break;
            /* exact shortcut: without emissive materials camera vertices past
             * max_path_length can neither connect nor see the background */
            if (inv && !S.anyEmissive && n > P.max_path_length) break;
        }
        return exitScene;
    }

    /* isConnected (BDPTRenderer.cpp:208-235) */
    bool isConnected(const V3<R>& A, const V3<R>& B) {
        V3<R> dir = sub(B, A);
        R len = length3(dir);
        R maxsz = maxAbs(A);
        R safeD = R(2) * R(3) * feps<R>() * maxsz;
        if (len <= safeD) return true;
        dir = mk(dir.x / len, dir.y / len, dir.z / len);
        R tnear = safeD;
        V3<R> org = mk(A.x + R(0) * dir.x, A.y + R(0) * dir.y, A.z + R(0) * dir.z);
        R tfar = len + (R(0.1f) * len > safeD ? R(0.1f) * len : safeD);
        SurfacePoint<R> sp;
        ++shadowRays;
        if (!sceneIntersect(org, dir, tnear, tfar, sp)) return true;
        if (gPinPlant & 16) return true;
        V3<R> dd = sub(sp.position, B);
        R sd = dot3(dd, dd);
        maxsz = maxAbs(B);
        safeD = R(10) * feps<R>() * maxsz;
        if ((gPinPlant & 32) && sd <= safeD * safeD) {
            SurfacePoint<R> sp2;   // the slip: any further hit up to tfar counts as an occluder
            R tn2 = len + safeD;
            if (sceneIntersect(org, dir, tn2, tfar, sp2)) return false;
        }
        return sd <= safeD * safeD;
    }

    /* HDRBackground::getRadiance (HdrBackground.cpp:55-97) / Background::getRadiance */
    V3<R> background(const V3<R>& dir) {
        if (S.bgType == 1) return mk(R(S.bgConst[0]), R(S.bgConst[1]), R(S.bgConst[2]));
        V3<R> rdir = mulT33(dir, S.bgRot);
        V3<R> rxy = normalize3(mk(rdir.x, rdir.y, R(0)));
        R xang = std::acos(clampv(rxy.x, R(-1), R(1)));
        xang = rxy.y < R(0) ? R(2) * kPi<R>() - xang : xang;
        R yang = std::asin(clampv(rdir.z, R(-1), R(1)));
        yang += R(0.5f) * kPi<R>();
        R u = xang / (R(2) * kPi<R>());
        R v = R(1) - yang / kPi<R>();
        const int W = S.bgW, H = S.bgH;
        R upage = std::floor(u - R(0.5f) / R(W)), vpage = std::floor(v - R(0.5f) / R(H));
        u -= upage; v -= vpage; u *= R(W); v *= R(H);
        int iu = (int)u, iv = (int)v;
        R kui = u - R(iu), kvi = v - R(iv), ku = R(1) - kui, kv = R(1) - kvi;
        int l = iu % W, r = (iu + 1) % W, b = iv % H, t = (iv + 1) % H;
        int id[4] = {b * W + l, b * W + r, t * W + l, t * W + r};
        R w[4] = {ku * kv, kui * kv, ku * kvi, kui * kvi};
        for (int i = 0; i < 4; ++i) id[i] = std::min(std::max(id[i], 0), W * H - 1);
        V3<R> val;
        for (int c = 0; c < 3; ++c)
            at(val, c) = R(S.bgRgb[(size_t)id[0] * 3 + c]) * w[0] + R(S.bgRgb[(size_t)id[1] * 3 + c]) * w[1] +
                         R(S.bgRgb[(size_t)id[2] * 3 + c]) * w[2] + R(S.bgRgb[(size_t)id[3] * 3 + c]) * w[3];
        const R sc = R(S.bgScale);
        return mk(val.x * sc, val.y * sc, val.z * sc);
    }

    /* computePaths (BDPTRenderer.cpp:237-301) for one sample */
    V3<R> computePaths(const V3<R>& O, const V3<R>& Rd, uint64_t sample, Stream& cs, Vtx<R>* cp, Vtx<R>* lp) {
        const int q = P.max_path_length;
        int lpLen = 0, lightIdx = -1;
        V3<R> lightRad;
        if (!S.lights.empty()) {
            Stream ls(P.seed, STREAM_LIGHT, sample);
            Uniforms U; U.s = &ls;
            /* tryComputeLightPath (BDPTRenderer.cpp:194-206) + RenderScene::sampleLight */
            float pdf;
            lightIdx = distanceRandom((int)S.lightCdf.size(), S.lightCdf.data(), S.lightMax, U.next(), &pdf);
            const Light& L = S.lights[lightIdx];
            V3<R> dir = uniformSphere<R>(U);
            const R sc = R(P.light_scale_value);
            lightRad = mk(R(L.intensity[0]) * sc, R(L.intensity[1]) * sc, R(L.intensity[2]) * sc);
            /* exact shortcut: light vertices with index >= max_path_length are
             * never connected (i + j <= q with j >= 1) */
            int maxN = std::min(P.light_bounces + 1, std::max(q, 1));
            V3<R> rad = lightRad;
            splatLight = &L;
            computePath(lp, maxN, ld3<R>(L.pos), dir, rad, lpLen, false, R(L.range), U);
        }
        /* the camera sub-stream continues after the two pixel-jitter draws */
        Uniforms U; U.s = &cs;
        int cpLen = 0;
        V3<R> radiance = mk(R(0), R(0), R(0));
        neeSample = sample;
        bool cpExit = computePath(cp, P.camera_bounces + 1, O, Rd, radiance, cpLen, true, R(0), U);
        int totalPN = 0;
        if (cpExit && S.bgType != 0 && cpLen <= q) {
            V3<R> bg = background(cp[cpLen - 1].outRay);
            const V3<R>& kf = cp[cpLen - 1].outFactor;
            bg = mk(bg.x * kf.x, bg.y * kf.y, bg.z * kf.z);
            if (X.nee) bg = mk(bg.x * lastEscW, bg.y * lastEscW, bg.z * lastEscW);   /* OMV_EXT_HDRI_NEE */
            radiance = mk(bg.x * R(1) + radiance.x, bg.y * R(1) + radiance.y, bg.z * R(1) + radiance.z);
            ++totalPN;
        } else if (intensity(radiance) > feps<R>()) ++totalPN;
        if (lightIdx >= 0) {
            const Light& L = S.lights[lightIdx];
            for (int i = 0; i < lpLen; ++i) {
                V3<R> lRad = lp[i].outRad;
                const V3<R>& lPos = lp[i].pos;
                for (int j = 1; j < cpLen && i + j <= q; ++j) {
                    const V3<R>& cPos = cp[j].pos;
                    V3<R> LC = normalize3(sub(cPos, lPos));
                    if (i == 0) {
                        lRad = ld3<R>(L.intensity);
                        R lsc = attenuation(distance3(cPos, lPos), R(L.range));
                        R k = R(P.light_scale_value) * lsc;
                        lRad = mk(lRad.x * k, lRad.y * k, lRad.z * k);
                    }
                    R c = std::abs(dot3(LC, cp[j].sp.normal));
                    const V3<R>& of = cp[j - 1].outFactor;
                    V3<R> lc = mk(lRad.x * c * of.x, lRad.y * c * of.y, lRad.z * c * of.z);
                    if (intensity(lc) < R(P.ray_cut_pix_value)) continue;
                    V3<R> outV = neg(cp[j - 1].outRay);
                    V3<R> brdf = matGetBrdf(E ? *E : S, X, cp[j].sp, LC, outV);
                    brdf = mk(brdf.x * c, brdf.y * c, brdf.z * c);
                    lc = mk(brdf.x * of.x * lRad.x, brdf.y * of.y * lRad.y, brdf.z * of.z * lRad.z);
                    if (intensity(lc) < R(P.ray_cut_pix_value)) continue;
                    if (!isConnected(lPos, cPos)) continue;
                    radiance = mk(lc.x * R(1) + radiance.x, lc.y * R(1) + radiance.y, lc.z * R(1) + radiance.z);
                }
            }
        }
        R scale = totalPN ? R(1) / R(totalPN) : R(1);
        return mk(radiance.x * scale, radiance.y * scale, radiance.z * scale);
    }

    /* Gradient accumulators of the adjoint: per texture (w*h*4) and per
     * material slot (dn[3], dk[3]); empty = not a target. */
    struct AdjGrad {
        std::vector<std::vector<R>> tex, om;
    };

    void scatterTex(const omv_pbr_desc& d, const SurfacePoint<R>& sp, int slot, int ch, R g, AdjGrad& G) {
        const int id = d.texture_map[slot];
        if (!(0 <= id && id < (int)G.tex.size()) || G.tex[id].empty() || !(g != R(0))) return;
        R u, v;
        if (!getTexCoord(S, sp, d.texture_channel[slot], u, v)) return;
        V3<R> uvT = mul33(d.texture_transform[slot], mk(u, v, R(1)));
        int ids[4];
        float w[4];
        if (!texAddr(S.textures[id], (float)uvT.x, (float)uvT.y, ids, w)) return;
        for (int i = 0; i < 4; ++i) G.tex[id][(size_t)ids[i] * 4 + ch] += g * R(w[i]);
    }

    /* Path-replay adjoint of one camera sample (computePath with inv = true,
     * BDPTRenderer.cpp:149-176, in a scene without point lights).  C is the
     * sample's radiance and A = d loss / d C.  Every term of C at or after
     * bounce i carries that bounce's weight g_i, so the bounce adds
     * A . R_i . d ln g_i / d theta, R_i being C minus the terms before i. */
    void adjointSample(const V3<R>& O, const V3<R>& Rd, Stream& cs, const V3<R>& C, const R A[3], AdjGrad& G) {
        Uniforms U; U.s = &cs;
        const int q = P.max_path_length;
        R rem[3] = {C.x, C.y, C.z};
        V3<R> pos = O, dir = Rd, of = mk(R(1), R(1), R(1));
        R tnear = R(0);
        int n = 1;
        for (int i = 1; i < P.camera_bounces + 1; ++i) {
            R delta = R(3) * feps<R>() * maxAbs(pos);
            V3<R> org = mk(pos.x + delta * dir.x, pos.y + delta * dir.y, pos.z + delta * dir.z);
            SurfacePoint<R> sp;
            if (!sceneIntersect(org, dir, tnear, std::numeric_limits<R>::max(), sp)) break;
            modifyFrame(S, sp);
            V3<R> outV, brdf, erad;
            BounceInfo<R> bi;
            matGetRayAndBrdf(S, X, sp, dir, outV, brdf, erad, true, U, nullptr, &bi);
            R wc[3];
            for (int c = 0; c < 3; ++c) wc[c] = A[c] * rem[c];
            const Material& m = S.materials[S.meshes[sp.mesh].matId];
            if (bi.kind == 1 && bi.b.lobe != 0) {
                const Bsdf<R>& b = bi.b;
                const omv_pbr_desc& d = m.pbr;
                LobeGeo<R> g = lobeGeo(b);
                R dC[3], dM[3], dA[3];
                lobeDlog(g, b.color, b.metal, b.alpha, dC, dM, dA);
                R gMetal = R(0), gAlpha = R(0);
                const R occ = (bi.omr[0] - R(1)) * R(d.occlusion_strength) + R(1);
                for (int c = 0; c < 3; ++c) {
                    scatterTex(d, sp, OMV_PBR_BASE_COLOR, c, wc[c] * dC[c] * occ * R(d.base_color_factor[c]), G);
                    gMetal += wc[c] * dM[c];
                    gAlpha += wc[c] * dA[c];
                }
                scatterTex(d, sp, OMV_PBR_METALLIC_ROUGHNESS, 2, gMetal * R(d.metallic_factor), G);
                const R r2 = bi.omr[2];
                if (r2 > R(0.00001f) && r2 < R(0.99999f))
                    scatterTex(d, sp, OMV_PBR_METALLIC_ROUGHNESS, 1, gAlpha * R(2) * b.rough * R(d.roughness_factor), G);
            } else if (bi.kind == 2) {
                const int slot = S.meshes[sp.mesh].matId;
                if (slot < (int)G.om.size() && !G.om[slot].empty()) {
                    for (int c = 0; c < 3; ++c) {
                        R rp, rs, dpn, dpk, dsn, dsk;
                        fresnel(R(m.n[c]), R(m.k[c]), bi.fCos, rp, rs);
                        fresnelGrad(R(m.n[c]), R(m.k[c]), bi.fCos, dpn, dpk, dsn, dsk);
                        const R F = rp + rs;
                        if (!(F != R(0))) continue;
                        G.om[slot][c] += wc[c] * (dpn + dsn) / F;
                        G.om[slot][3 + c] += wc[c] * (dpk + dsk) / F;
                    }
                }
            }
            R ct = std::abs(dot3(outV, sp.normal));
            of = mk(of.x * brdf.x * ct, of.y * brdf.y * ct, of.z * brdf.z * ct);
            rem[0] -= of.x * erad.x; rem[1] -= of.y * erad.y; rem[2] -= of.z * erad.z;
            n++;
            if (of.x == R(0) && of.y == R(0) && of.z == R(0)) break;
            if (!S.anyEmissive && n > q) break;
            pos = sp.position;
            dir = outV;
        }
    }

    /* FC paint (FCSubject.h:153-482 + the FC branches of the three materials) */
    void fcPaint(const SurfacePoint<R>& sp, const V3<R>& inV, V3<R>& col, Uniforms& U) {
        const Mesh& mesh = S.meshes[sp.mesh];
        const Material& m = S.materials[mesh.matId];
        const int sub = S.fcSubject;
        const uint32_t* tab = fcTable.data();
        const uint32_t tsz = (uint32_t)fcTable.size();
        bool isMissing = (m.type == MAT_DIFFUSE);
        V3<R> inDir = inV;
        if (isMissing) {
            /* RenderMaterial::getRayAndBrdf draws the hemisphere direction, then calls
             * getBrdf(-outV, sp, -inV): the "in" test sees -outV */
            V3<R> v = uniformHemisphere<R>(U);
            V3<R> outV = toBasis(sp.tangent, sp.binormal, sp.normal, v);
            inDir = neg(outV);
        }
        switch (sub) {
            case OMV_FC_PID: col = fcPaintEnumerated<R>((uint32_t)sp.globalFace, tab, tsz, false); break;
            case OMV_FC_GID: col = fcPaintEnumerated<R>((uint32_t)sp.mesh, tab, tsz, false); break;
            case OMV_FC_IID: col = fcPaintEnumerated<R>((uint32_t)sp.inst, tab, tsz, false); break;
            case OMV_FC_IPID: col = fcPaintEnumerated<R>(fcInstBase[sp.inst] + (uint32_t)sp.face, tab, tsz, false); break;
            case OMV_FC_MID:
                /* the missing material passes id -1, painted pink (FCSubject.h:239-244) */
                if (isMissing) col = mk(R(1000), R(0), R(1000));
                else col = fcPaintEnumerated<R>((uint32_t)mesh.matId, tab, tsz, true);
                break;
            case OMV_FC_RMP: case OMV_FC_MMP: {
                /* PBR hands omr[2] / omr[1] (PhysicallyBasedMaterial.cpp:216-219); the
                 * OM and missing materials hand a default (zero) argument */
                R a = R(0);
                if (m.type == MAT_PBR) {
                    R color[4], omr[3];
                    pbrInputs(S, sp, m.pbr, color, omr);
                    a = sub == OMV_FC_RMP ? omr[2] : omr[1];
                }
                col = fcPaintClamped<R>(sub, a);
                break;
            }
            case OMV_FC_MD: col = fcPaintClamped<R>(OMV_FC_MD, sp.density); break;
            case OMV_FC_SG: col = fcPaintClamped<R>(OMV_FC_SG, sp.gradient); break;
            case OMV_FC_MN: col = fcPaintMaterialName<R>(m.type); break;
            case OMV_FC_IN: {
                V3<R> tfn = faceNormal(S, sp);
                col = fcPaintInverted<R>(dot3(inDir, tfn) > R(0));
                break;
            }
            default: col = mk(R(0), R(0), R(0)); break;
        }
    }

    static int distanceRandom(int n, const float* cdf, float maxD, float u, float* pdf) {
        if (n <= 0) { *pdf = 0; return -1; }
        float d = u * maxD;
        int idx;
        if (n == 1) idx = 0;
        else if (d <= cdf[0]) idx = 0;
        else if (d >= cdf[n - 2]) idx = n - 1;
        else {
            int L = 0, Rr = n - 1;
            while (L < Rr) { int C = (L + Rr) / 2; if (d <= cdf[C]) Rr = C; else L = C + 1; }
            idx = (L + Rr) / 2;
        }
        *pdf = (cdf[idx] - (idx > 0 ? cdf[idx - 1] : 0.0f)) / maxD;
        return idx;
    }
};

/* PinholeCamera::getRay (PinholeCamera.cpp:80-95) with explicit jitter uniforms */
template <class R>
static void cameraRay(const omv_camera_desc& c, int x, int y, float u0, float u1, V3<R>& O, V3<R>& D) {
    float px = (float)x + 0.5f - c.half_w;
    float py = (float)y + 0.5f - c.half_h;
    float dx = lerpv(-0.5f, 0.5f, u0);
    float dy = lerpv(-0.5f, 0.5f, u1);
    V3<float> cr = normalize3(mk(px + dx, py + dy, c.focus));
    V3<float> w = mul33(c.rotation, cr);
    O = ld3<R>(c.position);
    D = mk(R(w.x), R(w.y), R(w.z));
}

template <class R>
static void renderRows(const Scene& S, const omv_renderer_params& P, const omv_camera_desc& cam, int row0, int rows,
                       int threads, float* film, float* samplesOut, omv_render_stats* st, const Scene* E = nullptr,
                       double* film64 = nullptr) {
    const int W = cam.width, H = cam.height, spp = std::max(P.samples_per_pixel, 0);
    const int rEnd = std::min(H, row0 + rows);
    std::atomic<int> next(row0);
    std::atomic<uint64_t> rays(0), cr(0), lr(0), sr(0), nodes(0), tris(0);
    /* OMV_EXT_LIGHT_SPLAT: per-thread splat sums, added to the film below */
    const bool splat = (P.extensions & OMV_EXT_LIGHT_SPLAT) && !S.lights.empty() && (film || film64);
    std::mutex splatMu;
    std::vector<double> splatSum(splat ? (size_t)W * H * 3 : 0, 0.0);
    auto worker = [&]() {
        Tracer<R> T(S, P);
        T.E = E;
        std::vector<double> mySplat(splat ? (size_t)W * H * 3 : 0, 0.0);
        if (splat) { T.cam = &cam; T.splat = mySplat.data(); }
        std::vector<Vtx<R>> cp(P.camera_bounces + 2), lp(P.light_bounces + 2);
        int row;
        while ((row = next.fetch_add(1)) < rEnd) {
            for (int x = 0; x < W; ++x) {
                const int id = row * W + x;
                R acc[3] = {R(0), R(0), R(0)};
                for (int k = 0; k < spp; ++k) {
                    const uint64_t s = (uint64_t)id * spp + k;
                    Stream cs(P.seed, STREAM_CAMERA, s);
                    float u0 = cs.next(), u1 = cs.next();
                    V3<R> O, D;
                    cameraRay<R>(cam, x, row, u0, u1, O, D);
                    V3<R> rad = T.computePaths(O, D, s, cs, cp.data(), lp.data());
                    if (samplesOut) { float* o = samplesOut + s * 3; o[0] = (float)rad.x; o[1] = (float)rad.y; o[2] = (float)rad.z; }
                    /* PinholeCamera::hit (:97-108): film += radiance in sample order */
                    acc[0] = acc[0] + rad.x; acc[1] = acc[1] + rad.y; acc[2] = acc[2] + rad.z;
                }
                /* PinholeCamera::stop (:70-78) */
                R cnt = R(spp);
                R sc = cnt < R(0.5f) ? R(1) : R(1) / cnt;
                if (film) {
                    float* f = film + (size_t)id * 4;
                    f[0] = (float)(sc * acc[0]); f[1] = (float)(sc * acc[1]); f[2] = (float)(sc * acc[2]); f[3] = 1.0f;
                }
                if (film64) {
                    double* f = film64 + (size_t)id * 4;
                    f[0] = (double)(sc * acc[0]); f[1] = (double)(sc * acc[1]); f[2] = (double)(sc * acc[2]); f[3] = 1.0;
                }
            }
        }
        rays += T.rays; cr += T.camRays; lr += T.lightRays; sr += T.shadowRays; nodes += T.cnt.nodes; tris += T.cnt.tris;
        if (splat) {
            std::lock_guard<std::mutex> g(splatMu);
            for (size_t k = 0; k < mySplat.size(); ++k) splatSum[k] += mySplat[k];
        }
    };
    int nt = threads > 0 ? threads : (P.number_of_cores > 0 ? P.number_of_cores : 1);
    nt = std::max(1, std::min(nt, std::max(1, rEnd - row0)));
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back(worker);
    for (auto& t : th) t.join();
    if (splat) {
        /* film = camera estimate + splat sums over all W*H*spp light paths */
        const double inv = 1.0 / ((double)W * H * std::max(spp, 1));
        for (size_t p = 0; p < (size_t)W * H; ++p)
            for (int c = 0; c < 3; ++c) {
                if (film) film[p * 4 + c] = film[p * 4 + c] + (float)(splatSum[p * 3 + c] * inv);
                if (film64) film64[p * 4 + c] = film64[p * 4 + c] + splatSum[p * 3 + c] * inv;
            }
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->samples = (uint64_t)(rEnd - row0) * W * spp;
        st->rays = rays; st->camera_rays = cr; st->light_rays = lr; st->shadow_rays = sr;
        st->nodes_visited = nodes; st->tris_tested = tris;
    }
}

/* The adjoint over the whole image: d loss / d film (W*H*4, rgb used) in,
 * gradients of the target textures and OM materials out (added to). */
template <class R>
static void adjointImage(const Scene& S, const omv_renderer_params& P, const omv_camera_desc& cam, const float* dfilm,
                         const std::vector<int>& texT, const std::vector<int>& omT, int threads,
                         std::vector<std::vector<double>>& texOut, std::vector<std::vector<double>>& omOut) {
    const int W = cam.width, H = cam.height, spp = std::max(P.samples_per_pixel, 0);
    std::atomic<int> next(0);
    std::mutex mu;
    auto worker = [&]() {
        Tracer<R> T(S, P);
        typename Tracer<R>::AdjGrad G;
        G.tex.resize(S.textures.size());
        G.om.resize(S.materials.size());
        for (int id : texT) G.tex[id].assign((size_t)S.textures[id].w * S.textures[id].h * 4, R(0));
        for (int id : omT) G.om[id].assign(6, R(0));
        std::vector<Vtx<R>> cp(P.camera_bounces + 2), lp(P.light_bounces + 2);
        int row;
        while ((row = next.fetch_add(1)) < H) {
            for (int x = 0; x < W; ++x) {
                const int id = row * W + x;
                const R sc = R(spp) < R(0.5f) ? R(1) : R(1) / R(spp);
                R A[3] = {R(dfilm[(size_t)id * 4 + 0]) * sc, R(dfilm[(size_t)id * 4 + 1]) * sc, R(dfilm[(size_t)id * 4 + 2]) * sc};
                if (A[0] == R(0) && A[1] == R(0) && A[2] == R(0)) continue;
                for (int k = 0; k < spp; ++k) {
                    const uint64_t s = (uint64_t)id * spp + k;
                    Stream cs(P.seed, STREAM_CAMERA, s);
                    float u0 = cs.next(), u1 = cs.next();
                    V3<R> O, D;
                    cameraRay<R>(cam, x, row, u0, u1, O, D);
                    V3<R> C = T.computePaths(O, D, s, cs, cp.data(), lp.data());
                    if (C.x == R(0) && C.y == R(0) && C.z == R(0)) continue;
                    Stream cs2(P.seed, STREAM_CAMERA, s);
                    (void)cs2.next(); (void)cs2.next();
                    T.adjointSample(O, D, cs2, C, A, G);
                }
            }
        }
        std::lock_guard<std::mutex> lk(mu);
        for (int id : texT) for (size_t i = 0; i < G.tex[id].size(); ++i) texOut[id][i] += (double)G.tex[id][i];
        for (int id : omT) for (int i = 0; i < 6; ++i) omOut[id][i] += (double)G.om[id][i];
    };
    int nt = std::max(1, std::min(threads > 0 ? threads : 1, H));
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back(worker);
    for (auto& t : th) t.join();
}

}  // namespace refcpu

/* ================================================================ C API */
using namespace refcpu;

struct ref_scene { Scene s; };

static bool meshOk(ref_scene* s, int id) { return s && id >= 0 && id < (int)s->s.meshes.size(); }

extern "C" {

ref_scene* ref_scene_create(void) { return new ref_scene(); }
void ref_scene_destroy(ref_scene* s) { delete s; }

int ref_scene_allocate(ref_scene* r, int meshes, int instances, int materials, int textures, int lights) {
    if (!r || meshes < 0 || instances < 0 || materials < 0 || textures < 0 || lights < 0) return OMV_E_ARG;
    Scene& S = r->s;
    S = Scene();
    S.meshes.resize(meshes);
    S.instances.resize(instances);
    S.materials.resize(materials + 1);
    S.materials[materials].type = MAT_DIFFUSE;  // RenderScene.cpp:65-68
    S.textures.resize(textures);
    S.lights.resize(lights);
    return OMV_OK;
}

int ref_scene_set_mesh(ref_scene* r, int meshId, int vertN, int facesN, int texN, int matId, const uint32_t* faces,
                       const float* vert, int vstride, const float* norm, int nstride, const float* tans, int tanstride,
                       const float* const* tex, int texstride) {
    if (!meshOk(r, meshId)) return OMV_E_RANGE;
    if (vertN < 0 || facesN < 0 || texN < 0 || (facesN > 0 && !faces)) return OMV_E_ARG;
    Mesh m;
    int utex = 0;
    for (int i = 0; i < texN; ++i) utex += (tex && tex[i]) ? 1 : 0;
    m.id = meshId; m.vn = vertN; m.fn = facesN;
    m.texChN = utex < 8 ? utex : 8;
    m.stride = 10 + m.texChN * 2;
    m.F.assign(faces, faces + (size_t)3 * facesN);
    m.V.assign((size_t)vertN * m.stride, 0.0f);
    for (int i = 0; i < m.texChN; ++i) m.texOffs[i] = 10 + i * 2;
    m.matId = matId;
    if (vert) { m.vertexDef = true; for (int i = 0; i < vertN; ++i) for (int c = 0; c < 3; ++c) m.V[(size_t)i * m.stride + c] = vert[(size_t)i * vstride + c]; }
    if (norm) { m.normalDef = true; for (int i = 0; i < vertN; ++i) for (int c = 0; c < 3; ++c) m.V[(size_t)i * m.stride + 3 + c] = norm[(size_t)i * nstride + c]; }
    if (tans) { m.tangentDef = true; for (int i = 0; i < vertN; ++i) for (int c = 0; c < 4; ++c) m.V[(size_t)i * m.stride + 6 + c] = tans[(size_t)i * tanstride + c]; }
    for (int ch = 0; ch < texN && ch < 8; ++ch) {
        if (!(tex && tex[ch])) continue;
        if (m.texMap[ch] == -1) m.texMap[ch] = m.filledTexChN++;
        int slot = m.texMap[ch];
        if (slot >= m.texChN) continue;
        m.texDef[slot] = true;
        for (int i = 0; i < vertN; ++i) for (int c = 0; c < 2; ++c) m.V[(size_t)i * m.stride + m.texOffs[slot] + c] = tex[ch][(size_t)i * texstride + c];
    }
    r->s.meshes[meshId] = std::move(m);
    return OMV_OK;
}

int ref_scene_set_instance(ref_scene* r, int instId, const float trans[16], int meshId) {
    if (!r || instId < 0 || instId >= (int)r->s.instances.size()) return OMV_E_RANGE;
    if (!trans) return OMV_E_ARG;
    Instance& in = r->s.instances[instId];
    in.instId = instId; in.meshId = meshId;
    std::memcpy(in.tm, trans, sizeof(float) * 16);
    /* RenderInstance::computeNormalMatrix (RenderMesh.h:36-43) with inverseMatrix3 (UtilityFunctions.h:342-349) */
    float M[3][3];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M[i][j] = trans[j * 4 + i];
    float det = M[0][0] * M[1][1] * M[2][2] + M[0][1] * M[1][2] * M[2][0] + M[0][2] * M[1][0] * M[2][1] -
                M[0][2] * M[1][1] * M[2][0] - M[0][1] * M[1][0] * M[2][2] - M[0][0] * M[1][2] * M[2][1];
    float k = (float)(1.0 / det);
    float* R = in.nm;
    R[0] = k * (M[1][1] * M[2][2] - M[1][2] * M[2][1]); R[1] = k * (-M[0][1] * M[2][2] + M[0][2] * M[2][1]); R[2] = k * (M[0][1] * M[1][2] - M[0][2] * M[1][1]);
    R[3] = k * (-M[1][0] * M[2][2] + M[1][2] * M[2][0]); R[4] = k * (M[0][0] * M[2][2] - M[0][2] * M[2][0]); R[5] = k * (-M[0][0] * M[1][2] + M[0][2] * M[1][0]);
    R[6] = k * (M[1][0] * M[2][1] - M[1][1] * M[2][0]); R[7] = k * (-M[0][0] * M[2][1] + M[0][1] * M[2][0]); R[8] = k * (M[0][0] * M[1][1] - M[0][1] * M[1][0]);
    affineInverse(trans, in.invd);
    for (int i = 0; i < 12; ++i) in.inv[i] = (float)in.invd[i];
    return OMV_OK;
}

static int matSlot(ref_scene* r, int matId) {
    if (!r || matId < 0 || matId >= (int)r->s.materials.size() - 1) return -1;
    return matId;
}

int ref_scene_set_material_om(ref_scene* r, int matId, float T, int nt, const omv_ior_table* tables) {
    if (matSlot(r, matId) < 0) return OMV_E_RANGE;
    Material m;
    m.type = MAT_OM; m.omHasBase = true; m.omTemperature = T;
    std::vector<IorTable> tabs = makeTables(nt, tables);
    for (int c = 0; c < 3; ++c) {
        float n = 0, k = 0;
        if (getIor(tabs, T, kRgbWl[c], n, k) != 0) return OMV_E_ARG;
        m.n[c] = n; m.k[c] = k;
    }
    r->s.materials[matId] = m;
    return OMV_OK;
}

int ref_scene_set_material_om_nk(ref_scene* r, int matId, const float n[3], const float k[3]) {
    if (matSlot(r, matId) < 0) return OMV_E_RANGE;
    Material m;
    m.type = MAT_OM; m.omHasBase = true;
    for (int c = 0; c < 3; ++c) { m.n[c] = n[c]; m.k[c] = k[c]; }
    r->s.materials[matId] = m;
    return OMV_OK;
}

int ref_scene_set_material_pbr(ref_scene* r, int matId, const omv_pbr_desc* d) {
    if (matSlot(r, matId) < 0) return OMV_E_RANGE;
    if (!d) return OMV_E_ARG;
    Material m;
    m.type = MAT_PBR; m.pbr = *d;
    r->s.materials[matId] = m;
    return OMV_OK;
}

int ref_scene_set_texture(ref_scene* r, int id, int w, int h, const uint8_t* rgba) {
    if (!r || id < 0 || id >= (int)r->s.textures.size()) return OMV_E_RANGE;
    Texture t; t.w = w; t.h = h; t.f32 = false;
    if (rgba && w > 0 && h > 0) t.rgba.assign(rgba, rgba + (size_t)w * h * 4);
    r->s.textures[id] = std::move(t);
    return OMV_OK;
}

int ref_scene_set_texture_f32(ref_scene* r, int id, int w, int h, const float* rgba) {
    if (!r || id < 0 || id >= (int)r->s.textures.size()) return OMV_E_RANGE;
    Texture t; t.w = w; t.h = h; t.f32 = true;
    if (rgba && w > 0 && h > 0) t.rgbaf.assign(rgba, rgba + (size_t)w * h * 4);
    r->s.textures[id] = std::move(t);
    return OMV_OK;
}

int ref_scene_set_light_point(ref_scene* r, int id, const float pos[3], const float inten[3], float range) {
    if (!r || id < 0 || id >= (int)r->s.lights.size()) return OMV_E_RANGE;
    Light& L = r->s.lights[id];
    for (int c = 0; c < 3; ++c) { L.pos[c] = pos[c]; L.intensity[c] = inten[c]; }
    L.range = range < 0.0f ? std::numeric_limits<float>::max() : range;
    L.power = 0.2126f * L.intensity[0] + 0.7152f * L.intensity[1] + 0.0722f * L.intensity[2];
    L.set = true;
    return OMV_OK;
}

int ref_scene_set_background_hdr(ref_scene* r, int w, int h, const float* rgb, const float rot[9], float scale) {
    if (!r || w <= 0 || h <= 0 || !rgb) return OMV_E_ARG;
    Scene& S = r->s;
    S.bgType = 2; S.bgW = w; S.bgH = h;
    S.bgRgb.assign(rgb, rgb + (size_t)w * h * 3);
    if (rot) std::memcpy(S.bgRot, rot, sizeof(float) * 9);
    S.bgScale = scale;
    return OMV_OK;
}

int ref_scene_set_background_constant(ref_scene* r, const float rgb[3]) {
    if (!r || !rgb) return OMV_E_ARG;
    r->s.bgType = 1;
    for (int c = 0; c < 3; ++c) r->s.bgConst[c] = rgb[c];
    return OMV_OK;
}

int ref_scene_set_false_color_subject(ref_scene* r, const char* name) {
    if (!r || !name) return OMV_E_ARG;
    static const char* names[] = {"", "pid", "gid", "mid", "rmp", "mmp", "md", "sg", "mn", "in", "iid", "ipid"};
    r->s.fcSubject = OMV_FC_NONE;
    for (int i = 1; i < 12; ++i) if (std::strcmp(name, names[i]) == 0) r->s.fcSubject = i;
    return OMV_OK;
}

int ref_scene_commit(ref_scene* r, char* log, size_t cap) {
    if (!r) return OMV_E_ARG;
    Scene& S = r->s;
    std::string lg;
    const int M = (int)S.materials.size();
    for (int i = 0; i < (int)S.meshes.size(); ++i) {
        Mesh& m = S.meshes[i];
        int normCh = -1;
        if (m.matId >= 0 && m.matId < M) {
            Material& mat = S.materials[m.matId];
            if (mat.type != MAT_NONE) {
                if (mat.type == MAT_PBR) normCh = mat.pbr.texture_channel[OMV_PBR_NORMAL];
            } else {
                lg += std::to_string(i) + " mesh refers to undefined material" + std::to_string(m.matId) + "\n";
                m.matId = M - 1;
            }
        } else m.matId = M - 1;
        /* RenderMesh::commitMesh (RenderMesh.cpp:246-265) */
        m.valid = m.vn > 0 && m.fn > 0 && m.vertexDef;
        if (!m.valid) {
            lg += std::to_string(i) + " mesh is inconsistent\n";
            int keepMat = m.matId;
            m = Mesh(); m.matId = keepMat; m.valid = false;
            continue;
        }
        if (!m.normalDef) computeNormals(m);
        if (!m.tangentDef) {
            int ch = (normCh >= 0 && normCh < 8 && m.texMap[normCh] >= 0) ? m.texMap[normCh] : -1;
            if (ch >= 0 && m.texDef[ch]) computeTangentsUV(m, ch);
            else { for (int v = 0; v < m.vn; ++v) randomTangent(&m.V[(size_t)v * m.stride + 3], &m.V[(size_t)v * m.stride + 6]); m.tangentDef = true; }
        }
    }
    int falseInst = 0;
    for (auto& in : S.instances) {
        if (in.meshId < 0 || in.meshId >= (int)S.meshes.size() || !S.meshes[in.meshId].valid) { in.meshId = -1; ++falseInst; }
    }
    if (falseInst) lg += std::to_string(falseInst) + " from " + std::to_string(S.instances.size()) + " instances are invalid\n";
    if (log && cap) { std::strncpy(log, lg.c_str(), cap - 1); log[cap - 1] = 0; }
    if (falseInst == (int)S.instances.size()) return OMV_E_SCENE;
    S.faceBase.assign(S.meshes.size() + 1, 0);
    for (size_t i = 0; i < S.meshes.size(); ++i) S.faceBase[i + 1] = S.faceBase[i] + S.meshes[i].fn;
    /* FC min/max density (RenderScene.cpp:196-203, RenderMesh.cpp:277-297) */
    S.minDensity = std::numeric_limits<float>::infinity(); S.maxDensity = 0.0f;
    for (auto& in : S.instances) {
        if (in.meshId < 0) continue;
        const Mesh& m = S.meshes[in.meshId];
        for (int f = 0; f < m.fn; ++f) {
            V3<float> V0 = mul34(in.tm, ld3<float>(m.vtx(m.F[3 * f]))), V1 = mul34(in.tm, ld3<float>(m.vtx(m.F[3 * f + 1]))),
                      V2 = mul34(in.tm, ld3<float>(m.vtx(m.F[3 * f + 2])));
            float a = length3(cross3(sub(V1, V0), sub(V2, V0))) / 2.0f;
            if (a > std::numeric_limits<float>::epsilon()) {
                float d = 1.0f / (0.0f + a);
                S.minDensity = std::min(S.minDensity, d); S.maxDensity = std::max(S.maxDensity, d);
            }
        }
    }
    /* light CDF (RenderScene.cpp:213-220) */
    S.lightCdf.clear();
    float acc = 0.0f;
    for (auto& L : S.lights) { acc = acc + L.power; S.lightCdf.push_back(acc); }
    S.lightMax = S.lightCdf.empty() ? 0.0f : S.lightCdf.back();
    S.anyEmissive = false;
    for (auto& m : S.materials) {
        if (m.type != MAT_PBR) continue;
        V3<float> ef = mk(m.pbr.emissive_factor[0], m.pbr.emissive_factor[1], m.pbr.emissive_factor[2]);
        int id = m.pbr.texture_map[OMV_PBR_EMISSIVITY];
        m.pbrEmissive = length3(ef) > std::numeric_limits<float>::epsilon() && id >= 0 && id < (int)S.textures.size();
        S.anyEmissive = S.anyEmissive || m.pbrEmissive;
    }
    /* BVHs */
    S.blas.assign(S.meshes.size(), Bvh());
    std::vector<Box> meshBox(S.meshes.size(), emptyBox());
    for (size_t i = 0; i < S.meshes.size(); ++i) {
        const Mesh& m = S.meshes[i];
        if (!m.valid) continue;
        std::vector<Box> bx(m.fn);
        for (int f = 0; f < m.fn; ++f) {
            Box b = emptyBox();
            for (int c = 0; c < 3; ++c) { const float* p = m.vtx(m.F[3 * f + c]); Box pb; for (int a = 0; a < 3; ++a) pb.mn[a] = pb.mx[a] = p[a]; grow(b, pb); }
            bx[f] = b; grow(meshBox[i], b);
        }
        buildBvh(S.blas[i], bx);
    }
    std::vector<Box> ib;
    std::vector<int> ids;
    for (auto& in : S.instances) {
        if (in.meshId < 0) continue;
        const Box& mb = meshBox[in.meshId];
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = 0; c < 8; ++c) {
            double p[3] = {(c & 1) ? mb.mx[0] : mb.mn[0], (c & 2) ? mb.mx[1] : mb.mn[1], (c & 4) ? mb.mx[2] : mb.mn[2]};
            for (int a = 0; a < 3; ++a) {
                double w = (double)in.tm[a * 4] * p[0] + (double)in.tm[a * 4 + 1] * p[1] + (double)in.tm[a * 4 + 2] * p[2] + (double)in.tm[a * 4 + 3];
                lo[a] = std::min(lo[a], w); hi[a] = std::max(hi[a], w);
            }
        }
        Box b;
        for (int a = 0; a < 3; ++a) {
            double pad = 1e-5 * (std::fabs(lo[a]) + std::fabs(hi[a])) + 1e-20;
            b.mn[a] = std::nextafter((float)(lo[a] - pad), -INFINITY); b.mx[a] = std::nextafter((float)(hi[a] + pad), INFINITY);
            in.wbmin[a] = b.mn[a]; in.wbmax[a] = b.mx[a];
        }
        ib.push_back(b); ids.push_back(in.instId);
    }
    buildBvh(S.tlas, ib);
    for (auto& p : S.tlas.prims) p = ids[p];
    buildEnvTables(S);
    S.committed = true;
    return OMV_OK;
}

int ref_scene_mesh_info(ref_scene* r, int id, int* vn, int* fn, int* stride, int* mat) {
    if (!meshOk(r, id)) return OMV_E_RANGE;
    const Mesh& m = r->s.meshes[id];
    if (vn) *vn = m.vn;
    if (fn) *fn = m.fn;
    if (stride) *stride = m.stride;
    if (mat) *mat = m.matId;
    return OMV_OK;
}

int ref_scene_get_vertices(ref_scene* r, int id, float* out, size_t cap) {
    if (!meshOk(r, id)) return OMV_E_RANGE;
    const Mesh& m = r->s.meshes[id];
    if (cap < m.V.size()) return OMV_E_ARG;
    std::memcpy(out, m.V.data(), m.V.size() * sizeof(float));
    return OMV_OK;
}

int ref_scene_get_om_nk(ref_scene* r, int id, float n[3], float k[3]) {
    if (matSlot(r, id) < 0) return OMV_E_RANGE;
    const Material& m = r->s.materials[id];
    for (int c = 0; c < 3; ++c) { n[c] = m.n[c]; k[c] = m.k[c]; }
    return m.type == MAT_OM ? OMV_OK : OMV_E_ARG;
}

int ref_render(ref_scene* r, const omv_renderer_params* p, const omv_camera_desc* cam, int fp64, int row0, int rows,
               int threads, float* film, omv_render_stats* st) {
    if (!r || !p || !cam || !r->s.committed) return OMV_E_ARG;
    if (rows <= 0) { row0 = 0; rows = cam->height; }
    if (fp64) renderRows<double>(r->s, *p, *cam, row0, rows, threads, film, nullptr, st);
    else renderRows<float>(r->s, *p, *cam, row0, rows, threads, film, nullptr, st);
    return OMV_OK;
}

/* Detached render for finite differences of the adjoint: directions and
 * lobes sampled with `sample`'s materials, weights evaluated with `eval`'s
 * (same geometry, same texture sizes). */
int ref_render_detached(ref_scene* sample, ref_scene* eval, const omv_renderer_params* p, const omv_camera_desc* cam, int fp64,
                        int threads, double* film) {
    if (!sample || !eval || !p || !cam || !sample->s.committed || !eval->s.committed) return OMV_E_ARG;
    if (sample->s.materials.size() != eval->s.materials.size() || sample->s.meshes.size() != eval->s.meshes.size()) return OMV_E_ARG;
    if (fp64) renderRows<double>(sample->s, *p, *cam, 0, cam->height, threads, nullptr, nullptr, nullptr, &eval->s, film);
    else renderRows<float>(sample->s, *p, *cam, 0, cam->height, threads, nullptr, nullptr, nullptr, &eval->s, film);
    return OMV_OK;
}

/* Path-replay adjoint of the camera-path estimator (scenes without point
 * lights): dfilm = d loss / d film (W*H*4 floats, rgb used); the gradients
 * of the listed textures (w*h*4 doubles each) and OM materials (6 doubles
 * each: dn rgb, dk rgb) are added to the caller's arrays. */
int ref_adjoint(ref_scene* r, const omv_renderer_params* p, const omv_camera_desc* cam, int fp64, int threads, const float* dfilm,
                int ntex, const int* tex_ids, double* const* tex_grads, int nom, const int* om_ids, double* om_grads) {
    if (!r || !p || !cam || !dfilm || !r->s.committed) return OMV_E_ARG;
    const Scene& S = r->s;
    if (!S.lights.empty()) return OMV_E_STATE;
    std::vector<int> texT, omT;
    for (int i = 0; i < ntex; ++i) {
        if (!(0 <= tex_ids[i] && tex_ids[i] < (int)S.textures.size()) || !tex_grads || !tex_grads[i]) return OMV_E_RANGE;
        texT.push_back(tex_ids[i]);
    }
    for (int i = 0; i < nom; ++i) {
        if (!(0 <= om_ids[i] && om_ids[i] < (int)S.materials.size()) || S.materials[om_ids[i]].type != MAT_OM || !om_grads) return OMV_E_RANGE;
        omT.push_back(om_ids[i]);
    }
    std::vector<std::vector<double>> texOut(S.textures.size()), omOut(S.materials.size());
    for (int id : texT) texOut[id].assign((size_t)S.textures[id].w * S.textures[id].h * 4, 0.0);
    for (int id : omT) omOut[id].assign(6, 0.0);
    if (fp64) adjointImage<double>(S, *p, *cam, dfilm, texT, omT, threads, texOut, omOut);
    else adjointImage<float>(S, *p, *cam, dfilm, texT, omT, threads, texOut, omOut);
    for (int i = 0; i < ntex; ++i) {
        const std::vector<double>& g = texOut[tex_ids[i]];
        for (size_t j = 0; j < g.size(); ++j) tex_grads[i][j] += g[j];
    }
    for (int i = 0; i < nom; ++i)
        for (int j = 0; j < 6; ++j) om_grads[(size_t)i * 6 + j] += omOut[om_ids[i]][j];
    return OMV_OK;
}

void ref_adjoint_plant(int mask) { gAdjPlant = mask; }
void ref_splat_plant(int mask) { gSplatPlant = mask; }

/* OMV_EXT_HDRI_NEE environment tables of a committed scene (H row entries,
 * W*H column entries and texel probabilities); returns 0 without an HDR */
int ref_env_tables(ref_scene* r, float* rowProb, int* rowAlias, float* colProb, int* colAlias, float* texPdf) {
    if (!r || !r->s.committed || r->s.envRowProb.empty()) return 0;
    const Scene& S = r->s;
    std::memcpy(rowProb, S.envRowProb.data(), S.envRowProb.size() * 4);
    std::memcpy(rowAlias, S.envRowAlias.data(), S.envRowAlias.size() * 4);
    std::memcpy(colProb, S.envColProb.data(), S.envColProb.size() * 4);
    std::memcpy(colAlias, S.envColAlias.data(), S.envColAlias.size() * 4);
    std::memcpy(texPdf, S.envTexPdf.data(), S.envTexPdf.size() * 4);
    return 1;
}
/* one environment sample from six uniforms: texel, direction, density */
int ref_env_sample(ref_scene* r, const float u[6], double dir[3], double* pdf, int* row, int* col) {
    if (!r || !r->s.committed || r->s.envRowProb.empty()) return OMV_E_STATE;
    const Scene& S = r->s;
    omv_renderer_params p;
    ref_renderer_params_default(&p);
    Tracer<double> T(S, p);
    const int W = S.bgW, H = S.bgH;
    int k = std::min((int)(u[0] * (float)H), H - 1);
    *row = u[1] < S.envRowProb[k] ? k : S.envRowAlias[k];
    k = std::min((int)(u[2] * (float)W), W - 1);
    const size_t rk = (size_t)*row * W + k;
    *col = u[3] < S.envColProb[rk] ? k : S.envColAlias[rk];
    double sinT;
    V3<double> w = T.envDirection(*row, *col, u[4], u[5], sinT);
    dir[0] = w.x; dir[1] = w.y; dir[2] = w.z;
    *pdf = sinT > 0 ? (double)S.envTexPdf[(size_t)*row * W + *col] * W * H / (2 * kPi<double>() * kPi<double>() * sinT) : 0.0;
    return OMV_OK;
}
/* solid-angle density of the environment sampler at a direction */
double ref_env_pdf(ref_scene* r, const double dir[3]) {
    if (!r || !r->s.committed) return 0.0;
    omv_renderer_params p;
    ref_renderer_params_default(&p);
    Tracer<double> T(r->s, p);
    return T.envPdf(mk(dir[0], dir[1], dir[2]));
}

/* d(Rp, Rs)/d(n, k) in double: out = {dRp/dn, dRp/dk, dRs/dn, dRs/dk} */
void ref_fresnel_grad(double n, double k, double c, double out[4]) {
    fresnelGrad<double>(n, k, c, out[0], out[1], out[2], out[3]);
}
void ref_fresnel64(double n, double k, double c, double out[2]) { fresnel<double>(n, k, c, out[0], out[1]); }

int ref_render_samples(ref_scene* r, const omv_renderer_params* p, const omv_camera_desc* cam, int fp64, int row0, int rows,
                       int threads, float* rad) {
    if (!r || !p || !cam || !r->s.committed || !rad) return OMV_E_ARG;
    if (rows <= 0) { row0 = 0; rows = cam->height; }
    if (fp64) renderRows<double>(r->s, *p, *cam, row0, rows, threads, nullptr, rad, nullptr);
    else renderRows<float>(r->s, *p, *cam, row0, rows, threads, nullptr, rad, nullptr);
    return OMV_OK;
}

/* Colour table of an enumerated false-colour subject: EnumeratedFCSubject's
 * construction (FCSubject.h:153-167) -- `size` distinct colours drawn
 * uniformly from [1, 0xFFFFFF] into a std::set (ascending), then shuffled --
 * with the reference's unseeded draws and std::random_shuffle replaced by one
 * splitmix64 sequence (Steele, Lea, Flood 2014) started at
 * (seed << 8 | subject): v = 1 + x mod 0xFFFFFF until size distinct values,
 * then for i = size down to 2: swap(t[i-1], t[x mod i]). */
static bool oracleFcTable(uint32_t seed, int subject, uint64_t size, std::vector<uint32_t>& out) {
    out.clear();
    if (size > 0xFFFFFFu) return false;
    uint64_t st = ((uint64_t)seed << 8) | (uint64_t)(subject & 0xFF);
    auto next = [&st]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    std::set<uint32_t> colours;
    while (colours.size() < size) colours.insert(1u + (uint32_t)(next() % 0xFFFFFFull));
    out.assign(colours.begin(), colours.end());
    for (uint64_t i = size; i > 1; --i) std::swap(out[i - 1], out[next() % i]);
    return true;
}

/* category sizes of RenderScene::setFalseColorSubject (RenderScene.cpp:130-146)
 * plus the instance-aware iid (instances) and ipid (faces over instances) */
static uint64_t oracleFcCategory(const Scene& S, std::vector<uint32_t>& instBase) {
    instBase.clear();
    switch (S.fcSubject) {
        case OMV_FC_PID: return (uint64_t)S.faceBase.back();
        case OMV_FC_GID: return S.meshes.size();
        case OMV_FC_MID: return S.materials.size() - 1;
        case OMV_FC_IID: return S.instances.size();
        case OMV_FC_IPID: {
            uint64_t n = 0;
            for (const Instance& in : S.instances) {
                instBase.push_back((uint32_t)n);
                if (in.meshId >= 0 && in.meshId < (int)S.meshes.size()) n += (uint64_t)S.meshes[in.meshId].fn;
            }
            return n;
        }
        default: return 0;
    }
}

int ref_fc_color_table(uint32_t seed, int subject, uint32_t size, uint32_t* out) {
    if (size && !out) return OMV_E_ARG;
    std::vector<uint32_t> t;
    if (!oracleFcTable(seed, subject, size, t)) return OMV_E_RANGE;
    for (uint32_t i = 0; i < size; ++i) out[i] = t[i];
    return OMV_OK;
}

/* FC mode (main.cpp:261-270): spp = 1, one camera bounce, no lights, no background */
int ref_render_fc(ref_scene* r, const omv_renderer_params* p, const omv_camera_desc* cam, float* film, int32_t* ids, uint8_t* tie) {
    if (!r || !p || !cam || !r->s.committed) return OMV_E_ARG;
    const Scene& S = r->s;
    const int W = cam->width, H = cam->height;
    Tracer<float> T(S, *p);
    {
        const uint64_t n = oracleFcCategory(S, T.fcInstBase);
        if (!oracleFcTable(p->seed, S.fcSubject, n, T.fcTable)) return OMV_E_RANGE;
    }
    std::vector<Vtx<float>> cp(3);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int id = y * W + x;
            const uint64_t s = (uint64_t)id;
            Stream cs(p->seed, STREAM_CAMERA, s);
            Uniforms U; U.s = &cs;
            float u0 = U.next(), u1 = U.next();
            V3<float> O, D;
            cameraRay<float>(*cam, x, y, u0, u1, O, D);
            V3<float> col = mk(0.0f, 0.0f, 0.0f), rad = mk(0.0f, 0.0f, 0.0f);
            /* one bounce of computePath in FC mode; also record the hit ids */
            float delta = 3.0f * feps<float>() * maxAbs(O);
            V3<float> org = mk(O.x + delta * D.x, O.y + delta * D.y, O.z + delta * D.z);
            float tnear = 0.0f;
            SurfacePoint<float> sp;
            int32_t i0 = -1, i1 = -1, i2 = -1, i3 = -1;
            uint8_t tflag = 0;
            {
                /* sceneIntersect, keeping the tie flag of the accepted hit */
                const float t1 = std::numeric_limits<float>::max();
                float tfar = t1;
                int itr = 10;
                bool ok = false;
                while (itr > 0) {
                    Hit<float> h;
                    if (!closestHit(S, org, D, tnear, tfar, true, h, T.cnt)) break;
                    computeSurfacePoint(S, h, sp);
                    tflag = h.tie ? 1 : 0;
                    if (!isMasked(S, sp)) { ok = true; break; }
                    V3<float> o2 = mk(org.x + h.t * D.x, org.y + h.t * D.y, org.z + h.t * D.z);
                    float v = maxAbs(o2); v = v > h.t ? v : h.t;
                    tnear = h.t + 3.0f * feps<float>() * v;
                    tfar = t1;
                    if (tfar <= tnear) break;
                    --itr;
                }
                if (ok) {
                    modifyFrame(S, sp);
                    T.fcPaint(sp, D, col, U);
                    rad = col;
                    const Mesh& m = S.meshes[sp.mesh];
                    i0 = sp.globalFace; i1 = sp.mesh; i2 = S.materials[m.matId].type == MAT_DIFFUSE ? -1 : m.matId; i3 = sp.inst;
                } else tflag = 0;
            }
            if (film) { float* f = film + (size_t)id * 4; f[0] = rad.x; f[1] = rad.y; f[2] = rad.z; f[3] = 1.0f; }
            if (ids) { ids[id] = i0; ids[(size_t)W * H + id] = i1; ids[(size_t)2 * W * H + id] = i2; ids[(size_t)3 * W * H + id] = i3; }
            if (tie) tie[id] = tflag;
        }
    return OMV_OK;
}

int ref_closest_hit(ref_scene* r, const float org[3], const float dir[3], float tnear, float tfar, int bvh, float* t, float* u,
                    float* v, int* inst, int* prim, int* tie, uint64_t* nodes, uint64_t* tris) {
    if (!r || !r->s.committed) return OMV_E_ARG;
    Hit<float> h; Counters c;
    bool ok = closestHit(r->s, ld3<float>(org), ld3<float>(dir), tnear, tfar, bvh != 0, h, c);
    if (nodes) *nodes = c.nodes;
    if (tris) *tris = c.tris;
    if (!ok) return 0;
    *t = h.t; *u = h.u; *v = h.v; *inst = h.inst; *prim = h.prim; *tie = h.tie;
    return 1;
}

/* ---- false-colour paint rules and planted mistakes (pin tests) */
void ref_pin_plant(int mask) { gPinPlant = mask; }
void ref_fc_paint_clamped(int subject, float value, float out[3]) {
    V3<float> c = fcPaintClamped<float>(subject, value);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
void ref_fc_paint_enumerated(uint32_t id, const uint32_t* table, uint32_t size, int mid_style, float out[3]) {
    V3<float> c = fcPaintEnumerated<float>(id, table, size, mid_style != 0);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
/* kind: 0 OpenMaterial, 1 PBR, 2 the missing material's diffuse */
void ref_fc_paint_material(int kind, float out[3]) {
    V3<float> c = fcPaintMaterialName<float>(kind == 0 ? MAT_OM : kind == 1 ? MAT_PBR : MAT_DIFFUSE);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
void ref_fc_paint_inverted(int flag, float out[3]) {
    V3<float> c = fcPaintInverted<float>(flag != 0);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

/* ---- unit-level */
void ref_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }
float ref_u01(uint32_t x) { return u01(x); }
void ref_fresnel(float n, float k, float c, float* rp, float* rs) { fresnel<float>(n, k, c, *rp, *rs); }

void ref_evaluate_direct(const float I[3], const float O[3], const float N[3], const float color[4], float metal, float alpha, float out[3]) {
    Bsdf<float> b;
    b.I = ld3<float>(I); b.O = ld3<float>(O); b.N = ld3<float>(N);
    for (int i = 0; i < 4; ++i) b.color[i] = color[i];
    b.metal = metal; b.alpha = alpha;
    V3<float> r = evaluateDirect(b);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

int ref_sample_indirect(const float I[3], const float N[3], const float T[3], const float color[4], float metal, float alpha,
                        float eta, int transmissive, const float* uniforms, float brdf[3], float O[3]) {
    Bsdf<float> b;
    b.I = ld3<float>(I); b.N = ld3<float>(N); b.T = ld3<float>(T);
    for (int i = 0; i < 4; ++i) b.color[i] = color[i];
    b.metal = metal; b.alpha = alpha; b.eta = eta; b.transmissive = transmissive != 0;
    Uniforms U; U.list = uniforms; U.n = 1 << 30;
    V3<float> r = sampleIndirect(b, U);
    brdf[0] = r.x; brdf[1] = r.y; brdf[2] = r.z;
    O[0] = b.O.x; O[1] = b.O.y; O[2] = b.O.z;
    return U.used;
}

void ref_uniform_hemisphere(float u0, float u1, float out[3]) {
    float l[2] = {u0, u1}; Uniforms U; U.list = l; U.n = 2;
    V3<float> v = uniformHemisphere<float>(U); out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
int ref_uniform_sphere(const float* u, int n, float out[3]) {
    Uniforms U; U.list = u; U.n = n;
    V3<float> v = uniformSphere<float>(U); out[0] = v.x; out[1] = v.y; out[2] = v.z;
    return U.used;
}
void ref_ggx_hemisphere(float u0, float u1, float a2, float out[3]) {
    float l[2] = {u0, u1}; Uniforms U; U.list = l; U.n = 2;
    V3<float> v = ggxHemisphere<float>(U, a2); out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
void ref_cosine_sqrt_hemisphere(float u0, float u1, float out[3]) {
    float l[2] = {u0, u1}; Uniforms U; U.list = l; U.n = 2;
    V3<float> v = cosineSqrtHemisphere<float>(U); out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
void ref_texture_rgba8(int w, int h, const uint8_t* rgba, float u, float v, float out[4]) {
    Texture t; t.w = w; t.h = h; if (rgba) t.rgba.assign(rgba, rgba + (size_t)w * h * 4);
    texLookup<float>(t, u, v, out);
}
void ref_hdr_radiance(int w, int h, const float* rgb, const float rot[9], float scale, const float dir[3], float out[3]) {
    Scene S; S.bgType = 2; S.bgW = w; S.bgH = h; S.bgRgb.assign(rgb, rgb + (size_t)w * h * 3);
    std::memcpy(S.bgRot, rot, sizeof(float) * 9); S.bgScale = scale;
    omv_renderer_params p; ref_renderer_params_default(&p);
    Tracer<float> T(S, p);
    V3<float> v = T.background(ld3<float>(dir)); out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
void ref_camera_ray(const omv_camera_desc* c, int x, int y, float u0, float u1, float org[3], float dir[3]) {
    V3<float> O, D; cameraRay<float>(*c, x, y, u0, u1, O, D);
    org[0] = O.x; org[1] = O.y; org[2] = O.z; dir[0] = D.x; dir[1] = D.y; dir[2] = D.z;
}
int ref_distance_random(int n, const float* w, float u, float* pdf) {
    std::vector<float> cdf; float acc = 0.0f;
    for (int i = 0; i < n; ++i) { acc = acc + w[i]; cdf.push_back(acc); }
    return Tracer<float>::distanceRandom(n, cdf.data(), n ? cdf.back() : 0.0f, u, pdf);
}
int ref_get_ior(int nt, const omv_ior_table* t, float T, float wl, float* n, float* k) {
    std::vector<IorTable> tabs = makeTables(nt, t);
    return getIor(tabs, T, wl, *n, *k);
}
/* SurfacePoint::applyTextureNormal on the frame f = (tangent, binormal,
 * normal, handedness) applied `calls` times with the same tn; out = the
 * tangent, binormal and normal after it */
void ref_apply_texture_normal(const float* f, const float* tn, int calls, float* out) {
    SurfacePoint<float> sp;
    sp.tangent = mk(f[0], f[1], f[2]);
    sp.binormal = mk(f[3], f[4], f[5]);
    sp.normal = mk(f[6], f[7], f[8]);
    sp.tw = f[9];
    for (int i = 0; i < calls; ++i) applyTextureNormal(sp, mk(tn[0], tn[1], tn[2]));
    const V3<float> v[3] = {sp.tangent, sp.binormal, sp.normal};
    for (int i = 0; i < 3; ++i) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
}

/* the normal-map texel (RGBA) to tangent-space normal step of
 * PhysicallyBasedMaterial::modifyFrame */
void ref_texel_normal(const float* texel, float scale, float* out) {
    const V3<float> n = texelNormal(texel, scale);
    out[0] = n.x; out[1] = n.y; out[2] = n.z;
}

/* OpenMaterial::defineNextDirection for a hit on the object-space triangle
 * tri (3 x 3 floats) of an instance with normal matrix nm (3 x 3, row-major)
 * and shading normal `normal` */
void ref_om_next_direction(const float* inV, const float* normal, const float* tri, const float* nm, float* out) {
    const V3<float> tfn = faceNormalOf(mk(tri[0], tri[1], tri[2]), mk(tri[3], tri[4], tri[5]), mk(tri[6], tri[7], tri[8]), nm);
    const V3<float> o = omNextDirectionOf(mk(inV[0], inV[1], inV[2]), mk(normal[0], normal[1], normal[2]), tfn);
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}

void ref_gamma_correction(float* img, int w, int h, float A, float gamma) {
    for (int i = 0; i < w * h; ++i) {
        float* p = img + (size_t)i * 4;
        float inI = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
        float outI = A * std::pow(inI, gamma);
        float f = inI > std::numeric_limits<float>::epsilon() ? outI / inI : 1.0f;
        for (int c = 0; c < 3; ++c) p[c] = f * p[c];
    }
    for (int i = 0; i < w * h; ++i) {
        float* p = img + (size_t)i * 4;
        for (int c = 0; c < 3; ++c) p[c] = clamp01(p[c]);
        p[3] = 1.0f;
    }
}
/* SimpleDenoiserMedian<1> (Denoiser.h:24-68): interior pixels take the
 * per-channel median of their 3x3 window (std::sort, element 4), the border
 * keeps its value, alpha 1 inside */
void ref_median3(float* img, int w, int h) {
    std::vector<float> out(img, img + (size_t)w * h * 4);
    for (int i = 1; i < h - 1; ++i)
        for (int j = 1; j < w - 1; ++j) {
            float win[3][9];
            int id = 0;
            for (int y = i - 1; y <= i + 1; ++y)
                for (int x = j - 1; x <= j + 1; ++x, ++id)
                    for (int c = 0; c < 3; ++c) win[c][id] = img[((size_t)y * w + x) * 4 + c];
            for (int c = 0; c < 3; ++c) {
                std::sort(win[c], win[c] + 9);
                out[((size_t)i * w + j) * 4 + c] = win[c][4];
            }
            out[((size_t)i * w + j) * 4 + 3] = 1.0f;
        }
    std::memcpy(img, out.data(), out.size() * sizeof(float));
}
/* saveImage's conversion (main.cpp:88-93): static_cast<unsigned char>(255.0f * f) */
void ref_to_u8(const float* img, size_t n, uint8_t* out) {
    for (size_t i = 0; i < n; ++i) out[i] = static_cast<unsigned char>(255.0f * img[i]);
}
void ref_compute_camera(int w, int h, float yfov, float* hw, float* hh, float* focus) {
    /* adjustToResolution + setYFoV (PinholeCamera.cpp:156-174) */
    float H = 0.5f * (float)std::max(1, h), W = 0.5f * (float)std::max(1, w);
    *hh = H; *hw = W;
    *focus = H / std::tan(0.5f * yfov * 3.14159265358979323846f / 180.0f);
}

/* the oracle ships its own copy of the default-parameter helper so it does
 * not link the product library */
void ref_renderer_params_default(omv_renderer_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->camera_bounces = 10; p->light_bounces = 10; p->max_path_length = 8; p->samples_per_pixel = 20;
    p->light_distance_attenuation = 1; p->light_scale_value = 100.0f; p->light_min_distance = 0.01f;
    p->ray_cut_pix_value = 0.002f; p->gamma = 0.5f; p->number_of_cores = 16; p->false_color_mode = 0;
    p->seed = 0; p->extensions = 0; p->om_roughness = 0.3f;
}

}  // extern "C"
