// scene.cpp -- omv::RenderScene: the reference's scene builder (set* calls,
// commitScene) with the mesh preprocessing restated, the BVH built and every
// pool uploaded to HBM once at commit.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <utility>

#include "bvh.h"
#include "omv_kernels.h"
#include "render_core.h"

namespace omv {

static thread_local std::string g_lastError;
void setError(const std::string& msg) { g_lastError = msg; }
const char* lastError() { return g_lastError.c_str(); }

void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}
hipError_t DeviceBuffer::reserve(size_t n) {
    if (n <= bytes && ptr) return hipSuccess;
    release();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) { ptr = nullptr; return e; }
    bytes = n;
    return hipSuccess;
}
hipError_t DeviceBuffer::upload(const void* src, size_t n) {
    hipError_t e = reserve(n ? n : 16);
    if (e != hipSuccess) return e;
    if (n) return hipMemcpy(ptr, src, n, hipMemcpyHostToDevice);
    return hipSuccess;
}

HostMesh::HostMesh() {
    for (int i = 0; i < OMV_MAX_UV_CHANNELS; ++i) { texMap[i] = -1; texOffs[i] = 0; texDef[i] = false; }
}

namespace {

struct v3 { float x, y, z; };
inline v3 mk(float a, float b, float c) { return v3{a, b, c}; }
inline float dot(const v3& a, const v3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline v3 cross(const v3& a, const v3& b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float len(const v3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline v3 nrmz(const v3& a) {   // normalize3 (UtilityFunctions.h:201-210)
    float L = len(a);
    if (L != 0.0f) { L = 1.0f / L; return mk(a.x * L, a.y * L, a.z * L); }
    return mk(1.0f, 0.0f, 0.0f);
}

// RenderMesh::computeNormals (RenderMesh.cpp:300-330): unweighted sum of unit
// face normals seeded with (0,0,eps), then normalised.
void computeNormals(HostMesh& m) {
    const float e = std::numeric_limits<float>::epsilon();
    for (int i = 0; i < m.vn; ++i) {
        float* n = &m.V[(size_t)i * m.stride + 3];
        n[0] = 0.0f; n[1] = 0.0f; n[2] = e;
    }
    for (int f = 0; f < m.fn; ++f) {
        const float* p0 = &m.V[(size_t)m.F[3 * f] * m.stride];
        const float* p1 = &m.V[(size_t)m.F[3 * f + 1] * m.stride];
        const float* p2 = &m.V[(size_t)m.F[3 * f + 2] * m.stride];
        v3 N = cross(mk(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]), mk(p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]));
        if (N.x * N.x + N.y * N.y + N.z * N.z > e) N = nrmz(N);
        for (int k = 0; k < 3; ++k) {
            float* n = &m.V[(size_t)m.F[3 * f + k] * m.stride + 3];
            n[0] += N.x; n[1] += N.y; n[2] += N.z;
        }
    }
    for (int i = 0; i < m.vn; ++i) {
        float* n = &m.V[(size_t)i * m.stride + 3];
        v3 u = nrmz(mk(n[0], n[1], n[2]));
        n[0] = u.x; n[1] = u.y; n[2] = u.z;
    }
    m.normalDef = true;
}

// RenderMesh::randomTangent (RenderMesh.cpp:418-432): axis x normal, w = +1.
void axisTangent(const float* n, float* t) {
    v3 nv = mk(n[0], n[1], n[2]), r;
    if (n[2] < 0.95f && n[2] > -0.95f) r = cross(mk(0.0f, 0.0f, 1.0f), nv);
    else if (n[1] < 0.95f && n[1] > -0.95f) r = cross(mk(0.0f, 1.0f, 0.0f), nv);
    else r = cross(mk(1.0f, 0.0f, 0.0f), nv);
    r = nrmz(r);
    t[0] = r.x * 1.0f; t[1] = r.y * 1.0f; t[2] = r.z * 1.0f; t[3] = 1.0f;
}

// RenderMesh::computeTangents(int) (RenderMesh.cpp:332-409): per-face uv
// derivative frames summed per vertex, Gram-Schmidt against the normal,
// handedness from the summed bitangent.
void computeTangentsFromUv(HostMesh& m, int slot) {
    const float e = std::numeric_limits<float>::epsilon();
    std::vector<float> tv((size_t)m.vn * 3, 0.0f), bv((size_t)m.vn * 3, 0.0f);
    const int off = m.texOffs[slot];
    for (int f = 0; f < m.fn; ++f) {
        const uint32_t i0 = m.F[3 * f], i1 = m.F[3 * f + 1], i2 = m.F[3 * f + 2];
        const float* X0 = &m.V[(size_t)i0 * m.stride];
        const float* X1 = &m.V[(size_t)i1 * m.stride];
        const float* X2 = &m.V[(size_t)i2 * m.stride];
        const float d1[3] = {X1[0] - X0[0], X1[1] - X0[1], X1[2] - X0[2]};
        const float d2[3] = {X2[0] - X0[0], X2[1] - X0[1], X2[2] - X0[2]};
        const float s1[2] = {X1[off] - X0[off], X1[off + 1] - X0[off + 1]};
        const float s2[2] = {X2[off] - X0[off], X2[off + 1] - X0[off + 1]};
        const float det = s1[0] * s2[1] - s2[0] * s1[1];
        if (det < e) continue;
        const float r = 1.0f / det;
        const float a00 = r * s2[1], a01 = -r * s2[0], a10 = -r * s1[1], a11 = r * s1[0];
        for (int c = 0; c < 3; ++c) {
            const float ud = d1[c] * a00 + d2[c] * a10;
            const float vd = d1[c] * a01 + d2[c] * a11;
            tv[i0 * 3 + c] += ud; tv[i1 * 3 + c] += ud; tv[i2 * 3 + c] += ud;
            bv[i0 * 3 + c] += vd; bv[i1 * 3 + c] += vd; bv[i2 * 3 + c] += vd;
        }
    }
    for (int i = 0; i < m.vn; ++i) {
        float* t = &m.V[(size_t)i * m.stride + 6];
        const float* n = &m.V[(size_t)i * m.stride + 3];
        const v3 nv = mk(n[0], n[1], n[2]);
        const float proj = dot(nv, mk(tv[3 * i], tv[3 * i + 1], tv[3 * i + 2]));
        t[0] = tv[3 * i] - n[0] * proj;
        t[1] = tv[3 * i + 1] - n[1] * proj;
        t[2] = tv[3 * i + 2] - n[2] * proj;
        const float l = len(mk(t[0], t[1], t[2]));
        if (l <= e) axisTangent(n, t);
        else { t[0] /= l; t[1] /= l; t[2] /= l; }
        const v3 nt = cross(nv, mk(t[0], t[1], t[2]));
        t[3] = dot(nt, mk(bv[3 * i], bv[3 * i + 1], bv[3 * i + 2])) < 0.0f ? -1.0f : 1.0f;
    }
    m.tangentDef = true;
}

// Linear table lookup of Interpolation::_get (Interpolation.h:44-65).
float tableLerp(const std::vector<std::pair<float, float>>& t, float x) {
    int lo = 0, hi = (int)t.size() - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (x < t[mid].first) hi = mid; else lo = mid;
    }
    const float dx = t[hi].first - t[lo].first;
    const float dy = t[hi].second - t[lo].second;
    return t[lo].second + (x - t[lo].first) * dy / dx;
}

// physicalconstants::rgb_wl (Constants.h:53)
const float kRgbWavelength[3] = {6.50000004e-07f, 5.10000007e-07f, 4.40000008e-07f};

// World -> object affine inverse in double, rounded to float (shared
// convention with the CPU oracle so object-space rays are bit-identical).
void affineInverse(const float* M, float* out) {
    const double a = M[0], b = M[1], c = M[2], d = M[4], e = M[5], f = M[6], g = M[8], h = M[9], i = M[10];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    const double r = 1.0 / det;
    const double m[9] = {A * r, -(b * i - c * h) * r, (b * f - c * e) * r,
                         B * r, (a * i - c * g) * r, -(a * f - c * d) * r,
                         C * r, -(a * h - b * g) * r, (a * e - b * d) * r};
    const double t[3] = {M[3], M[7], M[11]};
    for (int row = 0; row < 3; ++row) {
        for (int col = 0; col < 3; ++col) out[row * 4 + col] = (float)m[row * 3 + col];
        out[row * 4 + 3] = (float)(-(m[row * 3 + 0] * t[0] + m[row * 3 + 1] * t[1] + m[row * 3 + 2] * t[2]));
    }
}

int hipFail(hipError_t e, const char* what) {
    setError(std::string(what) + ": " + hipGetErrorString(e));
    return OMV_E_DEVICE;
}

}  // namespace

int RenderScene::allocate(int meshes, int instances, int materials, int textures, int lights) {
    if (meshes < 0 || instances < 0 || materials < 0 || textures < 0 || lights < 0) {
        setError("allocate: negative count");
        return OMV_E_ARG;
    }
    meshes_.assign(meshes, HostMesh());
    insts_.assign(instances, HostInstance());
    mats_.assign(materials + 1, OmvMaterial());
    for (auto& m : mats_) { std::memset(&m, 0, sizeof(m)); m.type = OMV_MAT_NONE; }
    // the missing material: DiffuseColorMaterial (1000, 0, 1000, 1) (RenderScene.cpp:65-68)
    OmvMaterial& miss = mats_.back();
    miss.type = OMV_MAT_DIFFUSE;
    miss.diffuse[0] = 1000.0f; miss.diffuse[1] = 0.0f; miss.diffuse[2] = 1000.0f; miss.diffuse[3] = 1.0f;
    for (int s = 0; s < 5; ++s) miss.tex_map[s] = -1;
    texs_.assign(textures, HostTexture());
    lights_.assign(lights, OmvLight());
    lightSet_.assign(lights, false);
    bgType_ = OMV_BG_NONE;
    fcSubject_ = OMV_FC_NONE;
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setMesh(int meshId, int vertN, int facesN, int texN, int matId, const uint32_t* faces, const float* vert,
                         int vstride, const float* norm, int nstride, const float* tans, int tanstride, const float* const* tex,
                         int texstride) {
    if (meshId < 0 || meshId >= (int)meshes_.size()) { setError("setMesh: mesh id out of range"); return OMV_E_RANGE; }
    if (vertN < 0 || facesN < 0 || texN < 0 || (facesN > 0 && !faces)) { setError("setMesh: bad arguments"); return OMV_E_ARG; }
    HostMesh m;
    int used = 0;
    for (int i = 0; i < texN; ++i) used += (tex && tex[i]) ? 1 : 0;
    m.id = meshId; m.vn = vertN; m.fn = facesN;
    m.texChN = std::min(used, OMV_MAX_UV_CHANNELS);
    m.stride = 10 + 2 * m.texChN;
    m.F.assign(faces, faces + (size_t)3 * facesN);
    for (uint32_t idx : m.F)
        if ((int64_t)idx >= (int64_t)vertN) { setError("setMesh: face index out of range"); return OMV_E_ARG; }
    m.V.assign((size_t)vertN * m.stride, 0.0f);
    for (int i = 0; i < m.texChN; ++i) m.texOffs[i] = 10 + 2 * i;
    m.matId = matId;
    if (vert) {
        m.vertexDef = true;
        for (int i = 0; i < vertN; ++i) std::memcpy(&m.V[(size_t)i * m.stride], vert + (size_t)i * vstride, 3 * sizeof(float));
    }
    if (norm) {
        m.normalDef = true;
        for (int i = 0; i < vertN; ++i) std::memcpy(&m.V[(size_t)i * m.stride + 3], norm + (size_t)i * nstride, 3 * sizeof(float));
    }
    if (tans) {
        m.tangentDef = true;
        for (int i = 0; i < vertN; ++i) std::memcpy(&m.V[(size_t)i * m.stride + 6], tans + (size_t)i * tanstride, 4 * sizeof(float));
    }
    // RenderMesh::setTextureCoordinates (RenderMesh.cpp:130-138): channels fill slots in call order
    for (int ch = 0; ch < texN && ch < OMV_MAX_UV_CHANNELS; ++ch) {
        if (!(tex && tex[ch])) continue;
        if (m.texMap[ch] == -1) m.texMap[ch] = m.filledTexChN++;
        const int slot = m.texMap[ch];
        if (slot >= m.texChN) continue;
        m.texDef[slot] = true;
        for (int i = 0; i < vertN; ++i) std::memcpy(&m.V[(size_t)i * m.stride + m.texOffs[slot]], tex[ch] + (size_t)i * texstride, 2 * sizeof(float));
    }
    meshes_[meshId] = std::move(m);
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setInstance(int instId, const float trans[16], int meshId) {
    if (instId < 0 || instId >= (int)insts_.size()) { setError("setInstance: instance id out of range"); return OMV_E_RANGE; }
    if (!trans) { setError("setInstance: null transform"); return OMV_E_ARG; }
    HostInstance& in = insts_[instId];
    in.id = instId; in.meshId = meshId;
    std::memcpy(in.tm, trans, sizeof(in.tm));
    // RenderInstance::computeNormalMatrix + inverseMatrix3 (RenderMesh.h:36-43, UtilityFunctions.h:342-349)
    float M[3][3];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M[i][j] = trans[j * 4 + i];
    const float det = M[0][0] * M[1][1] * M[2][2] + M[0][1] * M[1][2] * M[2][0] + M[0][2] * M[1][0] * M[2][1] -
                      M[0][2] * M[1][1] * M[2][0] - M[0][1] * M[1][0] * M[2][2] - M[0][0] * M[1][2] * M[2][1];
    const float k = (float)(1.0 / det);
    float* R = in.nrm;
    R[0] = k * (M[1][1] * M[2][2] - M[1][2] * M[2][1]);
    R[1] = k * (-M[0][1] * M[2][2] + M[0][2] * M[2][1]);
    R[2] = k * (M[0][1] * M[1][2] - M[0][2] * M[1][1]);
    R[3] = k * (-M[1][0] * M[2][2] + M[1][2] * M[2][0]);
    R[4] = k * (M[0][0] * M[2][2] - M[0][2] * M[2][0]);
    R[5] = k * (-M[0][0] * M[1][2] + M[0][2] * M[1][0]);
    R[6] = k * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    R[7] = k * (-M[0][0] * M[2][1] + M[0][1] * M[2][0]);
    R[8] = k * (M[0][0] * M[1][1] - M[0][1] * M[1][0]);
    affineInverse(trans, in.inv);
    committed_ = false;
    return OMV_OK;
}

static bool matIdOk(int id, size_t n) { return id >= 0 && id < (int)n - 1; }

static void pbrDefaults(OmvMaterial& m) {
    for (int s = 0; s < 5; ++s) {
        m.tex_map[s] = -1; m.tex_ch[s] = 0;
        for (int k = 0; k < 9; ++k) m.tex_xf[s][k] = (k % 4 == 0) ? 1.0f : 0.0f;
    }
}

int RenderScene::setMaterialOm(int matId, float temperature, int nTables, const omv_ior_table* tables) {
    if (!matIdOk(matId, mats_.size())) { setError("setMaterial: material id out of range"); return OMV_E_RANGE; }
    if (nTables <= 0 || !tables) { setError("setMaterial: no IOR table"); return OMV_E_ARG; }
    // AssetMaterialIor::findClosest + getIor (AssetMaterialIor.cpp:115-178), baked per RGB wavelength
    std::vector<int> idx(nTables);
    for (int i = 0; i < nTables; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return tables[a].temperature < tables[b].temperature; });
    int best = idx[0];
    float delta = std::fabs(tables[best].temperature - temperature);
    for (int i = 1; i < nTables; ++i) {
        const float di = std::fabs(tables[idx[i]].temperature - temperature);
        if (di < delta) { best = idx[i]; delta = di; }
    }
    const omv_ior_table& T = tables[best];
    std::vector<std::pair<float, float>> tn, tk;
    for (int i = 0; i < T.n_count; ++i) tn.emplace_back(T.n_wavelength[i], T.n_value[i]);
    for (int i = 0; i < T.k_count; ++i) tk.emplace_back(T.k_wavelength[i], T.k_value[i]);
    std::sort(tn.begin(), tn.end());
    std::sort(tk.begin(), tk.end());
    if (tn.size() < 2 || tk.size() < 2) { setError("setMaterial: IOR table needs two points"); return OMV_E_ARG; }
    const float lo = std::max(tn.front().first, tk.front().first), hi = std::min(tn.back().first, tk.back().first);
    OmvMaterial m;
    std::memset(&m, 0, sizeof(m));
    pbrDefaults(m);
    m.type = OMV_MAT_OM;
    m.has_base = 1;
    for (int c = 0; c < 3; ++c) {
        const float wl = kRgbWavelength[c];
        if (!(lo <= wl && wl <= hi)) {
            setError("setMaterial: no IOR available for an RGB wavelength (the reference throws here)");
            return OMV_E_ARG;
        }
        m.n[c] = tableLerp(tn, wl);
        m.k[c] = tableLerp(tk, wl);
    }
    mats_[matId] = m;
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setMaterialOmNk(int matId, const float n[3], const float k[3]) {
    if (!matIdOk(matId, mats_.size())) { setError("setMaterial: material id out of range"); return OMV_E_RANGE; }
    OmvMaterial m;
    std::memset(&m, 0, sizeof(m));
    pbrDefaults(m);
    m.type = OMV_MAT_OM;
    m.has_base = 1;
    for (int c = 0; c < 3; ++c) { m.n[c] = n[c]; m.k[c] = k[c]; }
    mats_[matId] = m;
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setMaterialPbr(int matId, const omv_pbr_desc& d) {
    if (!matIdOk(matId, mats_.size())) { setError("setMaterial: material id out of range"); return OMV_E_RANGE; }
    OmvMaterial m;
    std::memset(&m, 0, sizeof(m));
    m.type = OMV_MAT_PBR;
    m.alpha_mode = d.alpha_mode;
    m.double_sided = d.double_sided ? 1 : 0;
    for (int i = 0; i < 4; ++i) m.bcf[i] = d.base_color_factor[i];
    for (int i = 0; i < 3; ++i) m.ef[i] = d.emissive_factor[i];
    m.metal = d.metallic_factor; m.rough = d.roughness_factor; m.normal_scale = d.normal_scale;
    m.occ_strength = d.occlusion_strength; m.ior = d.ior; m.alpha_cutoff = d.alpha_cutoff;
    for (int s = 0; s < 5; ++s) {
        m.tex_map[s] = d.texture_map[s];
        m.tex_ch[s] = d.texture_channel[s];
        std::memcpy(m.tex_xf[s], d.texture_transform[s], 9 * sizeof(float));
    }
    // PbrMaterial::isOcclusionSeparateMap (PbrMaterial.cpp:318-321)
    m.occ_separate = d.texture_map[OMV_PBR_OCCLUSION] != d.texture_map[OMV_PBR_METALLIC_ROUGHNESS];
    mats_[matId] = m;
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setTexture(int texId, int w, int h, const uint8_t* rgba) {
    if (texId < 0 || texId >= (int)texs_.size()) { setError("setTexture: texture id out of range"); return OMV_E_RANGE; }
    HostTexture t;
    t.fmt = OMV_TEX_RGBA8;
    if (rgba && w > 0 && h > 0) { t.w = w; t.h = h; t.bytes.assign(rgba, rgba + (size_t)w * h * 4); }
    texs_[texId] = std::move(t);
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setTextureF32(int texId, int w, int h, const float* rgba) {
    if (texId < 0 || texId >= (int)texs_.size()) { setError("setTexture: texture id out of range"); return OMV_E_RANGE; }
    HostTexture t;
    t.fmt = OMV_TEX_RGBA32F;
    if (rgba && w > 0 && h > 0) {
        t.w = w; t.h = h;
        const uint8_t* p = reinterpret_cast<const uint8_t*>(rgba);
        t.bytes.assign(p, p + (size_t)w * h * 16);
    }
    texs_[texId] = std::move(t);
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setLightPoint(int lightId, const float pos[3], const float intensity[3], float range) {
    if (lightId < 0 || lightId >= (int)lights_.size()) { setError("setLight: light id out of range"); return OMV_E_RANGE; }
    OmvLight& L = lights_[lightId];
    for (int c = 0; c < 3; ++c) { L.pos[c] = pos[c]; L.intensity[c] = intensity[c]; }
    L.range = range < 0.0f ? std::numeric_limits<float>::max() : range;
    L.power = 0.2126f * L.intensity[0] + 0.7152f * L.intensity[1] + 0.0722f * L.intensity[2];  // RenderLightPoint.cpp:25
    lightSet_[lightId] = true;
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setBackgroundHdr(int w, int h, const float* rgb, const float rot[9], float scale) {
    if (w <= 0 || h <= 0 || !rgb) { setError("setBackground: bad image"); return OMV_E_ARG; }
    bgType_ = OMV_BG_HDR; bgW_ = w; bgH_ = h;
    bgRgb_.assign((size_t)w * h * 4, 0.0f);
    for (size_t i = 0; i < (size_t)w * h; ++i) for (int c = 0; c < 3; ++c) bgRgb_[i * 4 + c] = rgb[i * 3 + c];
    if (rot) std::memcpy(bgRot_, rot, sizeof(bgRot_));
    bgScale_ = scale;
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setBackgroundConstant(const float rgb[3]) {
    bgType_ = OMV_BG_CONSTANT;
    for (int c = 0; c < 3; ++c) bgConst_[c] = rgb[c];
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setFalseColorSubject(const char* name) {
    static const char* names[] = {"", "pid", "gid", "mid", "rmp", "mmp", "md", "sg", "mn", "in", "iid", "ipid"};
    fcSubject_ = OMV_FC_NONE;   // unknown names get the empty subject (FCHandler.cpp:87-95)
    for (int i = 1; i < 12; ++i) if (std::strcmp(name, names[i]) == 0) fcSubject_ = i;
    committed_ = false;
    return OMV_OK;
}

int RenderScene::setBuildMode(int mode) {
    if (mode != OMV_BUILD_HOST_SAH && mode != OMV_BUILD_DEVICE_LBVH) { setError("set_build_mode: unknown mode"); return OMV_E_ARG; }
    buildMode_ = mode;
    committed_ = false;
    return OMV_OK;
}

// Category sizes of RenderScene::setFalseColorSubject (RenderScene.cpp:
// 130-146): pid = the faces of all meshes, gid = meshes, mid = allocated
// materials; iid = instances and ipid = faces over all instances (the
// instance-aware subjects).
uint64_t RenderScene::fcCategory(int subject, std::vector<uint32_t>& instBase) const {
    instBase.clear();
    switch (subject) {
        case OMV_FC_PID: {
            uint64_t n = 0;
            for (const HostMesh& m : meshes_) n += (uint64_t)m.fn;
            return n;
        }
        case OMV_FC_GID: return meshes_.size();
        case OMV_FC_MID: return mats_.size() - 1;
        case OMV_FC_IID: return insts_.size();
        case OMV_FC_IPID: {
            uint64_t n = 0;
            for (const HostInstance& in : insts_) {
                instBase.push_back((uint32_t)std::min<uint64_t>(n, 0xFFFFFFFFu));
                if (in.meshId >= 0 && in.meshId < (int)meshes_.size() && meshes_[in.meshId].valid) n += (uint64_t)meshes_[in.meshId].fn;
            }
            return n;
        }
        default: return 0;
    }
}

// RenderScene::commitScene (RenderScene.cpp:155-223)
int RenderScene::commit(std::string& log) {
    const int M = (int)mats_.size();
    std::vector<char> prepNormals(meshes_.size(), 0), prepTangents(meshes_.size(), 0);
    std::vector<int> prepUv(meshes_.size(), -1);
    devPrepared_.assign(meshes_.size(), 0);
    devVtxOffset_.assign(meshes_.size(), 0);
    for (int i = 0; i < (int)meshes_.size(); ++i) {
        HostMesh& m = meshes_[i];
        int normCh = -1;
        if (m.matId >= 0 && m.matId < M) {
            const OmvMaterial& mat = mats_[m.matId];
            if (mat.type != OMV_MAT_NONE) {
                if (mat.type == OMV_MAT_PBR) normCh = mat.tex_ch[OMV_PBR_NORMAL];
            } else {
                log += std::to_string(i) + " mesh refers to undefined material" + std::to_string(m.matId) + "\n";
                m.matId = M - 1;
            }
        } else {
            m.matId = M - 1;
        }
        // RenderMesh::commitMesh (RenderMesh.cpp:246-265)
        m.valid = m.vn > 0 && m.fn > 0 && m.vertexDef;
        if (!m.valid) {
            log += std::to_string(i) + " mesh is inconsistent\n";
            const int keep = m.matId;
            m = HostMesh();
            m.matId = keep;
            continue;
        }
        if (buildMode_ == OMV_BUILD_DEVICE_LBVH) {   // normals / tangents on the device after the upload
            const int slot = (normCh >= 0 && normCh < OMV_MAX_UV_CHANNELS && m.texMap[normCh] >= 0) ? m.texMap[normCh] : -1;
            prepNormals[i] = !m.normalDef;
            prepTangents[i] = !m.tangentDef;
            prepUv[i] = (slot >= 0 && m.texDef[slot]) ? m.texOffs[slot] : -1;
            m.normalDef = m.tangentDef = true;
            continue;
        }
        if (!m.normalDef) computeNormals(m);
        if (!m.tangentDef) {
            const int slot = (normCh >= 0 && normCh < OMV_MAX_UV_CHANNELS && m.texMap[normCh] >= 0) ? m.texMap[normCh] : -1;
            if (slot >= 0 && m.texDef[slot]) computeTangentsFromUv(m, slot);
            else {
                for (int v = 0; v < m.vn; ++v) axisTangent(&m.V[(size_t)v * m.stride + 3], &m.V[(size_t)v * m.stride + 6]);
                m.tangentDef = true;
            }
        }
    }
    int invalid = 0;
    for (auto& in : insts_) {
        if (in.meshId < 0 || in.meshId >= (int)meshes_.size() || !meshes_[in.meshId].valid) { in.meshId = -1; ++invalid; }
    }
    if (invalid) log += std::to_string(invalid) + " from " + std::to_string(insts_.size()) + " instances are invalid\n";
    if (invalid == (int)insts_.size()) { setError("commit: nothing to render"); return OMV_E_SCENE; }
    for (size_t i = 0; i < lights_.size(); ++i)
        if (!lightSet_[i]) { setError("commit: light " + std::to_string(i) + " was never set"); return OMV_E_STATE; }

    // FC density range (RenderScene.cpp:196-203, RenderMesh.cpp:277-297)
    float minD = std::numeric_limits<float>::infinity(), maxD = 0.0f;
    for (const auto& in : insts_) {
        if (in.meshId < 0) continue;
        const HostMesh& m = meshes_[in.meshId];
        const float* T = in.tm;
        for (int f = 0; f < m.fn; ++f) {
            v3 P[3];
            for (int k = 0; k < 3; ++k) {
                const float* p = &m.V[(size_t)m.F[3 * f + k] * m.stride];
                P[k] = mk(T[0] * p[0] + T[1] * p[1] + T[2] * p[2] + T[3], T[4] * p[0] + T[5] * p[1] + T[6] * p[2] + T[7],
                          T[8] * p[0] + T[9] * p[1] + T[10] * p[2] + T[11]);
            }
            const float a = len(cross(mk(P[1].x - P[0].x, P[1].y - P[0].y, P[1].z - P[0].z),
                                      mk(P[2].x - P[0].x, P[2].y - P[0].y, P[2].z - P[0].z))) / 2.0f;
            if (a > std::numeric_limits<float>::epsilon()) {
                const float d = 1.0f / (0.0f + a);
                minD = std::min(minD, d);
                maxD = std::max(maxD, d);
            }
        }
    }

    // pools
    std::vector<float> vtx;
    std::vector<uint32_t> faces;
    std::vector<OmvMesh> dmeshes(meshes_.size());
    std::vector<OmvNode4> nodes;
    std::vector<float> tris;
    int faceBase = 0, blasDepth = 0;
    const bool deviceBuild = buildMode_ == OMV_BUILD_DEVICE_LBVH;
    std::vector<int> meshRec(meshes_.size(), -1);
    int recTotal = 0;
    for (size_t i = 0; i < meshes_.size(); ++i) {
        const HostMesh& m = meshes_[i];
        OmvMesh& d = dmeshes[i];
        std::memset(&d, 0, sizeof(d));
        d.vtx_offset = (uint32_t)vtx.size();
        d.stride = (uint32_t)std::max(m.stride, 10);
        d.face_offset = (uint32_t)faces.size();
        d.num_faces = (uint32_t)m.fn;
        d.mat_id = m.matId;
        d.face_base = faceBase;
        d.valid = m.valid ? 1 : 0;
        d.blas_root = -1;
        for (int ch = 0; ch < OMV_MAX_UV_CHANNELS; ++ch) d.tex_off[ch] = (m.texMap[ch] >= 0) ? m.texOffs[m.texMap[ch]] : -1;
        faceBase += m.fn;
        if (!m.valid) continue;
        vtx.insert(vtx.end(), m.V.begin(), m.V.end());
        faces.insert(faces.end(), m.F.begin(), m.F.end());
        if (deviceBuild) {   // built on the GPU below, records in mesh order
            meshRec[i] = recTotal;
            recTotal += m.fn;
            continue;
        }
        std::vector<BuildBox> boxes(m.fn);
        for (int f = 0; f < m.fn; ++f) {
            BuildBox& b = boxes[f];
            for (int a = 0; a < 3; ++a) { b.lo[a] = INFINITY; b.hi[a] = -INFINITY; }
            for (int k = 0; k < 3; ++k) {
                const float* p = &m.V[(size_t)m.F[3 * f + k] * m.stride];
                for (int a = 0; a < 3; ++a) { b.lo[a] = std::min(b.lo[a], p[a]); b.hi[a] = std::max(b.hi[a], p[a]); }
            }
        }
        const int recBase = (int)(tris.size() / 12);
        Bvh4Result r = buildBvh4(boxes, OMV_BVH_LEAF_MAX, recBase, nodes, false);
        d.blas_root = r.root;
        blasDepth = std::max(blasDepth, r.depth);
        for (int prim : r.order) {
            const uint32_t* f = &m.F[(size_t)3 * prim];
            for (int k = 0; k < 3; ++k) {
                const float* p = &m.V[(size_t)f[k] * m.stride];
                tris.push_back(p[0]); tris.push_back(p[1]); tris.push_back(p[2]);
                float w = 0.0f;
                if (k == 0) std::memcpy(&w, &prim, sizeof(w));
                tris.push_back(w);
            }
        }
    }
    if ((deviceBuild ? (size_t)recTotal : tris.size() / 12) >= (1u << 27)) { setError("commit: more than 2^27 triangles"); return OMV_E_SCENE; }
    // device BLAS build: vertex and face pools first, then one build per mesh
    std::vector<OmvNode4> blasRoots(meshes_.size());
    int deviceNodes = 0;
    if (deviceBuild) {
        hipError_t e;
        if ((e = dVtx_.upload(vtx.data(), vtx.size() * 4)) != hipSuccess) return hipFail(e, "upload vertices");
        if ((e = dFaces_.upload(faces.data(), faces.size() * 4)) != hipSuccess) return hipFail(e, "upload faces");
        size_t nodeCap = 0, scratch = 0;
        for (size_t i = 0; i < meshes_.size(); ++i)
            if (meshRec[i] >= 0) {
                nodeCap += (size_t)meshes_[i].fn;
                scratch = std::max(scratch, blasDeviceScratchBytes((uint32_t)meshes_[i].fn));
            }
        const size_t tlasCap = std::max<size_t>(1, insts_.size());
        if ((e = dNodes_.reserve((nodeCap + tlasCap) * sizeof(OmvNode4))) != hipSuccess) return hipFail(e, "BVH nodes");
        if ((e = dTris_.reserve((size_t)recTotal * 12 * 4)) != hipSuccess) return hipFail(e, "triangle records");
        for (size_t i = 0; i < meshes_.size(); ++i)
            if (meshRec[i] >= 0 && (prepNormals[i] || prepTangents[i]))
                scratch = std::max(scratch, meshDeviceScratchBytes((uint32_t)meshes_[i].vn, (uint32_t)meshes_[i].fn));
        DeviceBuffer work;
        if ((e = work.reserve(scratch)) != hipSuccess) return hipFail(e, "BVH build scratch");
        for (size_t i = 0; i < meshes_.size(); ++i) {   // commitMesh's normals / tangents
            if (meshRec[i] < 0 || !(prepNormals[i] || prepTangents[i])) continue;
            const HostMesh& m = meshes_[i];
            uint32_t bad = 0;
            e = preprocessMeshDevice(dVtx_.as<float>() + dmeshes[i].vtx_offset, dmeshes[i].stride, (uint32_t)m.vn,
                                     dFaces_.as<uint32_t>() + dmeshes[i].face_offset, (uint32_t)m.fn, prepNormals[i] != 0,
                                     prepTangents[i] != 0, prepUv[i], work.ptr, nullptr, &bad);
            if (e != hipSuccess) return hipFail(e, "device mesh preprocessing");
            if (bad) { setError("commit: face index out of range"); return OMV_E_INTERNAL; }
            devPrepared_[i] = 1;
            devVtxOffset_[i] = dmeshes[i].vtx_offset;
        }
        for (size_t i = 0; i < meshes_.size(); ++i) {
            if (meshRec[i] < 0) continue;
            const HostMesh& m = meshes_[i];
            int root = -1, count = 0, depth = 0;
            e = buildBlasDevice(dVtx_.as<float>(), dmeshes[i].vtx_offset, dmeshes[i].stride, dFaces_.as<uint32_t>(), dmeshes[i].face_offset,
                                (uint32_t)m.fn, OMV_BVH_LEAF_MAX, meshRec[i], deviceNodes, (uint32_t)m.fn, dNodes_.as<OmvNode4>(),
                                dTris_.as<float>(), work.ptr, nullptr, &root, &count, &depth);
            if (e != hipSuccess) return hipFail(e, "device BVH build");
            dmeshes[i].blas_root = root;
            deviceNodes += count;
            blasDepth = std::max(blasDepth, depth);
            if ((e = hipMemcpy(&blasRoots[i], dNodes_.as<OmvNode4>() + root, sizeof(OmvNode4), hipMemcpyDeviceToHost)) != hipSuccess)
                return hipFail(e, "BVH root readback");
        }
    }
    if (insts_.size() >= (1u << 29)) { setError("commit: too many instances"); return OMV_E_SCENE; }
    // TLAS over world-space instance boxes (transformed BLAS root boxes, rounded outward)
    std::vector<BuildBox> iboxes;
    std::vector<int> iids;
    std::vector<OmvInstance> dinsts(insts_.size());
    for (size_t i = 0; i < insts_.size(); ++i) {
        const HostInstance& in = insts_[i];
        OmvInstance& d = dinsts[i];
        std::memset(&d, 0, sizeof(d));
        std::memcpy(d.xfm, in.tm, 12 * sizeof(float));
        std::memcpy(d.inv, in.inv, sizeof(d.inv));
        std::memcpy(d.nrm, in.nrm, sizeof(d.nrm));
        d.mesh_id = in.meshId;
        d.blas_root = -1;
        if (in.meshId < 0) continue;
        d.blas_root = dmeshes[in.meshId].blas_root;
        const OmvNode4& root = deviceBuild ? blasRoots[in.meshId] : nodes[dmeshes[in.meshId].blas_root];
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 4; ++k) {
            if (root.child[k] == OMV_EMPTY_CHILD) continue;
            lo[0] = std::min(lo[0], root.lox[k]); lo[1] = std::min(lo[1], root.loy[k]); lo[2] = std::min(lo[2], root.loz[k]);
            hi[0] = std::max(hi[0], root.hix[k]); hi[1] = std::max(hi[1], root.hiy[k]); hi[2] = std::max(hi[2], root.hiz[k]);
        }
        double wl[3] = {INFINITY, INFINITY, INFINITY}, wh[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = 0; c < 8; ++c) {
            const double p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
            for (int a = 0; a < 3; ++a) {
                const double w = (double)in.tm[a * 4] * p[0] + (double)in.tm[a * 4 + 1] * p[1] + (double)in.tm[a * 4 + 2] * p[2] + (double)in.tm[a * 4 + 3];
                wl[a] = std::min(wl[a], w);
                wh[a] = std::max(wh[a], w);
            }
        }
        BuildBox b;
        for (int a = 0; a < 3; ++a) {
            const double pad = 1e-5 * (std::fabs(wl[a]) + std::fabs(wh[a])) + 1e-20;
            b.lo[a] = std::nextafter((float)(wl[a] - pad), -INFINITY);
            b.hi[a] = std::nextafter((float)(wh[a] + pad), INFINITY);
        }
        iboxes.push_back(b);
        iids.push_back((int)i);
    }
    Bvh4Result tl = buildBvh4(iboxes, 1, 0, nodes, true);
    float sceneLo[3] = {INFINITY, INFINITY, INFINITY}, sceneHi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const BuildBox& b : iboxes)
        for (int a = 0; a < 3; ++a) { sceneLo[a] = std::min(sceneLo[a], b.lo[a]); sceneHi[a] = std::max(sceneHi[a], b.hi[a]); }
    // TLAS leaves must name instances: the builder wrote -(order index + 1)
    {
        std::vector<int> stack = {tl.root};
        while (!stack.empty()) {
            OmvNode4& nd = nodes[stack.back()];
            stack.pop_back();
            for (int k = 0; k < 4; ++k) {
                if (nd.child[k] == OMV_EMPTY_CHILD) continue;
                if (nd.child[k] >= 0) { stack.push_back(nd.child[k]); continue; }
                const int ord = -(nd.child[k] + 1);
                nd.child[k] = -(iids[tl.order[ord]] + 1);
                nd.entry[k] = OMV_ENT(OMV_ENT_INST, iids[tl.order[ord]]);
            }
        }
    }
    // Light rays take directions in the closed positive octant only
    // (uniformSphere), so a light that lies above every child box of the TLAS
    // root on some axis -- by 2e-30 or more, the generation kernel's own test
    // (lightBelowRoot) -- emits rays that all miss the scene.  When every
    // light does, no light round has rays and the renderer skips them.
    int lightsLeave = 0;
    if (tl.root >= 0 && (size_t)tl.root < nodes.size() && !lights_.empty()) {
        const OmvNode4& r = nodes[tl.root];
        lightsLeave = 1;
        for (const OmvLight& L : lights_)
            for (int k = 0; k < 4; ++k)
                if (r.entry[k] != OMV_ENT_NONE &&
                    !(L.pos[0] - r.hix[k] >= 2e-30f || L.pos[1] - r.hiy[k] >= 2e-30f || L.pos[2] - r.hiz[k] >= 2e-30f))
                    lightsLeave = 0;
    }
    // device build: the TLAS (built above from index 0) goes after the BLAS nodes
    if (deviceBuild) {
        for (OmvNode4& nd : nodes)
            for (int k = 0; k < 4; ++k)
                if (nd.child[k] != OMV_EMPTY_CHILD && nd.child[k] >= 0) {
                    nd.child[k] += deviceNodes;
                    nd.entry[k] = OMV_ENT(OMV_ENT_TLAS_NODE, nd.child[k]);
                }
        tl.root += deviceNodes;
        hipError_t e = hipMemcpy(dNodes_.as<OmvNode4>() + deviceNodes, nodes.data(), nodes.size() * sizeof(OmvNode4), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hipFail(e, "upload TLAS");
    }
    // traversal stack bound: 3 siblings per BVH4 level plus the instance markers
    stackBound_ = 3 * (tl.depth + blasDepth) + 4;
    if (stackBound_ > OMV_STACK_TOTAL) {
        setError("commit: BVH too deep for the traversal stack (" + std::to_string(stackBound_) + " entries)");
        return OMV_E_SCENE;
    }
    // the traversal addresses a node's plane rows by 32-bit byte offsets (node * 128 + row)
    const size_t totalNodes = (deviceBuild ? (size_t)deviceNodes : 0) + nodes.size();
    if (totalNodes > OMV_MAX_NODES) {
        setError("commit: BVH of " + std::to_string(totalNodes) + " nodes exceeds the " + std::to_string(OMV_MAX_NODES) + "-node limit");
        return OMV_E_SCENE;
    }
    // light CDF (RenderScene.cpp:213-220, DistanceRandom::calculate)
    std::vector<float> cdf;
    float acc = 0.0f;
    for (const auto& L : lights_) { acc = acc + L.power; cdf.push_back(acc); }
    // emission reachability per material
    int anyEmissive = 0;
    for (auto& m : mats_) {
        if (m.type != OMV_MAT_PBR) continue;
        const float el = std::sqrt(m.ef[0] * m.ef[0] + m.ef[1] * m.ef[1] + m.ef[2] * m.ef[2]);
        const int id = m.tex_map[OMV_PBR_EMISSIVITY];
        m.emissive = (el > std::numeric_limits<float>::epsilon() && id >= 0 && id < (int)texs_.size()) ? 1 : 0;
        anyEmissive |= m.emissive;
    }
    // textures
    std::vector<OmvTexture> dtex(texs_.size());
    std::vector<uint8_t> texels;
    for (size_t i = 0; i < texs_.size(); ++i) {
        const HostTexture& t = texs_[i];
        while (texels.size() % 16) texels.push_back(0);
        dtex[i].w = t.bytes.empty() ? 0 : t.w;
        dtex[i].h = t.bytes.empty() ? 0 : t.h;
        dtex[i].fmt = t.fmt;
        dtex[i].offset = texels.size();
        texels.insert(texels.end(), t.bytes.begin(), t.bytes.end());
    }
    texOffsets_.resize(dtex.size());
    for (size_t i = 0; i < dtex.size(); ++i) texOffsets_[i] = dtex[i].offset;
    hipError_t e;
    if (!deviceBuild) {
        if ((e = dVtx_.upload(vtx.data(), vtx.size() * 4)) != hipSuccess) return hipFail(e, "upload vertices");
        if ((e = dFaces_.upload(faces.data(), faces.size() * 4)) != hipSuccess) return hipFail(e, "upload faces");
    }
    if ((e = dMeshes_.upload(dmeshes.data(), dmeshes.size() * sizeof(OmvMesh))) != hipSuccess) return hipFail(e, "upload meshes");
    if ((e = dInsts_.upload(dinsts.data(), dinsts.size() * sizeof(OmvInstance))) != hipSuccess) return hipFail(e, "upload instances");
    if ((e = dMats_.upload(mats_.data(), mats_.size() * sizeof(OmvMaterial))) != hipSuccess) return hipFail(e, "upload materials");
    if ((e = dTexs_.upload(dtex.data(), dtex.size() * sizeof(OmvTexture))) != hipSuccess) return hipFail(e, "upload textures");
    if ((e = dTexels_.upload(texels.data(), texels.size())) != hipSuccess) return hipFail(e, "upload texels");
    if (!deviceBuild) {
        if ((e = dNodes_.upload(nodes.data(), nodes.size() * sizeof(OmvNode4))) != hipSuccess) return hipFail(e, "upload nodes");
        if ((e = dTris_.upload(tris.data(), tris.size() * 4)) != hipSuccess) return hipFail(e, "upload triangles");
    }
    if ((e = dLights_.upload(lights_.data(), lights_.size() * sizeof(OmvLight))) != hipSuccess) return hipFail(e, "upload lights");
    if ((e = dCdf_.upload(cdf.data(), cdf.size() * 4)) != hipSuccess) return hipFail(e, "upload light cdf");
    if ((e = dBg_.upload(bgRgb_.data(), bgRgb_.size() * 4)) != hipSuccess) return hipFail(e, "upload background");

    OmvSceneView& v = view_;
    std::memset(&v, 0, sizeof(v));
    v.vtx = dVtx_.as<float>(); v.faces = dFaces_.as<uint32_t>(); v.meshes = dMeshes_.as<OmvMesh>();
    v.insts = dInsts_.as<OmvInstance>(); v.mats = dMats_.as<OmvMaterial>(); v.texs = dTexs_.as<OmvTexture>();
    v.texels = dTexels_.as<uint8_t>(); v.nodes = dNodes_.as<OmvNode4>(); v.tris = dTris_.as<float>();
    v.lights = dLights_.as<OmvLight>(); v.light_cdf = dCdf_.as<float>(); v.bg = dBg_.as<float>();
    v.num_meshes = (int)meshes_.size(); v.num_insts = (int)insts_.size(); v.num_mats = (int)mats_.size();
    v.num_texs = (int)texs_.size(); v.num_lights = (int)lights_.size();
    v.tlas_root = tl.root;
    v.lights_leave = lightsLeave;
    v.total_faces = faceBase;
    v.fc_subject = fcSubject_;
    v.any_emissive = anyEmissive;
    for (const OmvMaterial& m : mats_) v.any_masked |= (m.type == OMV_MAT_PBR && m.alpha_mode == OMV_ALPHA_MASK) ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        const float ext = sceneHi[a] - sceneLo[a];
        v.scene_lo[a] = std::isfinite(sceneLo[a]) ? sceneLo[a] : 0.0f;
        v.scene_inv[a] = (std::isfinite(ext) && ext > 0.0f) ? 1.0f / ext : 0.0f;
    }
    v.bg_type = bgType_; v.bg_w = bgW_; v.bg_h = bgH_;
    for (int c = 0; c < 3; ++c) v.bg_const[c] = bgConst_[c];
    std::memcpy(v.bg_rot, bgRot_, sizeof(bgRot_));
    v.bg_scale = bgScale_;
    v.light_max = cdf.empty() ? 0.0f : cdf.back();
    v.min_density = minD; v.max_density = maxD;

    // environment importance sampling tables, built on the device
    if (bgType_ == OMV_BG_HDR && bgW_ > 0 && bgH_ > 0) {
        const size_t n = (size_t)bgW_ * bgH_, H = (size_t)bgH_;
        if ((e = dEnv_.reserve((2 * H + 3 * n) * 4 + 256)) != hipSuccess) return hipFail(e, "environment tables");
        if ((e = dEnvScratch_.reserve((4 * n + 2 * H + 2) * 4)) != hipSuccess) return hipFail(e, "environment scratch");
        float* rp = dEnv_.as<float>();
        int* ra = reinterpret_cast<int*>(rp + H);
        float* cp = reinterpret_cast<float*>(ra + H);
        int* ca = reinterpret_cast<int*>(cp + n);
        float* tp = reinterpret_cast<float*>(ca + n);
        if ((e = launchEnvBuild(dBg_.as<float>(), bgW_, bgH_, rp, ra, cp, ca, tp, dEnvScratch_.ptr, nullptr)) != hipSuccess ||
            (e = hipDeviceSynchronize()) != hipSuccess)
            return hipFail(e, "environment table build");
        dEnvScratch_.release();
        v.env_row_prob = rp; v.env_row_alias = ra; v.env_col_prob = cp; v.env_col_alias = ca; v.env_tex_pdf = tp;
    }

    nodeCount_ = (int)nodes.size() + deviceNodes;
    bvhDepth_ = tl.depth + blasDepth;
    recordCount_ = deviceBuild ? recTotal : (int)(tris.size() / 12);
    deviceBytes_ = dVtx_.bytes + dFaces_.bytes + dMeshes_.bytes + dInsts_.bytes + dMats_.bytes + dTexs_.bytes + dTexels_.bytes +
                   dNodes_.bytes + dTris_.bytes + dLights_.bytes + dCdf_.bytes + dBg_.bytes + dEnv_.bytes;
    committed_ = true;
    return OMV_OK;
}

int RenderScene::updateTexture(int texId, const void* src) {
    if (!committed_) { setError("update_texture: scene not committed"); return OMV_E_STATE; }
    if (texId < 0 || texId >= (int)texs_.size()) { setError("update_texture: texture id out of range"); return OMV_E_RANGE; }
    HostTexture& t = texs_[texId];
    if (t.bytes.empty()) { setError("update_texture: texture has no image"); return OMV_E_STATE; }
    if (!src) { setError("update_texture: null texels"); return OMV_E_ARG; }
    hipError_t e = hipMemcpy(dTexels_.as<uint8_t>() + texOffsets_[texId], src, t.bytes.size(), hipMemcpyDefault);
    if (e == hipSuccess) e = hipMemcpy(t.bytes.data(), src, t.bytes.size(), hipMemcpyDefault);
    if (e != hipSuccess) return hipFail(e, "update_texture");
    return OMV_OK;
}

int RenderScene::updateMaterialOmNk(int matId, const float n[3], const float k[3]) {
    if (!committed_) { setError("update_material: scene not committed"); return OMV_E_STATE; }
    if (!isOm(matId)) { setError("update_material: not an OpenMATERIAL material"); return OMV_E_RANGE; }
    OmvMaterial& m = mats_[matId];
    for (int c = 0; c < 3; ++c) { m.n[c] = n[c]; m.k[c] = k[c]; }
    m.has_base = 1;
    hipError_t e = hipMemcpy(dMats_.as<uint8_t>() + (size_t)matId * sizeof(OmvMaterial), &m, sizeof(OmvMaterial), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hipFail(e, "update_material");
    return OMV_OK;
}

int RenderScene::envTables(float* rowProb, int* rowAlias, float* colProb, int* colAlias, float* texPdf) const {
    if (!committed_ || !view_.env_tex_pdf) { setError("env_tables: no environment tables"); return OMV_E_STATE; }
    const size_t n = (size_t)bgW_ * bgH_, H = (size_t)bgH_;
    hipError_t e = hipMemcpy(rowProb, view_.env_row_prob, H * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(rowAlias, view_.env_row_alias, H * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(colProb, view_.env_col_prob, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(colAlias, view_.env_col_alias, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(texPdf, view_.env_tex_pdf, n * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipFail(e, "env_tables");
    return OMV_OK;
}

int RenderScene::textureSize(int texId, int* w, int* h) const {
    if (texId < 0 || texId >= (int)texs_.size()) return OMV_E_RANGE;
    *w = texs_[texId].bytes.empty() ? 0 : texs_[texId].w;
    *h = texs_[texId].bytes.empty() ? 0 : texs_[texId].h;
    return OMV_OK;
}

int RenderScene::meshInfo(int id, int* vn, int* fn, int* stride, int* mat) const {
    if (id < 0 || id >= (int)meshes_.size()) { setError("meshInfo: id out of range"); return OMV_E_RANGE; }
    const HostMesh& m = meshes_[id];
    if (vn) *vn = m.vn;
    if (fn) *fn = m.fn;
    if (stride) *stride = m.stride;
    if (mat) *mat = m.matId;
    return OMV_OK;
}

int RenderScene::vertices(int id, float* out, size_t cap) const {
    if (id < 0 || id >= (int)meshes_.size()) { setError("vertices: id out of range"); return OMV_E_RANGE; }
    const HostMesh& m = meshes_[id];
    if (!out || cap < m.V.size()) { setError("vertices: buffer too small"); return OMV_E_ARG; }
    if (committed_ && id < (int)devPrepared_.size() && devPrepared_[id]) {   // normals / tangents were made on the device
        hipError_t e = hipMemcpy(out, dVtx_.as<float>() + devVtxOffset_[id], m.V.size() * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { setError(std::string("vertices: ") + hipGetErrorString(e)); return OMV_E_DEVICE; }
        return OMV_OK;
    }
    std::memcpy(out, m.V.data(), m.V.size() * sizeof(float));
    return OMV_OK;
}

int RenderScene::bvhInfo(int* nodes, int* depth, int* records, size_t* bytes) const {
    if (!committed_) { setError("bvhInfo: scene not committed"); return OMV_E_STATE; }
    if (nodes) *nodes = nodeCount_;
    if (depth) *depth = bvhDepth_;
    if (records) *records = recordCount_;
    if (bytes) *bytes = deviceBytes_;
    return OMV_OK;
}

}  // namespace omv
