/*
 * asset_harness.cpp -- driver that loads glTF / OpenMATERIAL assets with the
 * reference pathtracer's own AssetGeometry (compiled from the sources where
 * they lie in the reference checkout, see oracle/build_ref.sh) and writes
 * what it loaded as JSON golden vectors.  Test infrastructure only; nothing
 * here ships.
 *
 * Covered (reference paths under pathtracer/src/):
 *   AssetGeometry(filename)              assets/AssetGeometry.cpp:209-215, 297-822
 *   Node transforms / global transforms  assets/Node.h (getGlobalTransformation)
 *   iterateAssetGeometryNodes            assets/AssetGeometry.h:179-200
 *   AssetGeometry::getBBox               assets/AssetGeometry.cpp:229-236
 *   PbrMaterial getters                  assets/materials/PbrMaterial.cpp
 *   AssetMaterial + AssetMaterialIor     assets/materials/AssetMaterial.cpp, AssetMaterialIor.cpp
 *   PointLight getters                   lights/PointLight.h:49-56
 *   ImageData                            assets/ImageData.cpp:40-111
 *
 * Usage: asset_harness OUT.json LIGHTS(0|1) PRIORITY(om|pbr) FILE...
 *        asset_harness OUT.json hdr FILE...   (stbi_loadf of each file, the
 *        decoder HDRBackground uses, HdrBackground.cpp:38-47)
 *        asset_harness OUT.json write DIR NAME W H...   (ImageSaver with
 *        flipVerticallyOnWrite(true), as main.cpp:149 sets it: DIR/NAME.rgba8
 *        (W*H*4 bytes) written as NAME.png/.bmp/.tga/.jpg (quality 100) and
 *        DIR/NAME.f32 (W*H*4 floats) as NAME.pfm, ImageSaver.h:23-97)
 * Buffers are summarised by (length, checksum): the checksum is the sum over
 * the little-endian 32-bit words w_i of the raw bytes of w_i * (2 i + 1),
 * modulo 2^64 (zero-padded to a whole word).
 */
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "AssetGeometry.h"
#include "Exceptions.h"
#include "materials/AssetMaterial.h"
#include "materials/AssetMaterialIor.h"
#include "materials/PbrMaterial.h"
#include "stb_image.h"
#include "ImageSaver.h"

namespace {

FILE* OUT = nullptr;

uint64_t checksum(const void* p, size_t nbytes) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    uint64_t h = 0;
    for (size_t i = 0, w = 0; i < nbytes; i += 4, ++w) {
        uint32_t v = 0;
        for (size_t j = 0; j < 4 && i + j < nbytes; ++j) v |= (uint32_t)b[i + j] << (8 * j);
        h += (uint64_t)v * (uint64_t)(2 * w + 1);
    }
    return h;
}

void f(float v) { std::fprintf(OUT, "%.9g", (double)v); }
void farr(const float* v, int n) {
    std::fprintf(OUT, "[");
    for (int i = 0; i < n; ++i) { if (i) std::fprintf(OUT, ","); f(v[i]); }
    std::fprintf(OUT, "]");
}
void mat4(const fmat4 m) { farr(&m[0][0], 16); }
void str(const std::string& s) {
    std::fputc('"', OUT);
    for (char c : s) {
        if (c == '"' || c == '\\') std::fprintf(OUT, "\\%c", c);
        else if ((unsigned char)c < 0x20) std::fprintf(OUT, "\\u%04x", (unsigned char)c);
        else std::fputc(c, OUT);
    }
    std::fputc('"', OUT);
}
template <typename T>
void buf(const std::vector<T>& v) {
    std::fprintf(OUT, "[%zu,\"%016llx\"]", v.size(), (unsigned long long)checksum(v.data(), v.size() * sizeof(T)));
}

const float kWl[3] = {6.50000004e-07f, 5.10000007e-07f, 4.40000008e-07f};

void dumpAsset(AssetGeometry& ag) {
    std::map<const Node*, int> nodeIdx;
    for (size_t i = 0; i < ag.m_vpNodes.size(); ++i) nodeIdx[ag.m_vpNodes[i]] = (int)i;
    std::map<const Material*, int> matIdx;
    for (size_t i = 0; i < ag.m_vpMaterials.size(); ++i) matIdx[ag.m_vpMaterials[i]] = (int)i;

    std::fprintf(OUT, "\"title\":");
    str(ag.getTitle());
    std::fprintf(OUT, ",\"nodes\":[");
    for (size_t i = 0; i < ag.m_vpNodes.size(); ++i) {
        const Node* n = ag.m_vpNodes[i];
        if (i) std::fprintf(OUT, ",");
        std::fprintf(OUT, "{\"name\":");
        str(n->m_sName);
        std::fprintf(OUT, ",\"id\":");
        str(n->m_sId);
        std::fprintf(OUT, ",\"uri\":");
        str(n->m_sUri);
        std::fprintf(OUT, ",\"category\":");
        str(n->m_sCategory);
        std::fprintf(OUT, ",\"parent\":%d,\"children\":[", n->m_pParent ? nodeIdx.at(n->m_pParent) : -1);
        for (size_t c = 0; c < n->m_vpChildren.size(); ++c)
            std::fprintf(OUT, "%s%d", c ? "," : "", nodeIdx.count(n->m_vpChildren[c]) ? nodeIdx.at(n->m_vpChildren[c]) : -9);
        std::fprintf(OUT, "],\"prims\":[");
        for (size_t c = 0; c < n->m_vpPrimitives.size(); ++c) std::fprintf(OUT, "%s%d", c ? "," : "", n->m_vpPrimitives[c]->m_iId);
        std::fprintf(OUT, "],\"T\":");
        mat4(n->m_Transformation);
        fmat4 g;
        n->getGlobalTransformation(g);
        std::fprintf(OUT, ",\"G\":");
        mat4(g);
        std::fprintf(OUT, "}");
    }
    std::fprintf(OUT, "],\"scene_nodes\":[");
    for (size_t i = 0; i < ag.m_viSceneNodes.size(); ++i) std::fprintf(OUT, "%s%d", i ? "," : "", ag.m_viSceneNodes[i]);

    std::fprintf(OUT, "],\"prims\":[");
    for (size_t i = 0; i < ag.m_vpPrimitives.size(); ++i) {
        const Mesh* m = ag.m_vpPrimitives[i];
        if (i) std::fprintf(OUT, ",");
        int mat = -1;                                       // no material
        if (m->m_pMaterial && m->m_pMaterial == ag.m_pMissingMaterial) mat = -2;
        else if (m->m_pMaterial) mat = matIdx.count(m->m_pMaterial) ? matIdx.at(m->m_pMaterial) : -3;
        std::fprintf(OUT, "{\"id\":%d,\"name\":", m->m_iId);
        str(m->m_sName);
        std::fprintf(OUT, ",\"mat\":%d,\"idx\":", mat);
        buf(m->m_vuiIndexBuffer);
        std::fprintf(OUT, ",\"pos\":");
        buf(m->m_vfVertexBuffer);
        std::fprintf(OUT, ",\"nrm\":");
        buf(m->m_vfNormalBuffer);
        std::fprintf(OUT, ",\"tan\":");
        buf(m->m_vfTangentBuffer);
        std::fprintf(OUT, ",\"uv\":[");
        for (int c = 0; c < MAX_UV_CHANNELS; ++c) {
            if (c) std::fprintf(OUT, ",");
            buf(m->m_vfTexCoordBuffers[c]);
        }
        std::fprintf(OUT, "]}");
    }

    std::fprintf(OUT, "],\"missing_material\":%s,\"materials\":[", ag.m_pMissingMaterial ? "true" : "false");
    for (size_t i = 0; i < ag.m_vpMaterials.size(); ++i) {
        if (i) std::fprintf(OUT, ",");
        const Material* mb = ag.m_vpMaterials[i];
        if (auto om = dynamic_cast<const AssetMaterial*>(mb)) {
            std::fprintf(OUT, "{\"kind\":\"om\",\"filename\":");
            str(om->getFilename());
            std::fprintf(OUT, ",\"name\":");
            str(om->getName());
            std::fprintf(OUT, ",\"temperature\":");
            f(om->getTemperature());
            std::fprintf(OUT, ",\"ior_uri\":");
            str(om->getIndexOfRefractionUri());
            const AssetMaterialIor* ior = om->getIorPointer();
            std::fprintf(OUT, ",\"nk\":");
            if (ior) {
                float nk[6];
                for (int c = 0; c < 3; ++c) ior->getIor(om->getTemperature(), kWl[c], nk[c], nk[3 + c]);
                farr(nk, 6);
            } else {
                std::fprintf(OUT, "null");
            }
            std::fprintf(OUT, "}");
        } else if (auto pbr = dynamic_cast<const PbrMaterial*>(mb)) {
            fvec4 bc;
            fvec3 em;
            pbr->getBaseColorFactors(bc);
            pbr->getEmissiveFactors(em);
            std::fprintf(OUT, "{\"kind\":\"pbr\",\"name\":");
            str(pbr->getMaterialName());
            std::fprintf(OUT, ",\"base_color\":");
            farr(bc, 4);
            std::fprintf(OUT, ",\"emissive\":");
            farr(em, 3);
            std::fprintf(OUT, ",\"metallic\":");
            f(pbr->getMetallicFactor());
            std::fprintf(OUT, ",\"roughness\":");
            f(pbr->getRoughnessFactor());
            std::fprintf(OUT, ",\"normal_scale\":");
            f(pbr->getNormalScale());
            std::fprintf(OUT, ",\"occlusion_strength\":");
            f(pbr->getOcclusionStrength());
            std::fprintf(OUT, ",\"ior\":");
            f(pbr->getIor());
            std::fprintf(OUT, ",\"alpha_cutoff\":");
            f(pbr->getAlphaCutOff());
            std::fprintf(OUT, ",\"alpha_mode\":%d,\"double_sided\":%d,\"maps\":[", (int)pbr->getAlphaMode(), pbr->isDoubleSided() ? 1 : 0);
            const PbrMaterial::Texture slots[5] = {PbrMaterial::BaseColor, PbrMaterial::MetallicRoughness, PbrMaterial::Normal,
                                                   PbrMaterial::Emissivity, PbrMaterial::Occlusion};
            for (int s = 0; s < 5; ++s) {
                fmat3 tm;
                pbr->getTextureTransformation(slots[s], tm);
                std::fprintf(OUT, "%s[%d,%d,", s ? "," : "", pbr->getTextureMapId(slots[s]), pbr->getTextureChannelId(slots[s]));
                farr(&tm[0][0], 9);
                std::fprintf(OUT, "]");
            }
            std::fprintf(OUT, "]}");
        } else {
            std::fprintf(OUT, "{\"kind\":\"other\"}");
        }
    }

    std::fprintf(OUT, "],\"lights\":[");
    for (size_t i = 0; i < ag.m_vpLights.size(); ++i) {
        const PointLight* l = ag.m_vpLights[i];
        if (i) std::fprintf(OUT, ",");
        float rgb[3] = {l->getLightRed(), l->getLightGreen(), l->getLightBlue()};
        std::fprintf(OUT, "{\"rgb_getters\":");
        farr(rgb, 3);
        std::fprintf(OUT, ",\"intensity\":");
        f(l->getLightIntensity());
        std::fprintf(OUT, ",\"range\":");
        f(l->getLightRange());
        std::fprintf(OUT, ",\"node\":%d}", l->getNode() && nodeIdx.count(l->getNode()) ? nodeIdx.at(l->getNode()) : -1);
    }

    std::fprintf(OUT, "],\"images\":[");
    for (size_t i = 0; i < ag.m_vImages.size(); ++i) {
        const ImageData& im = ag.m_vImages[i];
        if (i) std::fprintf(OUT, ",");
        size_t nb = (im.m_pImage && im.m_iWidth > 0 && im.m_iHeight > 0) ? (size_t)im.m_iWidth * im.m_iHeight * 4 : 0;
        std::fprintf(OUT, "{\"id\":%d,\"w\":%d,\"h\":%d,\"rgba\":[%zu,\"%016llx\"]}", im.m_iId, im.m_iWidth, im.m_iHeight, nb,
                     (unsigned long long)checksum(im.m_pImage, nb));
    }

    BoundingBox<float> bb = ag.getBBox();
    std::fprintf(OUT, "],\"bbox_valid\":%s,\"bbox\":", bb.isValid() ? "true" : "false");
    if (bb.isValid()) {
        float c[8][3];
        bb.getCorners(c);
        float lohi[6] = {c[0][0], c[0][1], c[0][2], c[7][0], c[7][1], c[7][2]};
        farr(lohi, 6);
        fvec3 ctr;
        bb.getCenter(ctr);
        std::fprintf(OUT, ",\"bbox_center\":");
        farr(ctr, 3);
    } else {
        std::fprintf(OUT, "null");
    }

    // instance lists as main.cpp assembles them (mesh id + world matrix per
    // primitive of every node, depth first from the scene roots), once with
    // the identity and once with a fixed non-trivial scene transformation
    fmat4 I, T;
    assignIdentityMatrix(I);
    const float Tv[16] = {0.36f, -0.48f, 0.8f, 1.5f, 0.8f, 0.6f, 0.0f, -2.25f, -0.48f, 0.64f, 0.6f, 0.75f, 0, 0, 0, 1};
    std::memcpy(T, Tv, sizeof(T));
    const fmat4* mats[2] = {&I, &T};
    const char* names[2] = {"instances_identity", "instances_T"};
    for (int k = 0; k < 2; ++k) {
        std::fprintf(OUT, ",\"%s\":[", names[k]);
        bool first = true;
        iterateAssetGeometryNodes(&ag, *mats[k], [&](Node* n, const fmat4 tM) {
            for (auto* m : n->m_vpPrimitives) {
                std::fprintf(OUT, "%s[%d,", first ? "" : ",", m->m_iId);
                farr(&tM[0][0], 16);
                std::fprintf(OUT, "]");
                first = false;
            }
        });
        std::fprintf(OUT, "]");
    }
    std::fprintf(OUT, ",\"light_world\":[");
    for (size_t i = 0; i < ag.m_vpLights.size(); ++i) {
        fmat4 g, tM;
        ag.m_vpLights[i]->getNode()->getGlobalTransformation(g);
        multSquareMatrix(T, g, tM);
        std::fprintf(OUT, "%s", i ? "," : "");
        mat4(tM);
    }
    std::fprintf(OUT, "]");
}

const char* kind(const std::exception& e) {
    if (dynamic_cast<const GltfError*>(&e)) return "GltfError";
    if (dynamic_cast<const NotImplementedError*>(&e)) return "NotImplementedError";
    if (dynamic_cast<const ParsingError*>(&e)) return "ParsingError";
    if (dynamic_cast<const ValueError*>(&e)) return "ValueError";
    if (dynamic_cast<const FileNotFoundError*>(&e)) return "FileNotFoundError";
    return "exception";
}

}  // namespace

int dumpHdr(int argc, char** argv) {
    std::fprintf(OUT, "[");
    for (int a = 3; a < argc; ++a) {
        int w = 0, h = 0, n = 0;
        float* px = stbi_loadf(argv[a], &w, &h, &n, 0);
        std::fprintf(OUT, "%s\n{\"file\":", a > 3 ? "," : "");
        str(argv[a]);
        size_t nb = px ? (size_t)w * h * n * sizeof(float) : 0;
        std::fprintf(OUT, ",\"ok\":%s,\"w\":%d,\"h\":%d,\"n\":%d,\"data\":[%zu,\"%016llx\"]}", px ? "true" : "false", w, h, n,
                     nb / sizeof(float), (unsigned long long)checksum(px, nb));
        if (px) stbi_image_free(px);
    }
    std::fprintf(OUT, "]\n");
    std::fclose(OUT);
    return 0;
}

std::vector<char> readFile(const std::string& path) {
    std::vector<char> d;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return d;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) d.insert(d.end(), buf, buf + n);
    std::fclose(f);
    return d;
}

int writeImages(int argc, char** argv) {
    ImageSaver::flipVerticallyOnWrite(true);
    const std::string dir = argv[3];
    std::fprintf(OUT, "[");
    for (int a = 4; a + 2 < argc; a += 3) {
        const std::string name = argv[a];
        const int w = std::atoi(argv[a + 1]), h = std::atoi(argv[a + 2]);
        std::vector<char> px = readFile(dir + "/" + name + ".rgba8"), fl = readFile(dir + "/" + name + ".f32");
        if ((int)px.size() != w * h * 4 || (int)fl.size() != w * h * 16) return 3;
        const std::string base = dir + "/" + name;
        ImageSaver::savePng((base + ".png").c_str(), w, h, 4, px.data());
        ImageSaver::saveBmp((base + ".bmp").c_str(), w, h, 4, px.data());
        ImageSaver::saveTga((base + ".tga").c_str(), w, h, 4, px.data());
        ImageSaver::saveJpg((base + ".jpg").c_str(), w, h, 4, px.data(), 100);
        ImageSaver::savePfm((base + ".pfm").c_str(), w, h, reinterpret_cast<const float*>(fl.data()));
        std::fprintf(OUT, "%s\n{\"name\":", a > 4 ? "," : "");
        str(name);
        std::fprintf(OUT, ",\"w\":%d,\"h\":%d", w, h);
        for (const char* ext : {"png", "bmp", "tga", "jpg", "pfm"}) {
            std::vector<char> b = readFile(base + "." + ext);
            std::fprintf(OUT, ",\"%s\":[%zu,\"%016llx\"]", ext, b.size(), (unsigned long long)checksum(b.data(), b.size()));
        }
        std::fprintf(OUT, "}");
    }
    std::fprintf(OUT, "]\n");
    std::fclose(OUT);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 4 && std::string(argv[2]) == "write") {
        OUT = std::fopen(argv[1], "w");
        return OUT ? writeImages(argc, argv) : 2;
    }
    if (argc >= 3 && std::string(argv[2]) == "hdr") {
        OUT = std::fopen(argv[1], "w");
        return OUT ? dumpHdr(argc, argv) : 2;
    }
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s OUT.json LIGHTS(0|1) PRIORITY(om|pbr) FILE...\n", argv[0]);
        return 2;
    }
    OUT = std::fopen(argv[1], "w");
    if (!OUT) return 2;
    AssetGeometry::m_bUseLights = argv[2][0] == '1';
    AssetGeometry::m_sMaterialLoadingPriority = argv[3];
    std::fprintf(OUT, "[");
    for (int a = 4; a < argc; ++a) {
        std::fprintf(OUT, "%s\n{\"file\":", a > 4 ? "," : "");
        str(argv[a]);
        std::fprintf(OUT, ",\"lights\":%d,\"priority\":", AssetGeometry::m_bUseLights ? 1 : 0);
        str(argv[3]);
        std::fprintf(OUT, ",");
        std::fflush(OUT);
        try {
            // never destroyed: Node's copy constructor shares m_pAssetInfo
            // (assets/Node.h copyFrom), so an asset whose copied nodes hold
            // asset info would be freed twice
            AssetGeometry* ag = new AssetGeometry(std::string(argv[a]));
            std::fprintf(OUT, "\"error\":null,");
            dumpAsset(*ag);
        } catch (const std::exception& e) {
            std::fprintf(OUT, "\"error\":");
            str(std::string(kind(e)) + ": " + e.what());
        }
        std::fprintf(OUT, "}");
        std::fflush(OUT);
    }
    std::fprintf(OUT, "]\n");
    std::fclose(OUT);
    return 0;
}
