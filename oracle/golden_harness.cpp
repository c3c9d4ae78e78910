/*
 * golden_harness.cpp -- driver that calls the reference pathtracer's own
 * Embree-free code (compiled from the sources where they lie in the
 * reference checkout, see oracle/build_ref.sh) and prints its outputs as
 * JSON golden vectors.  Test infrastructure only; nothing here ships.
 *
 * Functions exercised (reference paths under pathtracer/src/):
 *   BSDFModel<GGX>::evaluateDirect / sampleIndirect   brdf/BRDFModel.h:211-337
 *   RandomSampler warps                               utilities/RandomSampler.h:66-103
 *   DistanceRandom                                    utilities/DistanceRandom.h:85-127
 *   Interpolation                                     utilities/Interpolation.h:44-208
 *   AssetMaterialIor::getIor                          assets/materials/AssetMaterialIor.cpp:115-178
 *   BitmapTexture::texture                            materials/BitmapTexture.cpp:181-234
 *   HDRBackground::getRadiance                        backgrounds/HdrBackground.cpp:38-97
 *   PinholeCamera                                     sensors/PinholeCamera.cpp:51-174
 *   RenderLightPoint                                  lights/RenderLightPoint.cpp:20-45
 *   gammaCorrection                                   tonemapping/ToneMapping.cpp:24-54
 *
 * Random draws: RandomSampler is std::mt19937 + uniform_real_distribution<float>;
 * a twin generator with the same seed yields the same sequence, which is
 * recorded so the oracle can replay it.
 */
#include <cstdio>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "BRDFModel.h"
#include "BitmapTexture.h"
#include "DistanceRandom.h"
#include "HdrBackground.h"
#include "Interpolation.h"
#include "PinholeCamera.h"
#include "RandomSampler.h"
#include "RenderLightPoint.h"
#include "ToneMapping.h"
#include "materials/AssetMaterialIor.h"

static FILE* OUT = stdout;
static bool firstItem = true;

static void f(float v) { std::fprintf(OUT, "%.9g", (double)v); }
static void arr(const float* v, int n) {
    std::fprintf(OUT, "[");
    for (int i = 0; i < n; ++i) { if (i) std::fprintf(OUT, ","); f(v[i]); }
    std::fprintf(OUT, "]");
}
static void arrv(const std::vector<float>& v) { arr(v.data(), (int)v.size()); }
static void beginSection(const char* name) {
    std::fprintf(OUT, "%s\"%s\": [", firstItem ? "" : ",\n", name);
    firstItem = false;
}
static void endSection() { std::fprintf(OUT, "]"); }

struct Twin {
    std::mt19937 g;
    std::uniform_real_distribution<float> d;
    explicit Twin(unsigned s) : g(s), d(0.0f, 1.0f) {}
    std::vector<float> take(int n) { std::vector<float> v(n); for (auto& x : v) x = d(g); return v; }
};

static void randUnit(std::mt19937& g, float* v) {
    std::normal_distribution<float> nd(0.0f, 1.0f);
    float l;
    do { v[0] = nd(g); v[1] = nd(g); v[2] = nd(g); l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); } while (l < 1e-3f);
    v[0] /= l; v[1] /= l; v[2] /= l;
}

/* Count how many draws a reference call consumed: the twin list holds the
 * sequence from the same seed; the next draw of the reference sampler is
 * located in it. */
static int consumed(RandomSampler& rs, const std::vector<float>& list) {
    float nxt = rs.rand();
    for (size_t i = 0; i < list.size(); ++i) if (list[i] == nxt) return (int)i;
    return -1;
}

static void writeHdrRgbe(const char* path, int w, int h, const std::vector<unsigned char>& rgbe) {
    std::ofstream o(path, std::ios::binary);
    o << "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " << h << " +X " << w << "\n";
    o.write(reinterpret_cast<const char*>(rgbe.data()), (std::streamsize)rgbe.size());
}

int main(int argc, char** argv) {
    const char* refRoot = argc > 1 ? argv[1] : "/root/reference";
    const char* tmpDir = argc > 2 ? argv[2] : "/tmp";
    std::mt19937 g(20240611u);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    std::fprintf(OUT, "{\n");

    /* ---------------- evaluateDirect */
    beginSection("evaluate_direct");
    for (int c = 0; c < 200; ++c) {
        BSDFData d;
        float N[3], I[3], O[3];
        randUnit(g, N); randUnit(g, I); randUnit(g, O);
        if (c % 3 != 2) {  // mostly valid geometry: I toward the surface, O away from it
            if (I[0] * N[0] + I[1] * N[1] + I[2] * N[2] > 0) for (float& x : I) x = -x;
            if (O[0] * N[0] + O[1] * N[1] + O[2] * N[2] < 0) for (float& x : O) x = -x;
        }
        std::copy_n(I, 3, d.I); std::copy_n(O, 3, d.O); std::copy_n(N, 3, d.N);
        for (int i = 0; i < 4; ++i) d.color[i] = U(g);
        d.metallness = U(g);
        float r = 0.02f + 0.97f * U(g);
        d.alpha = r * r;
        fvec3 brdf;
        BSDFModel<ndf::GGX>::evaluateDirect(d, brdf);
        std::fprintf(OUT, "%s{\"I\":", c ? ",\n" : "\n"); arr(I, 3);
        std::fprintf(OUT, ",\"O\":"); arr(O, 3);
        std::fprintf(OUT, ",\"N\":"); arr(N, 3);
        std::fprintf(OUT, ",\"color\":"); arr(d.color, 4);
        std::fprintf(OUT, ",\"metal\":"); f(d.metallness);
        std::fprintf(OUT, ",\"alpha\":"); f(d.alpha);
        std::fprintf(OUT, ",\"brdf\":"); arr(brdf, 3);
        std::fprintf(OUT, "}");
    }
    endSection();

    /* ---------------- sampleIndirect */
    beginSection("sample_indirect");
    for (int c = 0; c < 300; ++c) {
        BSDFData d;
        float N[3], I[3], T[3], tmp[3];
        randUnit(g, N); randUnit(g, I); randUnit(g, tmp);
        // tangent orthogonal to N
        float dn = tmp[0] * N[0] + tmp[1] * N[1] + tmp[2] * N[2];
        for (int i = 0; i < 3; ++i) T[i] = tmp[i] - dn * N[i];
        float tl = std::sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
        for (int i = 0; i < 3; ++i) T[i] /= tl;
        if (c % 4 != 3 && I[0] * N[0] + I[1] * N[1] + I[2] * N[2] > 0) for (float& x : I) x = -x;
        std::copy_n(I, 3, d.I); std::copy_n(N, 3, d.N); std::copy_n(T, 3, d.T);
        for (int i = 0; i < 3; ++i) d.color[i] = U(g);
        d.color[3] = (c % 5 == 0) ? 0.3f + 0.6f * U(g) : 1.0f;
        d.isTransmissive = (c % 2 == 0) || d.color[3] < 1.0f;
        d.metallness = U(g);
        float r = 0.02f + 0.97f * U(g);
        d.roughness = r;
        d.alpha = r * r;
        d.eta = (c % 3 == 0) ? 1.0f : 1.0f / 1.5f;
        unsigned seed = 1000u + (unsigned)c;
        RandomSampler rs(seed);
        Twin tw(seed);
        std::vector<float> list = tw.take(16);
        float I0[3] = {d.I[0], d.I[1], d.I[2]}, N0[3] = {d.N[0], d.N[1], d.N[2]};
        float color0[4] = {d.color[0], d.color[1], d.color[2], d.color[3]};
        int trans0 = d.isTransmissive ? 1 : 0;
        float eta0 = d.eta, metal0 = d.metallness, alpha0 = d.alpha;
        fvec3 brdf;
        BSDFModel<ndf::GGX>::sampleIndirect(d, &rs, brdf);
        int used = consumed(rs, list);
        std::fprintf(OUT, "%s{\"I\":", c ? ",\n" : "\n"); arr(I0, 3);
        std::fprintf(OUT, ",\"N\":"); arr(N0, 3);
        std::fprintf(OUT, ",\"T\":"); arr(T, 3);
        std::fprintf(OUT, ",\"color\":"); arr(color0, 4);
        std::fprintf(OUT, ",\"metal\":"); f(metal0);
        std::fprintf(OUT, ",\"alpha\":"); f(alpha0);
        std::fprintf(OUT, ",\"eta\":"); f(eta0);
        std::fprintf(OUT, ",\"transmissive\":%d", trans0);
        std::fprintf(OUT, ",\"uniforms\":"); arr(list.data(), used >= 0 ? std::max(used, 1) : (int)list.size());
        std::fprintf(OUT, ",\"used\":%d", used);
        std::fprintf(OUT, ",\"brdf\":"); arr(brdf, 3);
        std::fprintf(OUT, ",\"O\":"); arr(d.O, 3);
        std::fprintf(OUT, "}");
    }
    endSection();

    /* ---------------- warps */
    beginSection("warps");
    for (int c = 0; c < 240; ++c) {
        unsigned seed = 5000u + (unsigned)c;
        RandomSampler rs(seed);
        Twin tw(seed);
        std::vector<float> list = tw.take(64);
        fvec3 v;
        const char* kind;
        float a2 = 0.0f;
        switch (c % 4) {
            case 0: kind = "uniformHemisphere"; rs.uniformHemisphere(v); break;
            case 1: kind = "uniformSphere"; rs.uniformSphere(v); break;
            case 2: kind = "GGXHemisphere"; a2 = std::pow(0.01f + 0.98f * U(g), 4.0f); rs.GGXHemisphere(v, a2); break;
            default: kind = "cosineSqrtHemisphere"; rs.cosineSqrtHemisphere(v); break;
        }
        int used = consumed(rs, list);
        std::fprintf(OUT, "%s{\"kind\":\"%s\",\"alpha2\":", c ? ",\n" : "\n", kind); f(a2);
        std::fprintf(OUT, ",\"uniforms\":"); arr(list.data(), used >= 0 ? std::max(used, 1) : (int)list.size());
        std::fprintf(OUT, ",\"used\":%d,\"v\":", used); arr(v, 3);
        std::fprintf(OUT, "}");
    }
    endSection();

    /* ---------------- DistanceRandom */
    beginSection("distance_random");
    for (int c = 0; c < 120; ++c) {
        int n = 1 + (c % 7);
        std::vector<float> w(n);
        for (auto& x : w) x = (c % 5 == 0) ? 1.0f : U(g) * 10.0f;
        DistanceRandom<float> dr;
        dr.setCount(n);
        for (int i = 0; i < n; ++i) dr.setDistance(i, w[i]);
        dr.calculate();
        float u = (c % 11 == 0) ? 0.0f : U(g);
        int id = dr.getRandom(u);
        float pdf = dr.getPdf(id);
        std::fprintf(OUT, "%s{\"w\":", c ? ",\n" : "\n"); arrv(w);
        std::fprintf(OUT, ",\"u\":"); f(u);
        std::fprintf(OUT, ",\"idx\":%d,\"pdf\":", id); f(pdf);
        std::fprintf(OUT, "}");
    }
    endSection();

    /* ---------------- IOR */
    beginSection("ior");
    {
        const char* names[3] = {"gold", "aluminium", "iron"};
        int c = 0;
        for (int m = 0; m < 3; ++m) {
            std::string path = std::string(refRoot) + "/materials/data/" + names[m] + "_ior.gltf";
            AssetMaterialIor ior(path);
            std::vector<float> wls = {6.50000004e-07f, 5.10000007e-07f, 4.40000008e-07f};
            for (int i = 0; i < 40; ++i) wls.push_back(3.0e-7f + 7.0e-7f * U(g));
            for (float T : {300.0f, 0.0f, 1000.0f}) {
                for (float wl : wls) {
                    float n = 0, k = 0;
                    bool ok = true;
                    try { ior.getIor(T, wl, n, k); } catch (...) { ok = false; }
                    std::fprintf(OUT, "%s{\"material\":\"%s\",\"T\":", c++ ? ",\n" : "\n", names[m]); f(T);
                    std::fprintf(OUT, ",\"wl\":"); f(wl);
                    std::fprintf(OUT, ",\"ok\":%d,\"n\":", ok ? 1 : 0); f(n);
                    std::fprintf(OUT, ",\"k\":"); f(k);
                    std::fprintf(OUT, "}");
                }
            }
        }
    }
    endSection();

    /* ---------------- BitmapTexture */
    beginSection("texture");
    {
        const int sizes[4][2] = {{5, 3}, {4, 4}, {1, 1}, {7, 2}};
        int c = 0;
        for (auto& sz : sizes) {
            int w = sz[0], h = sz[1];
            std::vector<unsigned char> img((size_t)w * h * 4);
            for (auto& b : img) b = (unsigned char)(g() & 0xFF);
            BitmapTexture tex(w, h, reinterpret_cast<ubvec4*>(img.data()), true);
            std::fprintf(OUT, "%s{\"w\":%d,\"h\":%d,\"img\":[", c++ ? ",\n" : "\n", w, h);
            for (size_t k = 0; k < img.size(); ++k) std::fprintf(OUT, "%s%d", k ? "," : "", (int)img[k]);
            std::fprintf(OUT, "],\"queries\":[");
            for (int i = 0; i < 40; ++i) {
                float u = -2.0f + 5.0f * U(g), v = -2.0f + 5.0f * U(g);
                if (i < 4) { u = (float)i / (float)w; v = 0.5f / (float)h; }
                fvec4 o;
                tex.texture(u, v, o);
                std::fprintf(OUT, "%s{\"u\":", i ? "," : ""); f(u);
                std::fprintf(OUT, ",\"v\":"); f(v);
                std::fprintf(OUT, ",\"out\":"); arr(o, 4);
                std::fprintf(OUT, "}");
            }
            std::fprintf(OUT, "]}");
        }
    }
    endSection();

    /* ---------------- HDRBackground */
    beginSection("hdr");
    {
        const int sizes[2][2] = {{7, 5}, {12, 6}};
        int c = 0;
        for (auto& sz : sizes) {
            int w = sz[0], h = sz[1];
            std::vector<unsigned char> rgbe((size_t)w * h * 4);
            for (size_t p = 0; p < (size_t)w * h; ++p) {
                rgbe[p * 4 + 0] = (unsigned char)(3 + g() % 250);
                rgbe[p * 4 + 1] = (unsigned char)(g() % 256);
                rgbe[p * 4 + 2] = (unsigned char)(g() % 256);
                rgbe[p * 4 + 3] = (unsigned char)(120 + g() % 16);
            }
            std::string path = std::string(tmpDir) + "/omv_golden_" + std::to_string(w) + "x" + std::to_string(h) + ".hdr";
            writeHdrRgbe(path.c_str(), w, h, rgbe);
            HDRBackground bg(path.c_str());
            const fmat3 R = {{1.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 1.0f}, {0.0f, -1.0f, 0.0f}};  // main.cpp:401-403
            bg.setRotation(R);
            bg.setRadianceScale(3.14159265358979323846f);
            std::fprintf(OUT, "%s{\"w\":%d,\"h\":%d,\"rgbe\":[", c++ ? ",\n" : "\n", w, h);
            for (size_t k = 0; k < rgbe.size(); ++k) std::fprintf(OUT, "%s%d", k ? "," : "", (int)rgbe[k]);
            std::fprintf(OUT, "],\"queries\":[");
            for (int i = 0; i < 50; ++i) {
                float d[3];
                randUnit(g, d);
                if (i == 0) { d[0] = 0; d[1] = 1; d[2] = 0; }
                if (i == 1) { d[0] = 0; d[1] = -1; d[2] = 0; }
                if (i == 2) { d[0] = -1; d[1] = 0; d[2] = 0; }
                fvec3 o;
                bg.getRadiance(d, o);
                std::fprintf(OUT, "%s{\"dir\":", i ? "," : ""); arr(d, 3);
                std::fprintf(OUT, ",\"out\":"); arr(o, 3);
                std::fprintf(OUT, "}");
            }
            std::fprintf(OUT, "]}");
        }
    }
    endSection();

    /* ---------------- PinholeCamera */
    beginSection("camera");
    {
        int c = 0;
        const int res[3][2] = {{64, 48}, {37, 91}, {256, 256}};
        for (auto& rr : res) {
            for (int vp = 0; vp < 3; ++vp) {
                float pos[3] = {U(g) * 4 - 2, U(g) * 4 - 2, U(g) * 4 - 2}, fw[3], up[3] = {0, 1, 0}, left[3];
                randUnit(g, fw);
                if (vp == 0) { pos[0] = 0; pos[1] = 0; pos[2] = 4; fw[0] = 0; fw[1] = 0; fw[2] = -1; }
                cross3(fw, up, left);
                cross3(left, fw, up);
                float fov = vp == 2 ? 60.0f : 40.0f + 30.0f * U(g);
                PinholeCamera cam(ViewPoint(pos, left, up, fw), rr[0], rr[1]);
                cam.adjustToResolution();
                cam.setYFoV(fov);
                auto hs = cam.getRealSizes();
                for (int i = 0; i < 6; ++i) {
                    int x = (int)(g() % (unsigned)rr[0]), y = (int)(g() % (unsigned)rr[1]);
                    unsigned seed = 9000u + (unsigned)c;
                    RandomSampler rs(seed);
                    Twin tw(seed);
                    std::vector<float> list = tw.take(2);
                    RenderRay ray = cam.getRay(x, y, rs);
                    std::fprintf(OUT, "%s{\"w\":%d,\"h\":%d,\"pos\":", c++ ? ",\n" : "\n", rr[0], rr[1]); arr(pos, 3);
                    std::fprintf(OUT, ",\"left\":"); arr(left, 3);
                    std::fprintf(OUT, ",\"up\":"); arr(up, 3);
                    std::fprintf(OUT, ",\"fwd\":"); arr(fw, 3);
                    std::fprintf(OUT, ",\"fov\":"); f(fov);
                    std::fprintf(OUT, ",\"half_w\":"); f(hs.first);
                    std::fprintf(OUT, ",\"half_h\":"); f(hs.second);
                    std::fprintf(OUT, ",\"focus\":"); f(cam.getFocus());
                    std::fprintf(OUT, ",\"x\":%d,\"y\":%d,\"uniforms\":", x, y); arrv(list);
                    std::fprintf(OUT, ",\"O\":"); arr(ray.O(), 3);
                    std::fprintf(OUT, ",\"R\":"); arr(ray.R(), 3);
                    std::fprintf(OUT, ",\"id\":%d}", ray.getIndex());
                }
            }
        }
        /* the reference's own camera test (PinholeCamera.cpp:27-46) */
        fvec3 p = {0.f, 0.f, -5.f}, fwd = {0.f, 0.f, 1.f}, upv = {0.f, 1.f, 0.f}, lft;
        cross3(fwd, upv, lft);
        PinholeCamera pc(ViewPoint(p, lft, upv, fwd), 100, 100);
        std::string jp = std::string(refRoot) + "/cameras/cameraProps.json";
        pc.loadProperties(jp.c_str());
        auto hs = pc.getRealSizes();
        std::fprintf(OUT, ",\n{\"load_properties\":1,\"focus\":"); f(pc.getFocus());
        std::fprintf(OUT, ",\"half_w\":"); f(hs.first);
        std::fprintf(OUT, ",\"half_h\":"); f(hs.second);
        /* film normalisation: hit() then stop() */
        PinholeCamera fc(ViewPoint(p, lft, upv, fwd), 3, 2);
        fc.init();
        std::vector<float> rads;
        for (int s = 0; s < 7; ++s) {
            int id = s % 5;
            fvec3 rad = {U(g) * 3, U(g), U(g) * 0.1f};
            RenderRay r0, r1;
            r1.setIndex(id);
            fc.hit(rad, r0, &r1);
            rads.push_back((float)id); rads.push_back(rad[0]); rads.push_back(rad[1]); rads.push_back(rad[2]);
        }
        fc.stop();
        std::fprintf(OUT, ",\"hits\":"); arrv(rads);
        std::fprintf(OUT, ",\"film\":[");
        for (int y = 0; y < 2; ++y) for (int x = 0; x < 3; ++x) {
            fvec4 v; fc.getImpression(x, y, v);
            if (x || y) std::fprintf(OUT, ",");
            arr(v, 4);
        }
        std::fprintf(OUT, "]}");
    }
    endSection();

    /* ---------------- RenderLightPoint */
    beginSection("light_point");
    for (int c = 0; c < 40; ++c) {
        fvec3 pos = {U(g) * 6 - 3, U(g) * 6 - 3, U(g) * 6 - 3}, inten = {U(g) * 50, U(g) * 50, U(g) * 50};
        RenderLightPoint L;
        L.set(pos, inten, c % 3 == 0 ? 5.0f : std::numeric_limits<float>::max());
        unsigned seed = 7000u + (unsigned)c;
        RandomSampler rs(seed);
        Twin tw(seed);
        std::vector<float> list = tw.take(64);
        L.setRandomSampler(&rs);
        fvec3 O, R, rad;
        float pdf;
        L.getRandomRay(O, R, pdf, rad);
        int used = consumed(rs, list);
        std::fprintf(OUT, "%s{\"pos\":", c ? ",\n" : "\n"); arr(pos, 3);
        std::fprintf(OUT, ",\"intensity\":"); arr(inten, 3);
        std::fprintf(OUT, ",\"power\":"); f(L.getPower());
        std::fprintf(OUT, ",\"range\":"); f(L.getAttenuationDistance());
        std::fprintf(OUT, ",\"uniforms\":"); arr(list.data(), used >= 0 ? std::max(used, 1) : (int)list.size());
        std::fprintf(OUT, ",\"used\":%d,\"O\":", used); arr(O, 3);
        std::fprintf(OUT, ",\"R\":"); arr(R, 3);
        std::fprintf(OUT, ",\"rad\":"); arr(rad, 3);
        std::fprintf(OUT, ",\"pdf\":"); f(pdf);
        std::fprintf(OUT, "}");
    }
    endSection();

    /* ---------------- gammaCorrection */
    beginSection("gamma");
    for (int c = 0; c < 4; ++c) {
        int w = 5, h = 3;
        std::vector<float> img((size_t)w * h * 4), in;
        for (size_t i = 0; i < img.size(); ++i) img[i] = (i % 4 == 3) ? 1.0f : U(g) * (c == 0 ? 1.0f : 8.0f);
        if (c == 1) { img[0] = img[1] = img[2] = 0.0f; }
        in = img;
        float gam = c == 3 ? 0.45f : 0.5f;
        gammaCorrection(reinterpret_cast<fvec4*>(img.data()), w, h, 1.0f, gam);
        std::fprintf(OUT, "%s{\"w\":%d,\"h\":%d,\"gamma\":", c ? ",\n" : "\n", w, h); f(gam);
        std::fprintf(OUT, ",\"in\":"); arrv(in);
        std::fprintf(OUT, ",\"out\":"); arrv(img);
        std::fprintf(OUT, "}");
    }
    endSection();

    std::fprintf(OUT, "\n}\n");
    return 0;
}
