// adjoint.hip -- path-replay backward pass (Vicini, Speierer & Jakob,
// SIGGRAPH 2021) of the camera-path estimator of BDPTRenderer::computePaths
// (BDPTRenderer.cpp:237-301) for scenes without point lights, with detached
// sampling: directions and lobes stay those the current parameters sample,
// and each PBR bounce weight is w_ref * f(theta) / f(theta_0) for the sampled
// lobe's integrand f, an OM mirror bounce's weight is its Fresnel term.
//
// Per sample, the lane replays the path twice with the same Philox stream:
// the first replay is the forward per-sample function (its radiance C), the
// second walks the same vertices and, at bounce i, adds
//     A . R_i . d ln g_i / d theta
// to the parameters of that vertex, where A = d loss / d C and R_i is C minus
// the emission terms collected before bounce i (every later term carries
// g_i).  Parameter gradients go to texels through the bilinear weights of the
// lookup (4 texels per channel) and to the OM material's (n, k), with atomics.
// No path is stored between the passes.
#include <cstring>

#include "path_device.h"

namespace {

struct AdjTargets {
    const int* texSlot;      // per scene texture: index into texGrad, or -1
    float* const* texGrad;   // w*h*4 floats per target texture
    const int* omSlot;       // per scene material: index into omGrad, or -1
    float* omGrad;           // 6 floats per target material: dn rgb, dk rgb
};

DEV void scatterTex(const OmvSceneView& S, const AdjTargets& T, const OmvMaterial& M, const SP& sp, int slot, int ch, float g) {
    const int id = M.tex_map[slot];
    if (!(0 <= id && id < S.num_texs) || !(g != 0.0f)) return;
    const int ti = T.texSlot[id];
    if (ti < 0) return;
    float u, v;
    if (!getTexCoord(S, sp.mesh, sp.face, sp.bary, M.tex_ch[slot], u, v)) return;
    f3 uvT = mul33(M.tex_xf[slot], mk3(u, v, 1.0f));
    int idx[4];
    float w[4];
    if (!texAddress(S.texs[id], uvT.x, uvT.y, idx, w)) return;
    float* G = T.texGrad[ti];
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(G + (size_t)idx[i] * 4 + ch, g * w[i]);
}

// ------------------------------------------ FP64 detached evaluation
// The functional the adjoint differentiates, evaluated in double on the
// device's own (float) paths: directions, lobes and hits come from the
// sampling scene S0 exactly as the forward samples them; each bounce's weight
// is re-evaluated with the parameters of the evaluation scene S1 --
// w_ref * f1 / f0 for the sampled PBR lobe, the Fresnel term of S1's (n, k)
// for an OM mirror -- with the material inputs, lobe values, Fresnel terms,
// throughput and radiance in double.  Central differences of an image loss
// over this film are then free of float rounding in the loss difference.
struct dcpx { double re, im; };
DEV dcpx dmul(dcpx a, dcpx b) { dcpx r; r.re = a.re * b.re - a.im * b.im; r.im = a.re * b.im + a.im * b.re; return r; }
DEV dcpx dsqrtc(dcpx z) {   // principal branch
    double m = sqrt(z.re * z.re + z.im * z.im);
    double re = sqrt(0.5 * (m + fabs(z.re)));
    dcpx r;
    if (re == 0.0) { r.re = 0.0; r.im = 0.0; return r; }
    if (z.re >= 0.0) { r.re = re; r.im = z.im / (2.0 * re); }
    else { r.re = fabs(z.im) / (2.0 * re); r.im = z.im < 0.0 ? -re : re; }
    return r;
}
DEV dcpx ddiv(dcpx a, dcpx b) {
    double den = b.re * b.re + b.im * b.im;
    dcpx r; r.re = (a.re * b.re + a.im * b.im) / den; r.im = (a.im * b.re - a.re * b.im) / den;
    return r;
}
// fresnelReflection (OpenMaterial.cpp:20-38) in double
DEV void fresnelD(double n, double k, double c, double& Rp, double& Rs) {
    dcpx N; N.re = n; N.im = k;
    dcpx N2 = dmul(N, N);
    c = fabs(c);
    const double s2 = 1.0 - c * c;
    dcpx z; z.re = N2.re - s2; z.im = N2.im;
    dcpx sq = dsqrtc(z);
    dcpx np, dp, ns, ds;
    np.re = N2.re * c - sq.re; np.im = N2.im * c - sq.im;
    dp.re = N2.re * c + sq.re; dp.im = N2.im * c + sq.im;
    ns.re = c - sq.re; ns.im = -sq.im;
    ds.re = c + sq.re; ds.im = sq.im;
    dcpx rp = ddiv(np, dp), rs = ddiv(ns, ds);
    Rp = rp.re * rp.re + rp.im * rp.im;
    Rs = rs.re * rs.re + rs.im * rs.im;
}
// the bilinear lookup of texLookup with the texel values and sums in double
// (same texels and float weights)
DEV bool texLookupD(const OmvSceneView& S, int id, float u, float v, double o[4]) {
    const OmvTexture t = S.texs[id];
    int i[4];
    float w[4];
    if (!texAddress(t, u, v, i, w)) { o[0] = o[1] = o[2] = 0.5; o[3] = 1.0; return true; }
    for (int c = 0; c < 4; ++c) o[c] = 0.0;
    if (t.fmt == OMV_TEX_RGBA8) {
        const uint32_t* px = (const uint32_t*)(S.texels + t.offset);
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 4; ++c) o[c] += (double)((px[i[k]] >> (8 * c)) & 0xFF) * (double)w[k] / 255.0;
    } else {
        const float4* px = (const float4*)(S.texels + t.offset);
        for (int k = 0; k < 4; ++k) {
            const float4 q = px[i[k]];
            o[0] += (double)q.x * w[k]; o[1] += (double)q.y * w[k]; o[2] += (double)q.z * w[k]; o[3] += (double)q.w * w[k];
        }
    }
    return true;
}
DEV bool pbrTexD(const OmvSceneView& S, const SP& sp, const OmvMaterial& M, int slot, double o[4]) {
    int id = M.tex_map[slot];
    if (!(0 <= id && id < S.num_texs)) return false;
    float u, v;
    if (!getTexCoord(S, sp.mesh, sp.face, sp.bary, M.tex_ch[slot], u, v)) return false;
    f3 uvT = mul33(M.tex_xf[slot], mk3(u, v, 1.0f));
    return texLookupD(S, id, uvT.x, uvT.y, o);
}
// pbrInputs + pbrFinish in double: occluded base colour, metal, alpha = r^2
DEV void pbrParamsD(const OmvSceneView& S, const SP& sp, const OmvMaterial& M, double col[3], double& metal, double& alpha) {
    double t[4], c4[4] = {1.0, 1.0, 1.0, 1.0}, omr[3] = {1.0, (double)M.metal, (double)M.rough};
    if (pbrTexD(S, sp, M, OMV_PBR_BASE_COLOR, t)) for (int i = 0; i < 4; ++i) c4[i] = t[i];
    for (int i = 0; i < 3; ++i) c4[i] *= (double)M.bcf[i];
    if (pbrTexD(S, sp, M, OMV_PBR_METALLIC_ROUGHNESS, t)) { omr[0] *= t[0]; omr[1] *= t[2]; omr[2] *= t[1]; }
    if (M.occ_separate && pbrTexD(S, sp, M, OMV_PBR_OCCLUSION, t)) omr[0] = t[0];
    const double os = M.occ_strength;
    for (int i = 0; i < 3; ++i) col[i] = (omr[0] * c4[i] - c4[i]) * os + c4[i];
    metal = omr[1];
    const double r = fmin(fmax(omr[2], 0.00001), 0.99999);
    alpha = r * r;
}
// the sampled lobe's integrand (diffuse (1-F)(1-metal)color, reflection
// F D G, transmission (1-F) D G) on the float geometry, in double
DEV void lobeValueD(const LobeGeo& g, const double color[3], double metal, double alpha, double f[3]) {
    double D = 1.0, G = 1.0;
    if (g.lobe >= 2) {
        const double NH = g.NH, a2 = alpha * alpha, l = (a2 - 1.0) * NH * NH + 1.0;
        D = a2 / (3.14159265358979323846 * l * l);
        auto G1 = [&](double c) {
            if (!(alpha > 0.0 || c < 0.99)) return 1.0;
            const double a = c / (alpha * sqrt(1.0 - c * c));
            return 1.0 / (1.0 + 0.5 * (sqrt(1.0 + 1.0 / (a * a)) - 1.0));
        };
        const double x = G1(g.VN), y = G1(g.LN);
        G = x * y / (x + y - x * y);
    }
    for (int c = 0; c < 3; ++c) {
        const double F0 = (color[c] - 0.04) * metal + 0.04;
        const double F = (1.0 - (double)g.p5) * F0 + (double)g.p5;
        f[c] = g.lobe == 1 ? (1.0 - F) * (1.0 - metal) * color[c] : (g.lobe == 2 ? F * D * G : (g.lobe == 3 ? (1.0 - F) * D * G : 0.0));
    }
}
// one sample of the camera-path estimator (pathSample, no point lights)
// with the detached FP64 weights
DEV void detachedSample(const OmvSceneView& S0, const OmvSceneView& S1, const omv_camera_desc& cam, const RP& P, int spp, uint64_t s,
                        Lane& L, double out[3]) {
    const int q = P.q;
    const uint64_t pix = s / (uint64_t)spp;
    const int x = (int)(pix % (uint64_t)cam.width), y = (int)(pix / (uint64_t)cam.width);
    Rng cs; cs.init(P.seed, STREAM_CAMERA, s);
    float u0 = cs.next(), u1 = cs.next();
    f3 O, D;
    cameraRay(cam, x, y, u0, u1, O, D);
    double of[3] = {1.0, 1.0, 1.0}, rad[3] = {0.0, 0.0, 0.0};
    f3 pPos = O, pDir = D;
    float tnear = 0.0f;
    int n = 1;
    bool exitScene = false;
    for (int i = 1; i < P.camera_bounces + 1; ++i) {
        float delta = 3.0f * FLT_EPSILON * maxAbs(pPos);
        f3 org = mk3(pPos.x + delta * pDir.x, pPos.y + delta * pDir.y, pPos.z + delta * pDir.z);
        ++L.camRays;
        Hit h;
        if (!sceneIntersect(S0, org, pDir, tnear, FLT_MAX, h, L.st, L.cnt, L.rays)) { exitScene = true; break; }
        SP sp;
        computeSurfacePoint(S0, h, sp, false);
        modifyFrame(S0, sp);
        f3 outRay, brdf, erad;
        BounceInfo bi;
        matGetRayAndBrdf(S0, P.x, sp, pDir, outRay, brdf, erad, true, cs, &bi);
        double w[3] = {brdf.x, brdf.y, brdf.z};
        const int mid = S0.meshes[sp.mesh].mat_id;
        if (bi.kind == 1 && bi.b.lobe != 0) {
            const LobeGeo g = lobeGeo(bi.b);
            double c0[3], c1[3], m0, m1, a0, a1, f0[3], f1[3];
            pbrParamsD(S0, sp, S0.mats[mid], c0, m0, a0);
            pbrParamsD(S1, sp, S1.mats[mid], c1, m1, a1);
            lobeValueD(g, c0, m0, a0, f0);
            lobeValueD(g, c1, m1, a1, f1);
            for (int c = 0; c < 3; ++c) w[c] = f0[c] != 0.0 ? w[c] * (f1[c] / f0[c]) : 0.0;
        } else if (bi.kind == 2) {
            const OmvMaterial& M1 = S1.mats[mid];
            for (int c = 0; c < 3; ++c) {
                double rp, rs;
                fresnelD(M1.n[c], M1.k[c], bi.fCos, rp, rs);
                w[c] = (rs - rp) * 0.5 + rp;
                w[c] /= fabs((double)bi.fCos);
            }
        }
        const double ct = fabs((double)dot3(outRay, sp.n));
        for (int c = 0; c < 3; ++c) of[c] = of[c] * w[c] * ct;
        rad[0] += of[0] * erad.x; rad[1] += of[1] * erad.y; rad[2] += of[2] * erad.z;
        ++n;
        pPos = sp.pos; pDir = outRay;
        if (of[0] == 0.0 && of[1] == 0.0 && of[2] == 0.0) break;
        if (!S0.any_emissive && n > q) break;
    }
    if (exitScene && S0.bg_type != OMV_BG_NONE && n <= q) {
        const f3 bg = background(S0, pDir);
        rad[0] += bg.x * of[0]; rad[1] += bg.y * of[1]; rad[2] += bg.z * of[2];
    }
    out[0] = rad[0]; out[1] = rad[1]; out[2] = rad[2];
}

}  // namespace

// The FP64 detached film: one lane per pixel, its spp samples in sample order.
__global__ __launch_bounds__(OMV_BLOCK) void omv_detached_kernel(OmvSceneView S0, OmvSceneView S1, omv_camera_desc cam, RP P, int spp,
                                                                 uint64_t npix, double* __restrict__ film, uint32_t* __restrict__ gstack,
                                                                 unsigned long long* __restrict__ counters) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Lane L;
    laneInit(L, gstack, nullptr, gtid, gthreads);
    for (uint64_t p = gtid; p < npix; p += gthreads) {
        double a[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < spp; ++k) {
            double c[3];
            detachedSample(S0, S1, cam, P, spp, p * (uint64_t)spp + (uint64_t)k, L, c);
            a[0] += c[0]; a[1] += c[1]; a[2] += c[2];
        }
        const double sc = spp > 0 ? 1.0 / (double)spp : 1.0;
        film[p * 4 + 0] = a[0] * sc; film[p * 4 + 1] = a[1] * sc; film[p * 4 + 2] = a[2] * sc; film[p * 4 + 3] = 1.0;
    }
    flushCounters(counters, L.rays, L.camRays, L.lightRays, L.shadowRays, L.cnt);
}

__global__ __launch_bounds__(OMV_BLOCK) void omv_adjoint_kernel(OmvSceneView S, omv_camera_desc cam, RP P, int spp, uint64_t sample0,
                                                                uint64_t nsamples, const float* __restrict__ dfilm, AdjTargets T,
                                                                float* __restrict__ lv, uint32_t* __restrict__ gstack,
                                                                unsigned long long* __restrict__ counters) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Lane L;
    laneInit(L, gstack, lv, gtid, gthreads);
    const int q = P.q;
    const float cnt = (float)spp;
    const float sc = cnt < 0.5f ? 1.0f : 1.0f / cnt;
    for (uint64_t sl = gtid; sl < nsamples; sl += gthreads) {
        const uint64_t s = sample0 + sl;
        const uint64_t pix = s / (uint64_t)spp;
        const float A[3] = {dfilm[pix * 4 + 0] * sc, dfilm[pix * 4 + 1] * sc, dfilm[pix * 4 + 2] * sc};
        if (A[0] == 0.0f && A[1] == 0.0f && A[2] == 0.0f) continue;
        const f3 C = pathSample<false, false>(S, cam, P, spp, s, L);
        if (C.x == 0.0f && C.y == 0.0f && C.z == 0.0f) continue;
        // second replay: the camera loop of pathSample, same draws
        const int x = (int)(pix % (uint64_t)cam.width), y = (int)(pix / (uint64_t)cam.width);
        Rng cs; cs.init(P.seed, STREAM_CAMERA, s);
        float u0 = cs.next(), u1 = cs.next();
        f3 O, D;
        cameraRay(cam, x, y, u0, u1, O, D);
        float rem[3] = {C.x, C.y, C.z};
        f3 pPos = O, pDir = D, of = mk3(1.0f, 1.0f, 1.0f);
        float tnear = 0.0f;
        int n = 1;
        for (int i = 1; i < P.camera_bounces + 1; ++i) {
            float delta = 3.0f * FLT_EPSILON * maxAbs(pPos);
            f3 org = mk3(pPos.x + delta * pDir.x, pPos.y + delta * pDir.y, pPos.z + delta * pDir.z);
            Hit h;
            if (!sceneIntersect(S, org, pDir, tnear, FLT_MAX, h, L.st, L.cnt, L.rays)) break;
            SP sp;
            computeSurfacePoint(S, h, sp, false);
            modifyFrame(S, sp);
            f3 outRay, brdf, erad;
            BounceInfo bi;
            matGetRayAndBrdf(S, P.x, sp, pDir, outRay, brdf, erad, true, cs, &bi);
            const float wc[3] = {A[0] * rem[0], A[1] * rem[1], A[2] * rem[2]};
            const OmvMaterial& M = S.mats[sp.mat];
            if (bi.kind == 1 && bi.b.lobe != 0) {
                const LobeGeo g = lobeGeo(bi.b);
                const float col[3] = {bi.b.c0, bi.b.c1, bi.b.c2};
                float dC[3], dM[3], dA[3];
                lobeDlog(g, col, bi.b.metal, bi.b.alpha, dC, dM, dA);
                const float occ = (bi.occ - 1.0f) * M.occ_strength + 1.0f;
                float gMetal = 0.0f, gAlpha = 0.0f;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    scatterTex(S, T, M, sp, OMV_PBR_BASE_COLOR, c, wc[c] * dC[c] * occ * M.bcf[c]);
                    gMetal += wc[c] * dM[c];
                    gAlpha += wc[c] * dA[c];
                }
                scatterTex(S, T, M, sp, OMV_PBR_METALLIC_ROUGHNESS, 2, gMetal * M.metal);
                if (bi.roughRaw > 0.00001f && bi.roughRaw < 0.99999f)
                    scatterTex(S, T, M, sp, OMV_PBR_METALLIC_ROUGHNESS, 1, gAlpha * 2.0f * bi.b.rough * M.rough);
            } else if (bi.kind == 2) {
                const int slot = T.omSlot[sp.mat];
                if (slot >= 0) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        float dn, dk;
                        fresnelDlog(M.n[c], M.k[c], bi.fCos, dn, dk);
                        if (wc[c] * dn != 0.0f) atomicAdd(T.omGrad + slot * 6 + c, wc[c] * dn);
                        if (wc[c] * dk != 0.0f) atomicAdd(T.omGrad + slot * 6 + 3 + c, wc[c] * dk);
                    }
                }
            }
            float ct = fabsf(dot3(outRay, sp.n));
            of = mk3(of.x * brdf.x * ct, of.y * brdf.y * ct, of.z * brdf.z * ct);
            rem[0] -= of.x * erad.x; rem[1] -= of.y * erad.y; rem[2] -= of.z * erad.z;
            ++n;
            pPos = sp.pos; pDir = outRay;
            if (of.x == 0.0f && of.y == 0.0f && of.z == 0.0f) break;
            if (!S.any_emissive && n > q) break;
        }
    }
    flushCounters(counters, L.rays, L.camRays, L.lightRays, L.shadowRays, L.cnt);
}

namespace omv {

int adjointBlocksPerCU() {
    int n = 0;
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_adjoint_kernel, OMV_BLOCK, lds) != hipSuccess || n < 1) n = 1;
    return n;
}

int detachedBlocksPerCU() {
    int n = 0;
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_detached_kernel, OMV_BLOCK, lds) != hipSuccess || n < 1) n = 1;
    return n;
}

hipError_t launchDetached(const OmvSceneView& S0, const OmvSceneView& S1, const omv_camera_desc& cam, const omv_renderer_params& p,
                          uint64_t npix, double* film, uint32_t* gstack, unsigned long long* counters, int blocks, hipStream_t stream) {
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    hipLaunchKernelGGL(omv_detached_kernel, dim3(blocks), dim3(OMV_BLOCK), lds, stream, S0, S1, cam, toRP(p), p.samples_per_pixel, npix,
                       film, gstack, counters);
    return hipGetLastError();
}

hipError_t launchAdjoint(const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, uint64_t sample0,
                         uint64_t nsamples, const float* dfilm, const int* texSlot, float* const* texGrad, const int* omSlot,
                         float* omGrad, float* lv, uint32_t* gstack, unsigned long long* counters, int blocks, hipStream_t stream) {
    AdjTargets T;
    T.texSlot = texSlot; T.texGrad = texGrad; T.omSlot = omSlot; T.omGrad = omGrad;
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    hipLaunchKernelGGL(omv_adjoint_kernel, dim3(blocks), dim3(OMV_BLOCK), lds, stream, S, cam, toRP(p), p.samples_per_pixel, sample0,
                       nsamples, dfilm, T, lv, gstack, counters);
    return hipGetLastError();
}

}  // namespace omv
