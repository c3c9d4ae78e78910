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

}  // namespace

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
            const OmvMaterial& M = S.mats[S.meshes[sp.mesh].mat_id];
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
                const int slot = T.omSlot[S.meshes[sp.mesh].mat_id];
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
