// kernels.hip -- gfx950 (MI355X) kernels of the BDPT render core.
//
// One persistent sample kernel replaces BDPTRenderer::renderRow/computePaths
// (pathtracer/src/renderer/BDPTRenderer.cpp:237-370): each lane owns one
// sample (pixel, k) at a time, walks its light subpath, then its camera
// subpath, and makes every light/camera vertex connection as soon as the
// camera vertex exists, so camera vertices never leave registers.  Light
// vertices (position + carried radiance, 24 B each) go to a per-lane HBM
// scratch laid out [vertex][component][lane] for coalesced access.
//
// Ray casting replaces Embree's rtcIntersect1 with a two-level BVH4
// traversal (128-byte nodes, stack in LDS with a global spill area) and the
// watertight ray/triangle test of Woop, Benthin & Wald (JCGT 2013), closest
// hit with a deterministic (t, instance, primitive) tie-break.
//
// Samples are written per sample and reduced per pixel in sample order by a
// second kernel, which is the reference's film summation order
// (PinholeCamera::hit, PinholeCamera.cpp:97-108).  The same per-sample
// function drives the batched multi-job kernel, and the false-colour pass
// lives here too.  The device building blocks are in path_device.h.
#include <cstring>

#include "path_device.h"

// ============================================================== kernels
// Plain render: one sample per lane at a time over [sample0, sample0 + nsamples),
// persistent grid; radiance per sample for the ordered film reduction.
template <bool kLights, bool kNee, bool kSplat>
__global__ __launch_bounds__(OMV_BLOCK) void omv_bdpt_kernel(OmvSceneView S, omv_camera_desc cam, RP P, int spp,
                                                             uint64_t sample0, uint64_t nsamples, float* __restrict__ rad,
                                                             float* __restrict__ lv, uint32_t* __restrict__ gstack,
                                                             unsigned long long* __restrict__ counters, float* __restrict__ splat) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Lane L;
    laneInit(L, gstack, lv, gtid, gthreads);
    L.splat = splat;
    for (uint64_t sl = gtid; sl < nsamples; sl += gthreads) {
        f3 r = pathSample<kLights, kNee, kSplat>(S, cam, P, spp, sample0 + sl, L);
        rad[sl * 3 + 0] = r.x;
        rad[sl * 3 + 1] = r.y;
        rad[sl * 3 + 2] = r.z;
    }
    flushCounters(counters, L.rays, L.camRays, L.lightRays, L.shadowRays, L.cnt);
}

// Batched validation jobs in one persistent launch.  Work unit: a tile of
// 64 x tileWaves consecutive pixels of one job, claimed by a whole wave
// through one atomic on the work counter; each lane renders all spp samples
// of its pixel in sample order (the reference's PinholeCamera::hit order)
// and writes the normalised pixel, so no per-sample buffer is needed and the
// film equals the plain render's bit for bit.
struct OmvJobDev {
    OmvSceneView S;
    omv_camera_desc cam;
    RP P;
    int spp;
    int pad;
    uint64_t npix;
    float* film;
};

template <bool kNee>
__global__ __launch_bounds__(OMV_BLOCK) void omv_batch_kernel(const OmvJobDev* __restrict__ jobs, int njobs,
                                                              const uint64_t* __restrict__ tileEnd, uint64_t totalTiles,
                                                              int tileWaves, unsigned long long* __restrict__ work,
                                                              float* __restrict__ lv, uint32_t* __restrict__ gstack,
                                                              unsigned long long* __restrict__ counters) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    Lane L;
    laneInit(L, gstack, lv, gtid, gthreads);
    for (;;) {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(work, 1ull);
        t = __shfl(t, 0, 64);
        if (t >= totalTiles) break;
        // job of tile t: first job whose tile range ends after t (wave-uniform)
        int lo = 0, hi = njobs - 1;
        while (lo < hi) { int mid = (lo + hi) >> 1; if (tileEnd[mid] > t) hi = mid; else lo = mid + 1; }
        const int job = __builtin_amdgcn_readfirstlane(lo);
        const OmvJobDev& J = jobs[job];
        const uint64_t first = (job ? tileEnd[job - 1] : 0ull);
        const uint64_t pix0 = (t - first) * 64ull * (uint64_t)tileWaves;
        for (int w = 0; w < tileWaves; ++w) {
            const uint64_t p = pix0 + (uint64_t)w * 64ull + (uint64_t)lane;
            if (p >= J.npix) continue;
            reinterpret_cast<float4*>(J.film)[p] = renderPixel<true, kNee>(J.S, J.cam, J.P, J.spp, p, L);
        }
    }
    flushCounters(counters, L.rays, L.camRays, L.lightRays, L.shadowRays, L.cnt);
}

// Zero-fill usable inside a captured graph (no memset nodes).
__global__ void omv_fill_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Film: PinholeCamera::init/hit/stop (PinholeCamera.cpp:65-108) for whole pixels.
// A block reduces kFilmPix pixels.  Their samples are staged through LDS in
// rounds of kFilmK samples per pixel -- each pixel's round is one contiguous
// run of 3 * kFilmK floats, loaded by the lanes of one wave in one access --
// and each thread then adds its own pixel's staged samples in sample order, so
// the sums are the sequential ones.  Rows of the stage are 3 * kFilmK + 1
// floats apart (odd), so the threads' reads fall in distinct banks.
constexpr int kFilmPix = 256, kFilmK = 16, kFilmRow = 3 * kFilmK + 1;
__global__ __launch_bounds__(kFilmPix) void omv_film_kernel(const float* __restrict__ rad, int spp, uint64_t pix0, uint64_t npix,
                                                            float* __restrict__ film) {
    __shared__ float stage[kFilmPix * kFilmRow];
    const uint64_t b0 = (uint64_t)blockIdx.x * kFilmPix;
    const int np = (int)min<uint64_t>((uint64_t)kFilmPix, npix - b0);
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int k0 = 0; k0 < spp; k0 += kFilmK) {
        const int per = 3 * min(kFilmK, spp - k0);   // floats per pixel in this round (<= 48 < 64 lanes)
        __syncthreads();                              // the previous round is consumed
        if (lane < per) {
            // 16 pixel runs in flight per wave before they are stored
            for (int q0 = wave; q0 < np; q0 += 16 * (kFilmPix / 64)) {
                float v[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int q = q0 + i * (kFilmPix / 64);
                    v[i] = q < np ? rad[((b0 + (uint64_t)q) * (uint64_t)spp + (uint64_t)k0) * 3 + (uint64_t)lane] : 0.0f;
                }
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int q = q0 + i * (kFilmPix / 64);
                    if (q < np) stage[q * kFilmRow + lane] = v[i];
                }
            }
        }
        __syncthreads();
        if (t < np) {
            const float* r = stage + t * kFilmRow;
            for (int k = 0; 3 * k < per; ++k) { a0 = a0 + r[3 * k]; a1 = a1 + r[3 * k + 1]; a2 = a2 + r[3 * k + 2]; }
        }
    }
    if (t >= np) return;
    const uint64_t p = b0 + (uint64_t)t;
    float cnt = (float)spp;
    float sc = cnt < 0.5f ? 1.0f : 1.0f / cnt;
    float4 o; o.x = sc * a0; o.y = sc * a1; o.z = sc * a2; o.w = 1.0f;
    reinterpret_cast<float4*>(film)[pix0 + p] = o;
}

// OMV_EXT_LIGHT_SPLAT: film += splat sums / (number of light paths).
__global__ void omv_splat_add_kernel(const float* __restrict__ splat, uint64_t npix, float inv, float* __restrict__ film) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    float4 o = reinterpret_cast<float4*>(film)[p];
    o.x = o.x + splat[p * 3 + 0] * inv;
    o.y = o.y + splat[p * 3 + 1] * inv;
    o.z = o.z + splat[p * 3 + 2] * inv;
    reinterpret_cast<float4*>(film)[p] = o;
}

// False-colour pass: one camera bounce per pixel (main.cpp:261-270 set-up).
__global__ __launch_bounds__(OMV_BLOCK) void omv_fc_kernel(OmvSceneView S, omv_camera_desc cam, uint32_t seed, float* __restrict__ film,
                                                           int32_t* __restrict__ ids, uint8_t* __restrict__ tie, uint32_t* __restrict__ gstack,
                                                           unsigned long long* __restrict__ counters) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    st.init(gstack, gtid, gthreads);
    Counters cnt; cnt.nodes = 0; cnt.tris = 0; cnt.overflow = 0;
    uint32_t rays = 0, camRays = 0;
    const uint64_t npix = (uint64_t)cam.width * cam.height;
    const bool fcDensity = S.fc_subject == OMV_FC_MD || S.fc_subject == OMV_FC_SG;
    for (uint64_t p = gtid; p < npix; p += gthreads) {
        const int x = (int)(p % (uint64_t)cam.width), y = (int)(p / (uint64_t)cam.width);
        Rng cs; cs.init(seed, STREAM_CAMERA, p);
        float u0 = cs.next(), u1 = cs.next();
        f3 O, D;
        cameraRay(cam, x, y, u0, u1, O, D);
        float delta = 3.0f * FLT_EPSILON * maxAbs(O);
        f3 org = mk3(O.x + delta * D.x, O.y + delta * D.y, O.z + delta * D.z);
        float tnear = 0.0f;
        Hit h;
        ++camRays;
        f3 col = mk3(0.0f, 0.0f, 0.0f);
        int32_t i0 = -1, i1 = -1, i2 = -1, i3 = -1;
        uint8_t tflag = 0;
        if (sceneIntersect(S, org, D, tnear, FLT_MAX, h, st, cnt, rays)) {
            SP sp;
            computeSurfacePoint(S, h, sp, fcDensity);
            modifyFrame(S, sp);
            const OmvMesh& mesh = S.meshes[sp.mesh];
            const OmvMaterial& M = S.mats[mesh.mat_id];
            const bool missing = M.type == OMV_MAT_DIFFUSE;
            f3 inDir = D;
            if (missing) {   // RenderMaterial::getRayAndBrdf draws the direction first
                f3 v = uniformHemisphere(cs);
                inDir = neg3(toBasis(sp.t, sp.b, sp.n, v));
            }
            const int sub = S.fc_subject;
            const int globalFace = sp.face + mesh.face_base;
            if (sub == OMV_FC_PID || sub == OMV_FC_GID || sub == OMV_FC_MID || sub == OMV_FC_IID || sub == OMV_FC_IPID) {
                // EnumeratedFCSubject::paint (FCSubject.h:172-177): the id's table colour, black past the table
                uint32_t id;
                if (sub == OMV_FC_PID) id = (uint32_t)globalFace;
                else if (sub == OMV_FC_GID) id = (uint32_t)sp.mesh;
                else if (sub == OMV_FC_MID) id = (uint32_t)mesh.mat_id;
                else if (sub == OMV_FC_IID) id = (uint32_t)sp.inst;
                else id = S.fc_inst_base[sp.inst] + (uint32_t)sp.face;
                const uint32_t cc = id < S.fc_table_size ? S.fc_table[id] : 0u;
                float r = (float)((cc >> 16) & 0xFF) / 255.0f, g = (float)((cc >> 8) & 0xFF) / 255.0f, b = (float)(cc & 0xFF) / 255.0f;
                if (sub == OMV_FC_MID) col = missing ? mk3(1000.0f, 0.0f, 1000.0f) : mk3(0.0f, g, b);
                else col = mk3(r, g, b);
            } else if (sub == OMV_FC_RMP || sub == OMV_FC_MMP) {
                float a = 0.0f;
                if (M.type == OMV_MAT_PBR) {
                    float c4[4], omr[3];
                    pbrInputs(S, sp, M, c4, omr);
                    a = sub == OMV_FC_RMP ? omr[2] : omr[1];
                }
                a = clampr(a, 0.0f, 1.0f);
                col = sub == OMV_FC_RMP ? mk3(lerpr(1.0f, 0.0f, a), lerpr(1.0f, 1.0f, a), lerpr(1.0f, 0.0f, a))
                                        : mk3(lerpr(1.0f, 0.0f, a), lerpr(1.0f, 0.0f, a), lerpr(1.0f, 1.0f, a));
            } else if (sub == OMV_FC_MD) {
                float a = clampr(sp.density, 0.0f, 1.0f);
                col = mk3(lerpr(1.0f, 1.0f, a), lerpr(1.0f, 0.0f, a), lerpr(1.0f, 0.0f, a));
            } else if (sub == OMV_FC_SG) {
                float g = clampr(sp.gradient, 0.0f, 1.0f);
                col = mk3(1.0f - g, 1.0f - g, 1.0f - g);
            } else if (sub == OMV_FC_MN) {
                col = M.type == OMV_MAT_OM ? mk3(0.0f, 1.0f, 0.0f) : (M.type == OMV_MAT_PBR ? mk3(0.0f, 0.0f, 1.0f) : mk3(1000.0f, 0.0f, 1000.0f));
            } else if (sub == OMV_FC_IN) {
                f3 tfn = faceNormal(S, sp);
                float inv = dot3(inDir, tfn) > 0.0f ? 1.0f : 0.0f;
                col = mk3(inv, 1.0f - inv, 0.0f);
            }
            i0 = globalFace; i1 = sp.mesh; i2 = missing ? -1 : mesh.mat_id; i3 = sp.inst;
            tflag = h.tie ? 1 : 0;
        }
        float4 o; o.x = col.x; o.y = col.y; o.z = col.z; o.w = 1.0f;
        reinterpret_cast<float4*>(film)[p] = o;
        if (ids) { ids[p] = i0; ids[npix + p] = i1; ids[2 * npix + p] = i2; ids[3 * npix + p] = i3; }
        if (tie) tie[p] = tflag;
    }
    flushCounters(counters, rays, camRays, 0, 0, cnt);
}

// ----------------------------------------------------------- launchers
namespace omv {


hipError_t launchBdpt(const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, uint64_t sample0,
                      uint64_t nsamples, float* rad, float* lv, uint32_t* gstack, unsigned long long* counters, int blocks,
                      float* splat, hipStream_t stream) {
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    RP P = toRP(p);
    if (!S.env_tex_pdf) P.x.nee = false;   // no HDR environment: nothing to sample
    if (!splat || S.num_lights <= 0) P.x.splat = false;
    const dim3 g(blocks), b(OMV_BLOCK);
    const int spp = p.samples_per_pixel;
    if (S.num_lights > 0) {
        if (P.x.splat && P.x.nee)
            hipLaunchKernelGGL((omv_bdpt_kernel<true, true, true>), g, b, lds, stream, S, cam, P, spp, sample0, nsamples, rad, lv, gstack,
                               counters, splat);
        else if (P.x.splat)
            hipLaunchKernelGGL((omv_bdpt_kernel<true, false, true>), g, b, lds, stream, S, cam, P, spp, sample0, nsamples, rad, lv, gstack,
                               counters, splat);
        else if (P.x.nee)
            hipLaunchKernelGGL((omv_bdpt_kernel<true, true, false>), g, b, lds, stream, S, cam, P, spp, sample0, nsamples, rad, lv, gstack,
                               counters, splat);
        else
            hipLaunchKernelGGL((omv_bdpt_kernel<true, false, false>), g, b, lds, stream, S, cam, P, spp, sample0, nsamples, rad, lv, gstack,
                               counters, splat);
    } else if (P.x.nee) {
        hipLaunchKernelGGL((omv_bdpt_kernel<false, true, false>), g, b, lds, stream, S, cam, P, spp, sample0, nsamples, rad, lv, gstack,
                           counters, splat);
    } else {
        hipLaunchKernelGGL((omv_bdpt_kernel<false, false, false>), g, b, lds, stream, S, cam, P, spp, sample0, nsamples, rad, lv, gstack,
                           counters, splat);
    }
    return hipGetLastError();
}

hipError_t launchFilm(const float* rad, int spp, uint64_t pix0, uint64_t npix, float* film, hipStream_t stream) {
    if (npix == 0) return hipSuccess;
    const uint64_t blocks = (npix + kFilmPix - 1) / kFilmPix;
    hipLaunchKernelGGL(omv_film_kernel, dim3((unsigned)blocks), dim3(kFilmPix), 0, stream, rad, spp, pix0, npix, film);
    return hipGetLastError();
}

hipError_t launchSplatAdd(const float* splat, uint64_t npix, float inv, float* film, hipStream_t stream) {
    const int tb = 256;
    hipLaunchKernelGGL(omv_splat_add_kernel, dim3((unsigned)((npix + tb - 1) / tb)), dim3(tb), 0, stream, splat, npix, inv, film);
    return hipGetLastError();
}

hipError_t launchFc(const OmvSceneView& S, const omv_camera_desc& cam, uint32_t seed, float* film, int32_t* ids, uint8_t* tie,
                    uint32_t* gstack, unsigned long long* counters, int blocks, hipStream_t stream) {
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    hipLaunchKernelGGL(omv_fc_kernel, dim3(blocks), dim3(OMV_BLOCK), lds, stream, S, cam, seed, film, ids, tie, gstack, counters);
    return hipGetLastError();
}

size_t batchJobBytes() { return sizeof(OmvJobDev); }

void packBatchJob(void* dst, const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, float* film,
                  uint64_t npix) {
    OmvJobDev j;
    std::memset(&j, 0, sizeof(j));
    j.S = S; j.cam = cam; j.P = toRP(p);
    if (!S.env_tex_pdf) j.P.x.nee = false;
    j.spp = p.samples_per_pixel; j.npix = npix; j.film = film;
    std::memcpy(dst, &j, sizeof(j));
}

hipError_t launchFillU64(unsigned long long* p, int n, hipStream_t stream) {
    hipLaunchKernelGGL(omv_fill_u64, dim3((n + 63) / 64), dim3(64), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launchBatch(const void* jobs, int njobs, const uint64_t* tileEnd, uint64_t totalTiles, int tileWaves,
                       unsigned long long* counters, float* lv, uint32_t* gstack, int blocks, bool nee, hipStream_t stream) {
    // counters[0..7]: statistics, counters[8]: work counter; cleared by a
    // kernel so that the sequence can be captured into a graph
    hipLaunchKernelGGL(omv_fill_u64, dim3(1), dim3(64), 0, stream, counters, 16);
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    if (nee)
        hipLaunchKernelGGL(omv_batch_kernel<true>, dim3(blocks), dim3(OMV_BLOCK), lds, stream, static_cast<const OmvJobDev*>(jobs), njobs,
                           tileEnd, totalTiles, tileWaves, counters + 8, lv, gstack, counters);
    else
        hipLaunchKernelGGL(omv_batch_kernel<false>, dim3(blocks), dim3(OMV_BLOCK), lds, stream, static_cast<const OmvJobDev*>(jobs), njobs,
                           tileEnd, totalTiles, tileWaves, counters + 8, lv, gstack, counters);
    return hipGetLastError();
}

int occupancyBlocksPerCU(int which) {
    int n = 0;
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    hipError_t e;
    switch (which) {
        case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_bdpt_kernel<true, false, false>, OMV_BLOCK, lds); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_bdpt_kernel<false, false, false>, OMV_BLOCK, lds); break;
        case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_batch_kernel<false>, OMV_BLOCK, lds); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_bdpt_kernel<true, true, false>, OMV_BLOCK, lds); break;
        case 5: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_bdpt_kernel<false, true, false>, OMV_BLOCK, lds); break;
        case 6: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_batch_kernel<true>, OMV_BLOCK, lds); break;
        case 7: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_bdpt_kernel<true, false, true>, OMV_BLOCK, lds); break;
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_bdpt_kernel<true, true, true>, OMV_BLOCK, lds); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, omv_fc_kernel, OMV_BLOCK, lds); break;
    }
    if (e != hipSuccess) n = 0;
    return n;
}

}  // namespace omv
