// renderer.cpp -- omv::PinholeCamera and omv::BDPTRenderer host classes.
// BDPTRenderer::render (BDPTRenderer.cpp:372-393) becomes: one persistent
// sample-kernel launch per chunk of whole pixels, then the ordered film
// reduction for that chunk; the film stays in HBM until the camera reads it.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "omv_kernels.h"
#include "fc_table.h"
#include "render_core.h"

namespace omv {

namespace {
constexpr uint64_t kChunkSamples = 1ull << 24;   // per-sample radiance buffer: 192 MiB
// Wavefront samples per chunk.  Every chunk pays the latency-bound tail of
// its last bounce rounds (a few long paths), so chunks are as large as HBM
// comfortably allows: ~6 GB of state at 2^25 samples, ~28 GB with point
// lights (light vertices, staged connections, connection rays) -- a tenth
// of one MI355X's 288 GB; chunk_samples lowers it.
constexpr uint64_t kWfSlots = 1ull << 25;
constexpr uint64_t kWfSlotsLights = 1ull << 25;
// Camera round from which the wavefront hands the remaining paths to one
// tail launch (omv_wf_tail).  Measured on MI355X (kernel ms, round by round
// -> tail from round 3 / 4 / 5 / 6): 2M-triangle instanced scene 1080p x 16
// 17.97 -> 21.22 / 17.99 / 17.64 / 17.81; torus knot 512^2 x 64 11.43 ->
// 11.55 / 10.73 / 10.82 / 10.99.  From round 5 both gain; earlier the tail
// launch holds too many paths, each traced one lane at a time.  Re-measured
// once the traversal and shading had become about 40 % faster (round by round
// -> tail from round 4 / 5 / 6 / 7): instanced (point lights: a connection
// trace per bounce) 11.13 -> 11.31 / 10.93 / 11.01 / 10.97; torus knot (no
// visibility rays) 6.68 -> 6.25 / 6.31 / 6.48 / 6.56.  So from round 5 with
// point lights or NEE, from round 4 without.
constexpr int kTailRoundVisibility = 5;
constexpr int kTailRoundPlain = 4;
// Round-1 camera rays are traced as wave packets (omv_wf_trace_cam1) only on
// request (tuning traversal = 2).  Measured on MI355X (kernel ms per render,
// per-ray -> packets), first with the branchy triangle setup: at 64 samples
// per pixel torus knot 512^2 8.55 -> 8.38, instanced 960x540 15.06 ->
// 14.89; at 16: instanced 1080p 15.06 -> 15.13.  Once the per-ray traversal
// got the branch-free triangle setup and stack fast paths: torus knot 512^2
// x 64 6.36 -> 6.41, 1024^2 x 16 6.23 -> 6.26, instanced 1080p x 16 12.48
// -> 12.61, x 4 4.74 -> 5.12.
int hipFail(hipError_t e, const char* what) {
    setError(std::string(what) + ": " + hipGetErrorString(e));
    return OMV_E_DEVICE;
}
}  // namespace

// PinholeCamera(w, h) (PinholeCamera.cpp:51-57).  The reference leaves the
// focal set-up to the caller; main() always follows with adjustToResolution
// and setYFoV(60) (main.cpp:484-487), which is the default here.
PinholeCamera::PinholeCamera(int w, int h) {
    std::memset(&desc, 0, sizeof(desc));
    desc.width = std::max(1, w);
    desc.height = std::max(1, h);
    desc.rotation[0] = desc.rotation[4] = desc.rotation[8] = 1.0f;
    adjustToResolution();
    setYFoV(60.0f);
}

// ViewPoint(loc, xdir, ydir, zdir) (RenderSensor.cpp:19-25): columns x, y, z
void PinholeCamera::setViewPoint(const float pos[3], const float x[3], const float y[3], const float z[3]) {
    for (int i = 0; i < 3; ++i) {
        desc.position[i] = pos[i];
        desc.rotation[i * 3 + 0] = x[i];
        desc.rotation[i * 3 + 1] = y[i];
        desc.rotation[i * 3 + 2] = z[i];
    }
}

void PinholeCamera::adjustToResolution() {   // PinholeCamera.cpp:156-160
    desc.half_h = 0.5f * (float)desc.height;
    desc.half_w = 0.5f * (float)desc.width;
}

int PinholeCamera::setYFoV(float fov) {      // PinholeCamera.cpp:167-174
    if (0.0f < fov && fov < 180.0f) {
        desc.focus = desc.half_h / std::tan(0.5f * fov * 3.14159265358979323846f / 180.0f);
        return OMV_OK;
    }
    setError("Field-of-view should be greater than zero and smaller than the straight angle.");
    return OMV_E_ARG;
}

// PinholeCamera::loadProperties (PinholeCamera.cpp:115-148) with the camera
// file's "pinhole" values already parsed; a null pointer is a key the file
// does not have (aspect and y-fov then count as 0, the focus stays).
void PinholeCamera::loadProperties(const float* aspect, const float* focus, const float* yfov) {
    const float a = aspect ? *aspect : 0.0f;
    const float fov = yfov ? *yfov : 0.0f;
    if (focus) desc.focus = *focus;
    desc.half_h = desc.focus * std::tan(0.5f * fov * 3.14159265358979323846f / 180.0f);
    desc.half_w = a * desc.half_h;
}

void PinholeCamera::setRealHeight(float height, float aspect) {   // PinholeCamera.cpp:150-154
    desc.half_h = std::max(0.001f, height);
    desc.half_w = std::max(0.1f, aspect) * desc.half_h;
}

BDPTRenderer::BDPTRenderer() { omv_renderer_params_default(&params_); }

BDPTRenderer::~BDPTRenderer() {
    for (auto& e : ev_) if (e) (void)hipEventDestroy(e);
    for (auto& e : traceEv_) (void)hipEventDestroy(e);
    for (auto& e : chunkEv_) if (e) (void)hipEventDestroy(e);
    if (hostStage_) (void)hipHostFree(hostStage_);
    if (stream2_) (void)hipStreamDestroy(stream2_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int BDPTRenderer::setSchedule(int s) {
    if (s != OMV_SCHEDULE_WAVEFRONT && s != OMV_SCHEDULE_PER_SAMPLE) { setError("set_schedule: unknown schedule"); return OMV_E_ARG; }
    schedule_ = s;
    return OMV_OK;
}

int BDPTRenderer::setTuning(const omv_render_tuning& t) {
    if (t.trace_keep < 0 || t.trace_keep > 64 || t.trace_blocks_per_cu < 0 || t.chunk_samples < 0 || t.traversal < 0 || t.traversal > 2 ||
        t.sort_rays < 0 || t.sort_rays > 2 || t.chunk_streams < 0 || t.chunk_streams > 2 || t.chunk_stagger < 0 || t.occlusion < 0 ||
        t.occlusion > 1 || t.xcd_bands < 0 || t.xcd_bands > 64 || t.tail_round < 0) {
        setError("set_tuning: value out of range");
        return OMV_E_RANGE;
    }
    tuning_ = t;
    return OMV_OK;
}

OmvSceneView BDPTRenderer::viewForLaunch() const {
    OmvSceneView v = scene_->view();
    v.closest_only = tuning_.occlusion == 1 ? 1 : 0;
    return v;
}

int BDPTRenderer::initStream() {
    hipError_t e;
    if (stream_) return OMV_OK;
    if ((e = hipStreamCreate(&stream_)) != hipSuccess) return hipFail(e, "hipStreamCreate");
    for (auto& ev : ev_) if ((e = hipEventCreate(&ev)) != hipSuccess) return hipFail(e, "hipEventCreate");
    int dev = 0;
    (void)hipGetDevice(&dev);
    if ((e = hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return hipFail(e, "device attribute");
    // XCDs (one L2 each) for the traversal's banded hand-out; 1 where unknown
    if (hipDeviceGetAttribute(&xcds_, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || xcds_ < 1) xcds_ = 1;
    return OMV_OK;
}

int BDPTRenderer::prepare(int which, uint64_t lanesWanted, int& blocks) {
    int rc = initStream();
    if (rc) return rc;
    hipError_t e;
    int perCu = occupancyBlocksPerCU(which);
    if (perCu <= 0) perCu = 1;
    uint64_t maxBlocks = (uint64_t)cus_ * (uint64_t)perCu;
    uint64_t want = (lanesWanted + OMV_BLOCK - 1) / OMV_BLOCK;
    blocks = (int)std::max<uint64_t>(1, std::min(maxBlocks, want));
    const uint64_t lanes = (uint64_t)blocks * OMV_BLOCK;
    if ((e = dStack_.reserve(lanes * (OMV_STACK_TOTAL - OMV_STACK_LDS) * sizeof(uint32_t))) != hipSuccess) return hipFail(e, "stack spill");
    if ((e = dCounters_.reserve(8 * sizeof(unsigned long long))) != hipSuccess) return hipFail(e, "counters");
    return OMV_OK;
}

int BDPTRenderer::render(PinholeCamera& cam, float* filmDev) {
    if (!scene_ || !scene_->committed()) { setError("render: no committed scene"); return OMV_E_STATE; }
    const omv_renderer_params& p = params_;
    if (p.samples_per_pixel < 0 || p.camera_bounces < 0 || p.light_bounces < 0) { setError("render: negative parameter"); return OMV_E_ARG; }
    const uint64_t W = cam.desc.width, H = cam.desc.height, npix = W * H;
    const int spp = p.samples_per_pixel;
    std::memset(&stats_, 0, sizeof(stats_));
    filmOf_ = nullptr;
    if (!filmDev) {
        try {
            cam.film.resize(npix * 4);   // written from the device below
        } catch (const std::bad_alloc&) {
            setError("render: cannot allocate the host film");
            return OMV_E_NOMEM;
        }
    }
    int rc = initStream();
    if (rc) return rc;
    hipError_t e;
    if (!filmDev && (e = dFilm_.reserve(npix * 4 * sizeof(float))) != hipSuccess) return hipFail(e, "film");
    // the film the kernels write: the renderer's (read back into the camera)
    // or the caller's device buffer (left in device memory)
    film_ = filmDev ? filmDev : dFilm_.as<float>();
    if (!hostStage_ && (e = hipHostMalloc(&hostStage_, 8 * sizeof(unsigned long long), hipHostMallocDefault)) != hipSuccess) {
        hostStage_ = nullptr;
        return hipFail(e, "counter staging");
    }
    if ((e = dCounters_.reserve(8 * sizeof(unsigned long long))) != hipSuccess) return hipFail(e, "counters");
    if ((e = hipMemsetAsync(dCounters_.ptr, 0, 8 * sizeof(unsigned long long), stream_)) != hipSuccess) return hipFail(e, "memset");
    // PinholeCamera::init (PinholeCamera.cpp:65-68)
    if ((e = hipMemsetAsync(film_, 0, npix * 4 * sizeof(float), stream_)) != hipSuccess) return hipFail(e, "memset film");
    // OMV_EXT_LIGHT_SPLAT: splat sums over the whole render, added to the film at the end
    splat_ = nullptr;
    if ((p.extensions & OMV_EXT_LIGHT_SPLAT) && scene_->view().num_lights > 0 && spp > 0) {
        if ((e = dSplat_.reserve(npix * 3 * sizeof(float))) != hipSuccess) return hipFail(e, "splat film");
        if ((e = hipMemsetAsync(dSplat_.ptr, 0, npix * 3 * sizeof(float), stream_)) != hipSuccess) return hipFail(e, "memset splat");
        splat_ = dSplat_.as<float>();
    }
    (void)hipEventRecord(ev_[0], stream_);
    double kernelMs = 0.0;
    wfTimingPending_ = false;
    if (spp > 0) {
        rc = schedule_ == OMV_SCHEDULE_WAVEFRONT ? renderWavefront(cam, kernelMs) : renderSamples(cam, kernelMs);
        if (rc) return rc;
        if (splat_ && (e = launchSplatAdd(splat_, npix, (float)(1.0 / ((double)npix * (double)spp)), film_, stream_)) !=
                          hipSuccess)
            return hipFail(e, "splat kernel launch");
    } else {
        // PinholeCamera::stop with a zero count leaves rgb = 0 and alpha = 1
        std::vector<float> z(npix * 4, 0.0f);
        for (uint64_t i = 0; i < npix; ++i) z[i * 4 + 3] = 1.0f;
        if ((e = hipMemcpyAsync(film_, z.data(), z.size() * 4, hipMemcpyHostToDevice, stream_)) != hipSuccess) return hipFail(e, "film");
    }
    (void)hipEventRecord(ev_[1], stream_);
    // film (pinned) and counters on the same stream, one wait
    unsigned long long* hostCnt = static_cast<unsigned long long*>(hostStage_);
    if (!filmDev && (e = hipMemcpyAsync(cam.film.data(), dFilm_.ptr, npix * 4 * sizeof(float), hipMemcpyDeviceToHost, stream_)) != hipSuccess)
        return hipFail(e, "film readback");
    if ((e = hipMemcpyAsync(hostCnt, dCounters_.ptr, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_)) != hipSuccess)
        return hipFail(e, "counter readback");
    (void)hipEventRecord(ev_[4], stream_);
    if ((e = hipEventSynchronize(ev_[4])) != hipSuccess) return hipFail(e, "render");
    float total = 0.0f;
    (void)hipEventElapsedTime(&total, ev_[0], ev_[1]);
    if (wfTimingPending_) collectWavefrontTiming(kernelMs);
    unsigned long long c[8];
    std::memcpy(c, hostCnt, sizeof(c));
    stats_.samples = npix * (uint64_t)spp;
    stats_.rays = c[0]; stats_.camera_rays = c[1]; stats_.light_rays = c[2]; stats_.shadow_rays = c[3];
    stats_.nodes_visited = c[4]; stats_.tris_tested = c[5]; stats_.stack_overflows = c[6];
    stats_.render_ms = total;
    stats_.kernel_ms = kernelMs;
    filmOf_ = filmDev ? nullptr : &cam;
    if (stats_.stack_overflows) { setError("render: traversal stack overflow"); return OMV_E_INTERNAL; }
    return OMV_OK;
}

// FP64 evaluation of the functional backward() differentiates: directions
// from this renderer's scene, weights with eval's parameters (adjoint.hip).
int BDPTRenderer::renderDetached(PinholeCamera& cam, const RenderScene& eval, double* out) {
    if (!scene_ || !scene_->committed()) { setError("render_detached: no committed scene"); return OMV_E_STATE; }
    if (!eval.committed()) { setError("render_detached: evaluation scene not committed"); return OMV_E_STATE; }
    if (!out) { setError("render_detached: null film"); return OMV_E_ARG; }
    const OmvSceneView& S = scene_->view();
    const OmvSceneView& E = eval.view();
    if (S.num_lights > 0) { setError("render_detached: scenes with point lights are not differentiated"); return OMV_E_STATE; }
    if (params_.extensions & (OMV_EXT_HDRI_NEE | OMV_EXT_OM_ROUGH)) { setError("render_detached: estimator extensions are not differentiated"); return OMV_E_STATE; }
    // eval must differ from the scene in material parameters and texels only
    bool same = E.num_meshes == S.num_meshes && E.num_mats == S.num_mats && E.num_texs == S.num_texs;
    for (int i = 0; same && i < S.num_meshes; ++i) {
        int fa = 0, fb = 0, sa = 0, sb = 0, ma = 0, mb = 0;
        same = scene_->meshInfo(i, nullptr, &fa, &sa, &ma) == OMV_OK && eval.meshInfo(i, nullptr, &fb, &sb, &mb) == OMV_OK && fa == fb &&
               sa == sb && ma == mb;
    }
    for (int i = 0; same && i < S.num_texs; ++i) {
        int wa = 0, ha = 0, wb = 0, hb = 0;
        same = scene_->textureSize(i, &wa, &ha) == OMV_OK && eval.textureSize(i, &wb, &hb) == OMV_OK && wa == wb && ha == hb;
    }
    for (int i = 0; same && i < S.num_mats; ++i) same = scene_->isOm(i) == eval.isOm(i);
    if (!same) { setError("render_detached: the evaluation scene's meshes, materials or textures differ in structure"); return OMV_E_ARG; }
    const omv_renderer_params& p = params_;
    const uint64_t npix = (uint64_t)cam.desc.width * cam.desc.height;
    std::memset(&stats_, 0, sizeof(stats_));
    int rc = initStream();
    if (rc) return rc;
    const int blocks = cus_ * detachedBlocksPerCU();
    const uint64_t lanes = (uint64_t)blocks * OMV_BLOCK;
    hipError_t e;
    if ((e = dStack_.reserve(lanes * (OMV_STACK_TOTAL - OMV_STACK_LDS) * sizeof(uint32_t))) != hipSuccess) return hipFail(e, "stack spill");
    if ((e = dCounters_.reserve(8 * sizeof(unsigned long long))) != hipSuccess) return hipFail(e, "counters");
    if ((e = dDetached_.reserve(npix * 4 * sizeof(double))) != hipSuccess) return hipFail(e, "detached film");
    if ((e = hipMemsetAsync(dCounters_.ptr, 0, 8 * sizeof(unsigned long long), stream_)) != hipSuccess) return hipFail(e, "memset");
    (void)hipEventRecord(ev_[0], stream_);
    if ((e = launchDetached(S, E, cam.desc, p, npix, dDetached_.as<double>(), dStack_.as<uint32_t>(), dCounters_.as<unsigned long long>(),
                            blocks, stream_)) != hipSuccess)
        return hipFail(e, "detached kernel launch");
    (void)hipEventRecord(ev_[1], stream_);
    if ((e = hipEventSynchronize(ev_[1])) != hipSuccess) return hipFail(e, "detached kernel");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ev_[0], ev_[1]);
    unsigned long long c[8] = {0};
    if ((e = hipMemcpy(c, dCounters_.ptr, sizeof(c), hipMemcpyDeviceToHost)) != hipSuccess) return hipFail(e, "counter readback");
    if ((e = hipMemcpy(out, dDetached_.ptr, npix * 4 * sizeof(double), hipMemcpyDefault)) != hipSuccess) return hipFail(e, "film readback");
    stats_.samples = npix * (uint64_t)std::max(p.samples_per_pixel, 0);
    stats_.rays = c[0]; stats_.camera_rays = c[1];
    stats_.nodes_visited = c[4]; stats_.tris_tested = c[5]; stats_.stack_overflows = c[6];
    stats_.render_ms = ms;
    stats_.kernel_ms = ms;
    if (stats_.stack_overflows) { setError("render_detached: traversal stack overflow"); return OMV_E_INTERNAL; }
    return OMV_OK;
}

// Allocation-free render: validates on the host, then one persistent
// per-pixel launch over the fixed scratch array (noalloc.hip).  Touches none
// of this renderer's buffers, streams or events.
int BDPTRenderer::renderNoalloc(const PinholeCamera& cam, float* film, hipStream_t stream, omv_render_stats* st) {
    if (!scene_ || !scene_->committed()) { setError("render_noalloc: no committed scene"); return OMV_E_STATE; }
    if (!film) { setError("render_noalloc: null film"); return OMV_E_ARG; }
    const omv_renderer_params& p = params_;
    if (p.samples_per_pixel < 0 || p.camera_bounces < 0 || p.light_bounces < 0) { setError("render_noalloc: negative parameter"); return OMV_E_ARG; }
    const OmvSceneView S = viewForLaunch();
    if ((p.extensions & OMV_EXT_LIGHT_SPLAT) && S.num_lights > 0) {
        setError("render_noalloc: OMV_EXT_LIGHT_SPLAT needs a W*H buffer of splat sums; use omv_renderer_render");
        return OMV_E_STATE;
    }
    unsigned long long c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    hipError_t e = hipSuccess;
    const char* what = nullptr;
    int rc = omv::renderNoalloc(S, cam.desc, p, film, stream, st ? c : nullptr, e, what);
    if (rc) {
        setError(std::string(what ? what : "render_noalloc") + (e != hipSuccess ? std::string(": ") + hipGetErrorString(e) : std::string()));
        return rc;
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->samples = (uint64_t)cam.desc.width * (uint64_t)cam.desc.height * (uint64_t)std::max(p.samples_per_pixel, 0);
        st->rays = c[0]; st->camera_rays = c[1]; st->light_rays = c[2]; st->shadow_rays = c[3];
        st->nodes_visited = c[4]; st->tris_tested = c[5]; st->stack_overflows = c[6];
        if (st->stack_overflows) { setError("render_noalloc: traversal stack overflow"); return OMV_E_INTERNAL; }
    }
    return OMV_OK;
}

int BDPTRenderer::post(PinholeCamera& cam, float gamma, bool denoise, float* outF, uint8_t* outU8) {
    if (filmOf_ != &cam || !dFilm_.ptr) { setError("post: this camera has no film from this renderer"); return OMV_E_STATE; }
    const uint64_t npix = (uint64_t)cam.desc.width * cam.desc.height;
    hipError_t e;
    if ((e = dPost_.reserve(npix * (16 + 16 + 4))) != hipSuccess) return hipFail(e, "post buffers");
    float* tmp = dPost_.as<float>();
    float* f = tmp + npix * 4;
    uint8_t* u8 = reinterpret_cast<uint8_t*>(f + npix * 4);
    if ((e = launchPost(dFilm_.as<float>(), tmp, f, outU8 ? u8 : nullptr, cam.desc.width, cam.desc.height, 1.0f, gamma, denoise, stream_)) !=
        hipSuccess)
        return hipFail(e, "post kernels");
    if ((e = hipStreamSynchronize(stream_)) != hipSuccess) return hipFail(e, "post");
    if (outF && (e = hipMemcpy(outF, f, npix * 16, hipMemcpyDefault)) != hipSuccess) return hipFail(e, "post readback");
    if (outU8 && (e = hipMemcpy(outU8, u8, npix * 4, hipMemcpyDefault)) != hipSuccess) return hipFail(e, "post readback");
    return OMV_OK;
}

// Per-sample schedule: one persistent sample-kernel launch per chunk of
// whole pixels, then the ordered film reduction for that chunk.
int BDPTRenderer::renderSamples(PinholeCamera& cam, double& kernelMs) {
    const omv_renderer_params& p = params_;
    const uint64_t npix = (uint64_t)cam.desc.width * cam.desc.height;
    const int spp = p.samples_per_pixel;
    int blocks = 0;
    const uint64_t chunkPix = std::max<uint64_t>(1, kChunkSamples / (uint64_t)spp);
    const uint64_t firstChunk = std::min(npix, chunkPix) * (uint64_t)spp;
    const bool nee = (p.extensions & OMV_EXT_HDRI_NEE) && scene_->view().env_tex_pdf;
    const bool lights = scene_->view().num_lights > 0;
    int rc = prepare(lights ? (splat_ ? (nee ? 8 : 7) : (nee ? 4 : 0)) : (nee ? 5 : 2), firstChunk, blocks);
    if (rc) return rc;
    const uint64_t lanes = (uint64_t)blocks * OMV_BLOCK;
    const int lvCap = std::max(std::min(p.light_bounces + 1, std::max(p.max_path_length, 1)), 1);
    hipError_t e;
    if ((e = dLv_.reserve(lanes * (uint64_t)lvCap * 6 * sizeof(float))) != hipSuccess) return hipFail(e, "light vertex scratch");
    if ((e = dRad_.reserve(std::min(npix, chunkPix) * (uint64_t)spp * 3 * sizeof(float))) != hipSuccess) return hipFail(e, "sample buffer");
    for (uint64_t pix0 = 0; pix0 < npix; pix0 += chunkPix) {
        const uint64_t np = std::min(chunkPix, npix - pix0);
        (void)hipEventRecord(ev_[2], stream_);
        if ((e = launchBdpt(viewForLaunch(), cam.desc, p, pix0 * (uint64_t)spp, np * (uint64_t)spp, dRad_.as<float>(), dLv_.as<float>(),
                            dStack_.as<uint32_t>(), dCounters_.as<unsigned long long>(), blocks, splat_, stream_)) != hipSuccess)
            return hipFail(e, "sample kernel launch");
        (void)hipEventRecord(ev_[3], stream_);
        if ((e = launchFilm(dRad_.as<float>(), spp, pix0, np, film_, stream_)) != hipSuccess) return hipFail(e, "film kernel launch");
        if ((e = hipEventSynchronize(ev_[3])) != hipSuccess) return hipFail(e, "sample kernel");
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ev_[2], ev_[3]);
        kernelMs += ms;
    }
    return OMV_OK;
}

// Wavefront schedule: chunks of whole pixels, each run as ray-queue rounds
// (wavefront.hip), then the ordered film reduction for that chunk.
int BDPTRenderer::renderWavefront(PinholeCamera& cam, double& kernelMs) {
    const omv_renderer_params& p = params_;
    const OmvSceneView S = viewForLaunch();
    const bool lights = S.num_lights > 0;
    const uint64_t npix = (uint64_t)cam.desc.width * cam.desc.height;
    const int spp = p.samples_per_pixel;
    const int lvCap = std::max(std::min(p.light_bounces + 1, std::max(p.max_path_length, 1)), 1);
    const int lc = std::max(1, std::min(lvCap, std::max(p.max_path_length, 1)));
    const uint64_t maxSlots = tuning_.chunk_samples > 0 ? (uint64_t)tuning_.chunk_samples : (lights ? kWfSlotsLights : kWfSlots);
    const uint64_t chunkPix = std::max<uint64_t>(1, maxSlots / (uint64_t)spp);
    const uint64_t slots = std::min(npix, chunkPix) * (uint64_t)spp;
    if (slots * (uint64_t)lc >= (1ull << 32)) { setError("render: too many samples per pixel for one chunk"); return OMV_E_RANGE; }
    if (!wfPerCu_[0]) wavefrontBlocksPerCU(wfPerCu_);
    // resident traversal workgroups per CU: the occupancy limit of the traversal
    // kernel in use (the persistent form holds more registers)
    const int traceOcc = tuning_.trace_keep > 0 ? wfPerCu_[4] : wfPerCu_[0];
    const int blocksTrace = cus_ * (tuning_.trace_blocks_per_cu > 0 ? std::min(tuning_.trace_blocks_per_cu, traceOcc) : traceOcc);
    hipError_t e;
    if ((e = dStack_.reserve((uint64_t)blocksTrace * OMV_BLOCK * (OMV_STACK_TOTAL - OMV_STACK_LDS) * sizeof(uint32_t))) != hipSuccess)
        return hipFail(e, "stack spill");
    const bool nee = (p.extensions & OMV_EXT_HDRI_NEE) && S.env_tex_pdf;
    const int blocksShade = cus_ * shadeCamBlocksPerCU(wfPerCu_, lights, nee);
    if ((e = dArena_.reserve(wavefrontArenaBytes((uint32_t)slots, lvCap, lc, lights, nee, splat_ != nullptr, tuning_.sort_rays == 2,
                                                 blocksShade, cus_ * wfPerCu_[3]))) != hipSuccess)
        return hipFail(e, "wavefront state");
    if ((e = dQcount_.reserve(wavefrontCounterCount(p.camera_bounces, lvCap) * sizeof(uint32_t))) != hipSuccess) return hipFail(e, "queue counters");
    if ((e = dRad_.reserve(slots * 3 * sizeof(float))) != hipSuccess) return hipFail(e, "sample buffer");
    WfLaunch L;
    std::memset(&L, 0, sizeof(L));
    L.S = S; L.cam = cam.desc; L.p = p; L.lvCap = lvCap; L.lc = lc;
    L.arena = dArena_.ptr; L.qcount = dQcount_.as<uint32_t>(); L.rad = dRad_.as<float>(); L.gstack = dStack_.as<uint32_t>();
    L.counters = dCounters_.as<unsigned long long>();
    L.blocksTrace = blocksTrace;
    L.blocksShade = blocksShade;
    L.blocksGen = cus_ * wfPerCu_[3];
    L.traceKeep = tuning_.trace_keep;
    L.splat = splat_;
    L.sortRays = tuning_.sort_rays == 2;
    // default hand-out (measured on MI355X, entry-by-entry traversal): bands
    // of the queue dealt to the XCDs on scenes of at most 4 instances -- two
    // per XCD when introduced (torus knot 512^2 x 64 kernels 9.50 -> 8.95 ms,
    // icosphere 256^2 x 64 1.64 -> 1.53 ms), four per XCD once the kernels had
    // become faster (2 / 4 / 8 per XCD: torus knot 6.27 / 6.16 / 6.23 ms,
    // icosphere 0.565 / 0.550 / 0.558 ms) -- and the plain stride on
    // many-instance scenes (instanced 1080p 15.85 vs 16.03 ms; later 10.92 vs
    // 11.00 with one band per XCD)
    L.xcdBands = tuning_.xcd_bands > 0 ? tuning_.xcd_bands : (S.num_insts <= 4 ? 4 * xcds_ : 1);
    L.tailRound = tuning_.tail_round == 1 ? 0
                 : (tuning_.tail_round >= 2 ? tuning_.tail_round : (lights || nee ? kTailRoundVisibility : kTailRoundPlain));
    const bool packets = tuning_.traversal == 2;
    L.blocksPacket = packets ? cus_ * wfPerCu_[8] : 0;
    // an event pair around every traversal launch: (2 light + 3 camera rounds) per chunk
    const uint64_t chunks = (npix + chunkPix - 1) / chunkPix;
    const size_t evNeed = (size_t)std::min<uint64_t>(1u << 14, chunks * 2 * (uint64_t)(2 * lvCap + 3 * p.camera_bounces));
    while (traceEv_.size() < evNeed) {
        hipEvent_t ev;
        if ((e = hipEventCreate(&ev)) != hipSuccess) return hipFail(e, "hipEventCreate");
        traceEv_.push_back(ev);
    }
    int evUsed = 0;
    L.traceEv = traceEv_.data(); L.traceEvCap = (int)(evNeed & ~(size_t)1); L.traceEvUsed = &evUsed;
    // Two chunk streams (the default when there are several chunks): chunks
    // alternate between stream_ and stream2_, each with its own state, and
    // chunk k + 1 starts once chunk k has finished chunk_stagger camera rounds.
    // Chunks write disjoint film rows and their own sample buffers, so the
    // film is the same as with one stream.
    const int nstreams = tuning_.chunk_streams != 1 && chunks > 1 ? 2 : 1;
    hipStream_t st[2] = {stream_, stream_};
    WfLaunch Ls[2] = {L, L};
    float* rad[2] = {dRad_.as<float>(), dRad_.as<float>()};
    if (nstreams == 2) {
        if (!stream2_ && (e = hipStreamCreate(&stream2_)) != hipSuccess) return hipFail(e, "hipStreamCreate");
        for (auto& ev : chunkEv_)
            if (!ev && (e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return hipFail(e, "hipEventCreate");
        if ((e = dStack2_.reserve(dStack_.bytes)) != hipSuccess) return hipFail(e, "stack spill");
        if ((e = dArena2_.reserve(dArena_.bytes)) != hipSuccess) return hipFail(e, "wavefront state");
        if ((e = dQcount2_.reserve(dQcount_.bytes)) != hipSuccess) return hipFail(e, "queue counters");
        if ((e = dRad2_.reserve(slots * 3 * sizeof(float))) != hipSuccess) return hipFail(e, "sample buffer");
        st[1] = stream2_;
        Ls[1].arena = dArena2_.ptr; Ls[1].qcount = dQcount2_.as<uint32_t>(); Ls[1].rad = rad[1] = dRad2_.as<float>();
        Ls[1].gstack = dStack2_.as<uint32_t>();
        const int stagger = tuning_.chunk_stagger > 0 ? tuning_.chunk_stagger : 1;
        for (int k = 0; k < 2; ++k) { Ls[k].roundEv = chunkEv_[k]; Ls[k].roundEvAt = stagger; }
    }
    (void)hipEventRecord(ev_[2], stream_);
    uint64_t k = 0;
    for (uint64_t pix0 = 0; pix0 < npix; pix0 += chunkPix, ++k) {
        const int j = (int)(k % (uint64_t)nstreams);
        const uint64_t np = std::min(chunkPix, npix - pix0);
        WfLaunch& Lk = Ls[j];
        Lk.sample0 = pix0 * (uint64_t)spp;
        Lk.nslots = (uint32_t)(np * (uint64_t)spp);
        const char* what = nullptr;
        if (nstreams == 2 && k > 0 && (e = hipStreamWaitEvent(st[j], chunkEv_[1 - j], 0)) != hipSuccess) what = "hipStreamWaitEvent";
        else if ((e = launchWavefront(Lk, st[j])) != hipSuccess) what = "wavefront launch";
        else if ((e = launchFilm(rad[j], spp, pix0, np, film_, st[j])) != hipSuccess) what = "film kernel launch";
        if (what) {
            // both streams drained before the buffers can be used again
            if (nstreams == 2) (void)hipStreamSynchronize(stream2_);
            (void)hipStreamSynchronize(stream_);
            return hipFail(e, what);
        }
    }
    if (nstreams == 2) {
        // stream_ goes on (splat sum, read-back) once stream2_ has finished
        (void)hipEventRecord(chunkEv_[2], stream2_);
        if ((e = hipStreamWaitEvent(stream_, chunkEv_[2], 0)) != hipSuccess) return hipFail(e, "hipStreamWaitEvent");
    }
    (void)hipEventRecord(ev_[3], stream_);
    // kernel and traversal times are read once render() has waited for the stream
    wfEvUsed_ = evUsed;
    wfTimingPending_ = true;
    (void)kernelMs;
    return OMV_OK;
}

// Times of the last wavefront render's kernels (ev_[2] .. ev_[3]) and of its
// traversal launches; the stream has been waited for.
void BDPTRenderer::collectWavefrontTiming(double& kernelMs) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ev_[2], ev_[3]);
    kernelMs += ms;
    double tr = 0.0;
    for (int k = 0; k + 1 < wfEvUsed_; k += 2) {
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, traceEv_[k], traceEv_[k + 1]);
        tr += t;
    }
    stats_.trace_ms = tr;
    stats_.trace_launches = (uint64_t)(wfEvUsed_ / 2);
    wfTimingPending_ = false;
}

// Path-replay adjoint (adjoint.hip) of the last render's estimator.
int BDPTRenderer::backward(PinholeCamera& cam, const float* dFilm, const omv_adjoint_targets& t) {
    if (!scene_ || !scene_->committed()) { setError("backward: no committed scene"); return OMV_E_STATE; }
    const OmvSceneView& S = scene_->view();
    if (S.num_lights > 0) { setError("backward: scenes with point lights are not differentiated"); return OMV_E_STATE; }
    if (params_.extensions & OMV_EXT_HDRI_NEE) { setError("backward: OMV_EXT_HDRI_NEE is not differentiated"); return OMV_E_STATE; }
    if (params_.extensions & OMV_EXT_OM_ROUGH) { setError("backward: the OM rough-conductor extension is not differentiated"); return OMV_E_STATE; }
    if (!dFilm) { setError("backward: null d_film"); return OMV_E_ARG; }
    if (t.n_textures < 0 || t.n_om_materials < 0 || (t.n_textures && (!t.texture_ids || !t.texture_grads)) ||
        (t.n_om_materials && (!t.om_material_ids || !t.om_grads))) {
        setError("backward: bad targets"); return OMV_E_ARG;
    }
    std::vector<int> texSlot(S.num_texs, -1), omSlot(S.num_mats, -1);
    std::vector<float*> texGrad;
    for (int i = 0; i < t.n_textures; ++i) {
        const int id = t.texture_ids[i];
        int w = 0, h = 0;
        if (scene_->textureSize(id, &w, &h) != OMV_OK || w <= 0 || !t.texture_grads[i]) { setError("backward: bad texture target"); return OMV_E_RANGE; }
        if (texSlot[id] >= 0) { setError("backward: texture listed twice"); return OMV_E_ARG; }
        texSlot[id] = (int)texGrad.size();
        texGrad.push_back(t.texture_grads[i]);
    }
    for (int i = 0; i < t.n_om_materials; ++i) {
        const int id = t.om_material_ids[i];
        if (!scene_->isOm(id) || omSlot[id] >= 0) { setError("backward: bad OM material target"); return OMV_E_RANGE; }
        omSlot[id] = i;
    }
    const omv_renderer_params& p = params_;
    const uint64_t npix = (uint64_t)cam.desc.width * cam.desc.height;
    const uint64_t nsamples = npix * (uint64_t)std::max(p.samples_per_pixel, 0);
    std::memset(&stats_, 0, sizeof(stats_));
    int rc = initStream();
    if (rc) return rc;
    if (nsamples == 0) return OMV_OK;
    const int blocks = cus_ * adjointBlocksPerCU();
    const uint64_t lanes = (uint64_t)blocks * OMV_BLOCK;
    hipError_t e;
    if ((e = dStack_.reserve(lanes * (OMV_STACK_TOTAL - OMV_STACK_LDS) * sizeof(uint32_t))) != hipSuccess) return hipFail(e, "stack spill");
    if ((e = dLv_.reserve(lanes * 6 * sizeof(float))) != hipSuccess) return hipFail(e, "lane scratch");
    if ((e = dCounters_.reserve(8 * sizeof(unsigned long long))) != hipSuccess) return hipFail(e, "counters");
    if ((e = dDfilm_.reserve(npix * 4 * sizeof(float))) != hipSuccess) return hipFail(e, "d_film");
    // slot maps and the pointer table in one upload
    std::vector<uint8_t> maps(texSlot.size() * 4 + omSlot.size() * 4 + texGrad.size() * sizeof(float*) + 16);
    size_t oTex = 0, oOm = texSlot.size() * 4, oPtr = ((oOm + omSlot.size() * 4 + 15) / 16) * 16;
    maps.resize(oPtr + std::max<size_t>(1, texGrad.size()) * sizeof(float*));
    std::memcpy(maps.data() + oTex, texSlot.data(), texSlot.size() * 4);
    std::memcpy(maps.data() + oOm, omSlot.data(), omSlot.size() * 4);
    if (!texGrad.empty()) std::memcpy(maps.data() + oPtr, texGrad.data(), texGrad.size() * sizeof(float*));
    if ((e = dAdjMaps_.upload(maps.data(), maps.size())) != hipSuccess) return hipFail(e, "adjoint maps");
    if ((e = hipMemcpyAsync(dDfilm_.ptr, dFilm, npix * 4 * sizeof(float), hipMemcpyDefault, stream_)) != hipSuccess) return hipFail(e, "d_film copy");
    if ((e = hipMemsetAsync(dCounters_.ptr, 0, 8 * sizeof(unsigned long long), stream_)) != hipSuccess) return hipFail(e, "memset");
    uint8_t* mb = dAdjMaps_.as<uint8_t>();
    (void)hipEventRecord(ev_[0], stream_);
    if ((e = launchAdjoint(S, cam.desc, p, 0, nsamples, dDfilm_.as<float>(), reinterpret_cast<const int*>(mb + oTex),
                           reinterpret_cast<float* const*>(mb + oPtr), reinterpret_cast<const int*>(mb + oOm), t.om_grads, dLv_.as<float>(),
                           dStack_.as<uint32_t>(), dCounters_.as<unsigned long long>(), blocks, stream_)) != hipSuccess)
        return hipFail(e, "adjoint kernel launch");
    (void)hipEventRecord(ev_[1], stream_);
    if ((e = hipEventSynchronize(ev_[1])) != hipSuccess) return hipFail(e, "adjoint kernel");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ev_[0], ev_[1]);
    unsigned long long c[8] = {0};
    if ((e = hipMemcpy(c, dCounters_.ptr, sizeof(c), hipMemcpyDeviceToHost)) != hipSuccess) return hipFail(e, "counter readback");
    stats_.samples = nsamples;
    stats_.rays = c[0]; stats_.camera_rays = c[1];
    stats_.nodes_visited = c[4]; stats_.tris_tested = c[5]; stats_.stack_overflows = c[6];
    stats_.render_ms = ms;
    stats_.kernel_ms = ms;
    if (stats_.stack_overflows) { setError("backward: traversal stack overflow"); return OMV_E_INTERNAL; }
    return OMV_OK;
}

int BDPTRenderer::renderFc(PinholeCamera& cam, int32_t* ids, uint8_t* tie) {
    if (!scene_ || !scene_->committed()) { setError("render_fc: no committed scene"); return OMV_E_STATE; }
    const uint64_t npix = (uint64_t)cam.desc.width * cam.desc.height;
    std::memset(&stats_, 0, sizeof(stats_));
    try {
        cam.film.assign(npix * 4, 0.0f);
    } catch (const std::bad_alloc&) {
        setError("render_fc: cannot allocate the host film");
        return OMV_E_NOMEM;
    }
    int blocks = 0;
    int rc = prepare(1, npix, blocks);
    if (rc) return rc;
    hipError_t e;
    if ((e = dFilm_.reserve(npix * 4 * sizeof(float))) != hipSuccess) return hipFail(e, "film");
    if ((e = dIds_.reserve(npix * 4 * sizeof(int32_t))) != hipSuccess) return hipFail(e, "ids");
    if ((e = dTie_.reserve(npix)) != hipSuccess) return hipFail(e, "tie");
    // colour table of an enumerated subject (and the instance-face bases of ipid)
    OmvSceneView view = scene_->view();
    const int sub = scene_->fcSubject();
    std::vector<uint32_t> instBase;
    const uint64_t catSize = scene_->fcCategory(sub, instBase);
    if (sub == OMV_FC_PID || sub == OMV_FC_GID || sub == OMV_FC_MID || sub == OMV_FC_IID || sub == OMV_FC_IPID) {
        if (catSize > 0xFFFFFFu) { setError("render_fc: more ids than there are 24-bit colours"); return OMV_E_RANGE; }
        if (sub != fcTableSubject_ || params_.seed != fcTableSeed_ || catSize != fcTableSize_) {
            std::vector<uint32_t> table;
            fcColorTable(params_.seed, sub, (uint32_t)catSize, table);
            if (table.empty()) table.push_back(0);
            if ((e = dFcTable_.upload(table.data(), table.size() * sizeof(uint32_t))) != hipSuccess) return hipFail(e, "fc table");
            fcTableSubject_ = sub; fcTableSeed_ = params_.seed; fcTableSize_ = catSize;
        }
        view.fc_table = dFcTable_.as<uint32_t>();
        view.fc_table_size = (uint32_t)catSize;
        if (sub == OMV_FC_IPID) {
            if (instBase.empty()) instBase.push_back(0);
            if ((e = dFcInstBase_.upload(instBase.data(), instBase.size() * sizeof(uint32_t))) != hipSuccess) return hipFail(e, "fc instance bases");
            view.fc_inst_base = dFcInstBase_.as<uint32_t>();
        }
    }
    if ((e = hipMemsetAsync(dCounters_.ptr, 0, 8 * sizeof(unsigned long long), stream_)) != hipSuccess) return hipFail(e, "memset");
    (void)hipEventRecord(ev_[0], stream_);
    if ((e = launchFc(view, cam.desc, params_.seed, dFilm_.as<float>(), dIds_.as<int32_t>(), dTie_.as<uint8_t>(),
                      dStack_.as<uint32_t>(), dCounters_.as<unsigned long long>(), blocks, stream_)) != hipSuccess)
        return hipFail(e, "fc kernel launch");
    (void)hipEventRecord(ev_[1], stream_);
    if ((e = hipEventSynchronize(ev_[1])) != hipSuccess) return hipFail(e, "fc kernel");
    float total = 0.0f;
    (void)hipEventElapsedTime(&total, ev_[0], ev_[1]);
    if ((e = hipMemcpy(cam.film.data(), dFilm_.ptr, npix * 4 * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess) return hipFail(e, "film readback");
    if (ids && (e = hipMemcpy(ids, dIds_.ptr, npix * 4 * sizeof(int32_t), hipMemcpyDeviceToHost)) != hipSuccess) return hipFail(e, "ids readback");
    if (tie && (e = hipMemcpy(tie, dTie_.ptr, npix, hipMemcpyDeviceToHost)) != hipSuccess) return hipFail(e, "tie readback");
    unsigned long long c[8] = {0};
    if ((e = hipMemcpy(c, dCounters_.ptr, sizeof(c), hipMemcpyDeviceToHost)) != hipSuccess) return hipFail(e, "counter readback");
    stats_.samples = npix;
    stats_.rays = c[0]; stats_.camera_rays = c[1];
    stats_.nodes_visited = c[4]; stats_.tris_tested = c[5]; stats_.stack_overflows = c[6];
    stats_.render_ms = total;
    stats_.kernel_ms = total;
    if (stats_.stack_overflows) { setError("render_fc: traversal stack overflow"); return OMV_E_INTERNAL; }
    return OMV_OK;
}

}  // namespace omv
