// render_core.h -- C++ host side of the render core, mirroring the three
// reference classes the C-ABI replaces:
//   omv::RenderScene    pathtracer/src/scene/RenderScene.h:109-204
//   omv::PinholeCamera  pathtracer/src/sensors/PinholeCamera.h
//   omv::BDPTRenderer   pathtracer/src/renderer/BDPTRenderer.h:147-155
// The scene owns every device buffer; the renderer owns per-launch scratch.
#ifndef OMV_RENDER_CORE_H
#define OMV_RENDER_CORE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <deque>
#include <new>
#include <string>
#include <vector>

#include "../../include/omv.h"
#include "omv_types.h"

namespace omv {

// Owns one device allocation.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release();
    hipError_t reserve(size_t n);                        // grows, never shrinks
    hipError_t upload(const void* src, size_t n);        // exact size
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct HostMesh {
    int id = -1, vn = 0, fn = 0, texChN = 0, stride = 0, filledTexChN = 0, matId = -1;
    std::vector<uint32_t> F;
    std::vector<float> V;
    int texMap[OMV_MAX_UV_CHANNELS];
    int texOffs[OMV_MAX_UV_CHANNELS];
    bool texDef[OMV_MAX_UV_CHANNELS];
    bool vertexDef = false, normalDef = false, tangentDef = false, valid = false;
    HostMesh();
};

struct HostInstance {
    int id = -1, meshId = -1;
    float tm[16];
    float nrm[9];
    float inv[12];
};

struct HostTexture {
    int w = 0, h = 0, fmt = OMV_TEX_RGBA8;
    std::vector<uint8_t> bytes;
};

class RenderScene {
public:
    int allocate(int meshes, int instances, int materials, int textures, int lights);
    int setMesh(int meshId, int vertN, int facesN, int texN, int matId, const uint32_t* faces, const float* vert, int vstride,
                const float* norm, int nstride, const float* tans, int tanstride, const float* const* tex, int texstride);
    int setInstance(int instId, const float trans[16], int meshId);
    int setMaterialOm(int matId, float temperature, int nTables, const omv_ior_table* tables);
    int setMaterialOmNk(int matId, const float n[3], const float k[3]);
    int setMaterialPbr(int matId, const omv_pbr_desc& d);
    int setTexture(int texId, int w, int h, const uint8_t* rgba);
    int setTextureF32(int texId, int w, int h, const float* rgba);
    int setLightPoint(int lightId, const float pos[3], const float intensity[3], float range);
    int setBackgroundHdr(int w, int h, const float* rgb, const float rot[9], float scale);
    int setBackgroundConstant(const float rgb[3]);
    int setFalseColorSubject(const char* name);
    int setBuildMode(int mode);
    int commit(std::string& log);

    bool committed() const { return committed_; }
    const OmvSceneView& view() const { return view_; }
    int meshInfo(int id, int* vn, int* fn, int* stride, int* mat) const;
    int vertices(int id, float* out, size_t cap) const;
    int bvhInfo(int* nodes, int* depth, int* records, size_t* bytes) const;
    int stackBound() const { return stackBound_; }
    int updateTexture(int texId, const void* src);
    int updateMaterialOmNk(int matId, const float n[3], const float k[3]);
    int textureSize(int texId, int* w, int* h) const;
    int envTables(float* rowProb, int* rowAlias, float* colProb, int* colAlias, float* texPdf) const;
    bool isOm(int matId) const { return matId >= 0 && matId < (int)mats_.size() && mats_[matId].type == OMV_MAT_OM; }
    int fcSubject() const { return fcSubject_; }
    // category size of an enumerated false-colour subject (0 for the others)
    // and, for ipid, the first instance-face index of every instance
    uint64_t fcCategory(int subject, std::vector<uint32_t>& instBase) const;

private:
    std::vector<HostMesh> meshes_;
    std::vector<HostInstance> insts_;
    std::vector<OmvMaterial> mats_;   // materials + 1 (missing slot last)
    std::vector<HostTexture> texs_;
    std::vector<OmvLight> lights_;
    std::vector<bool> lightSet_;
    int bgType_ = OMV_BG_NONE, bgW_ = 0, bgH_ = 0;
    std::vector<float> bgRgb_;
    float bgRot_[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float bgScale_ = 1.0f, bgConst_[3] = {100.0f, 100.0f, 100.0f};
    int fcSubject_ = OMV_FC_NONE;
    int buildMode_ = OMV_BUILD_HOST_SAH;
    std::vector<uint32_t> devVtxOffset_;   // device build: meshes whose normals / tangents live only on the device
    std::vector<char> devPrepared_;
    bool committed_ = false;
    int nodeCount_ = 0, bvhDepth_ = 0, recordCount_ = 0, stackBound_ = 0;
    size_t deviceBytes_ = 0;
    std::vector<uint64_t> texOffsets_;   // byte offsets in the device texel pool
    OmvSceneView view_{};
    DeviceBuffer dVtx_, dFaces_, dMeshes_, dInsts_, dMats_, dTexs_, dTexels_, dNodes_, dTris_, dLights_, dCdf_, dBg_, dEnv_, dEnvScratch_;
};

// Page-locked host memory (hipHostMalloc) for films, so a film is read back
// from the device in one DMA without a staging copy.
template <class T>
struct PinnedAllocator {
    using value_type = T;
    PinnedAllocator() = default;
    template <class U>
    PinnedAllocator(const PinnedAllocator<U>&) {}
    T* allocate(size_t n) {
        void* p = nullptr;
        if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess || !p) throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) { (void)hipHostFree(p); }
    template <class U>
    bool operator==(const PinnedAllocator<U>&) const { return true; }
    template <class U>
    bool operator!=(const PinnedAllocator<U>&) const { return false; }
};

class PinholeCamera {
public:
    PinholeCamera(int w, int h);
    void setViewPoint(const float pos[3], const float x[3], const float y[3], const float z[3]);
    void adjustToResolution();
    int setYFoV(float fov);
    void setRealHeight(float height, float aspect);
    void loadProperties(const float* aspect, const float* focus, const float* yfov);
    omv_camera_desc desc;
    std::vector<float, PinnedAllocator<float>> film;   // W*H*4 after a render
};

class BDPTRenderer {
public:
    BDPTRenderer();
    ~BDPTRenderer();
    void setParameters(const omv_renderer_params& p) { params_ = p; }
    void setScene(RenderScene* s) { scene_ = s; }
    int setSchedule(int s);
    int schedule() const { return schedule_; }
    int setTuning(const omv_render_tuning& t);
    const omv_render_tuning& tuning() const { return tuning_; }
    int render(PinholeCamera& cam, float* filmDev = nullptr);
    int renderFc(PinholeCamera& cam, int32_t* ids, uint8_t* tie);
    int renderNoalloc(const PinholeCamera& cam, float* film, hipStream_t stream, omv_render_stats* st);
    int renderDetached(PinholeCamera& cam, const RenderScene& eval, double* out);
    int backward(PinholeCamera& cam, const float* dFilm, const omv_adjoint_targets& t);
    int post(PinholeCamera& cam, float gamma, bool denoise, float* outF, uint8_t* outU8);
    const omv_render_stats& stats() const { return stats_; }

private:
    int prepare(int which, uint64_t lanesWanted, int& blocks);
    int initStream();
    int renderSamples(PinholeCamera& cam, double& kernelMs);
    int renderWavefront(PinholeCamera& cam, double& kernelMs);
    void collectWavefrontTiming(double& kernelMs);
    int wfEvUsed_ = 0;
    bool wfTimingPending_ = false;
    omv_renderer_params params_;
    int schedule_ = OMV_SCHEDULE_WAVEFRONT;
    omv_render_tuning tuning_{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    OmvSceneView viewForLaunch() const;   // the scene view with the tuning's traversal order
    int wfPerCu_[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    RenderScene* scene_ = nullptr;
    omv_render_stats stats_{};
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // pinned host staging for the counter read-back (hipHostMalloc)
    void* hostStage_ = nullptr;
    int cus_ = 0;
    int xcds_ = 1;
    DeviceBuffer dRad_, dLv_, dStack_, dCounters_, dFilm_, dIds_, dTie_, dArena_, dQcount_, dDfilm_, dAdjMaps_, dPost_, dSplat_;
    DeviceBuffer dFcTable_, dFcInstBase_, dDetached_;
    int fcTableSubject_ = -1;          // the table in dFcTable_ is for (subject, seed, size)
    uint32_t fcTableSeed_ = 0;
    uint64_t fcTableSize_ = 0;
    float* splat_ = nullptr;   // OMV_EXT_LIGHT_SPLAT sums of the current render, or null
    float* film_ = nullptr;    // the film the current render's kernels write
    const PinholeCamera* filmOf_ = nullptr;   // camera whose film dFilm_ holds
    std::vector<hipEvent_t> traceEv_;
    // the second chunk stream (chunk_streams 2): its stream, its own wavefront
    // state, queue counters, sample buffer and spill stacks, and one event per
    // stream marking where the next chunk may start
    hipStream_t stream2_ = nullptr;
    hipEvent_t chunkEv_[3] = {nullptr, nullptr, nullptr};
    DeviceBuffer dArena2_, dQcount2_, dRad2_, dStack2_;
};

// Many BDPTRenderer::render jobs in one persistent launch, replayed from a
// captured hipGraph.  Work is handed out in tiles of 64 * tileWaves pixels;
// jobs are laid out heaviest-per-tile first so the cheap tiles fill the tail.
class BatchRunner {
public:
    BatchRunner();
    ~BatchRunner();
    int addJob(const RenderScene* s, const omv_camera_desc& cam, const omv_renderer_params& p);
    int jobCount() const { return (int)jobs_.size(); }
    int setTuning(const omv_batch_tuning& t);
    const omv_batch_tuning& tuning() const { return tuning_; }
    int autotune(int reps, omv_batch_tuning* best);
    int run();
    int film(int job, float* out, size_t cap) const;
    const omv_render_stats& stats() const { return stats_; }

private:
    struct Job {
        const RenderScene* scene;
        omv_camera_desc cam;
        omv_renderer_params p;
        uint64_t npix, filmOff;
        double tileCost;
    };
    int prepare();
    int prepareWavefront(const std::vector<int>& order);
    int launch();
    int launchWavefrontJobs();
    void dropGraph();
    // side streams in use: the tuning's count, at most one per job
    int activeStreams() const { return std::max(1, std::min(tuning_.streams, (int)jobs_.size())); }
    std::vector<double> tunedMs_;   // the last autotune's measured ms per candidate
    std::vector<Job> jobs_;
    bool anyNee_ = false;   // some job uses OMV_EXT_HDRI_NEE: the batch kernel with its code
    std::vector<int> order_;            // heaviest job first
    omv_batch_tuning tuning_{0, 1, 1, 0};
    struct StreamRes {
        hipStream_t s = nullptr;
        hipEvent_t done = nullptr;
        DeviceBuffer arena, qcount, rad, gstack;
    };
    std::deque<StreamRes> sres_;
    hipEvent_t fork_ = nullptr;
    int wfPerCu_[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t wfMaxSlots_ = 0;
    bool dirty_ = true;
    int blocks_ = 0, cus_ = 0;
    uint64_t totalTiles_ = 0;
    omv_render_stats stats_{};
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    std::vector<OmvSceneView> views_;   // scene views the job table was built from
    DeviceBuffer dJobs_, dTileEnd_, dFilm_, dLv_, dStack_, dCounters_;
};

// thread-local last error
void setError(const std::string& msg);
const char* lastError();
// image_io.cpp: saveImage of the reference main (main.cpp:86-120)
int encodeImage(const std::string& ext, const float* rgba, int w, int h, std::vector<uint8_t>& out);
int saveImage(const char* filename, const float* rgba, int w, int h);

}  // namespace omv
#endif
