// batch.cpp -- omv::BatchRunner: many independent render jobs (each what
// BDPTRenderer::render, BDPTRenderer.cpp:372-393, does for one camera) in one
// persistent launch.  The reference renders one job per process run
// (main.cpp:489-521); nothing there batches, so only the per-job result is
// a parity target: every job's film equals BDPTRenderer::render's.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include "omv_kernels.h"
#include "render_core.h"

namespace omv {

namespace {
int hipFail(hipError_t e, const char* what) {
    setError(std::string("batch: ") + what + ": " + hipGetErrorString(e));
    return OMV_E_DEVICE;
}
int lvCapOf(const omv_renderer_params& p) {
    return std::max(std::min(p.light_bounces + 1, std::max(p.max_path_length, 1)), 1);
}
}  // namespace

BatchRunner::BatchRunner() = default;

BatchRunner::~BatchRunner() {
    dropGraph();
    for (auto& e : ev_) if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void BatchRunner::dropGraph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
}

int BatchRunner::addJob(const RenderScene* s, const omv_camera_desc& cam, const omv_renderer_params& p) {
    if (!s || !s->committed()) { setError("batch: job scene is not committed"); return OMV_E_STATE; }
    if (p.samples_per_pixel < 0 || p.camera_bounces < 0 || p.light_bounces < 0) { setError("batch: negative parameter"); return OMV_E_ARG; }
    if (cam.width <= 0 || cam.height <= 0) { setError("batch: empty camera"); return OMV_E_ARG; }
    Job j;
    j.scene = s;
    j.cam = cam;
    j.p = p;
    j.npix = (uint64_t)cam.width * (uint64_t)cam.height;
    j.filmOff = 0;
    // relative cost of one pixel: samples x bounces x traversal depth
    int tris = 0;
    s->bvhInfo(nullptr, nullptr, &tris, nullptr);
    const double bounces = (double)(std::max(p.camera_bounces, 1) + (s->view().num_lights > 0 ? std::max(p.light_bounces, 1) : 0));
    j.tileCost = (double)std::max(p.samples_per_pixel, 1) * bounces * std::log2(2.0 + (double)tris);
    jobs_.push_back(j);
    dirty_ = true;
    return OMV_OK;
}

int BatchRunner::setTuning(const omv_batch_tuning& t) {
    if (t.blocks_per_cu < 0 || t.tile_waves < 1 || t.tile_waves > 64) { setError("batch: tuning out of range"); return OMV_E_RANGE; }
    tuning_ = t;
    tuning_.use_graph = t.use_graph ? 1 : 0;
    dirty_ = true;
    return OMV_OK;
}

int BatchRunner::prepare() {
    hipError_t e;
    if (jobs_.empty()) { setError("batch: no jobs"); return OMV_E_STATE; }
    if (!stream_) {
        if ((e = hipStreamCreate(&stream_)) != hipSuccess) return hipFail(e, "hipStreamCreate");
        for (auto& ev : ev_) if ((e = hipEventCreate(&ev)) != hipSuccess) return hipFail(e, "hipEventCreate");
        int dev = 0;
        (void)hipGetDevice(&dev);
        if ((e = hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return hipFail(e, "device attribute");
    }
    dropGraph();
    for (const Job& j : jobs_)
        if (!j.scene->committed()) { setError("batch: a job scene is no longer committed"); return OMV_E_STATE; }

    // film offsets in job-id order
    uint64_t filmFloats = 0;
    for (Job& j : jobs_) { j.filmOff = filmFloats; filmFloats += j.npix * 4; }
    // tile order: heaviest tiles first (longest-processing-time order)
    std::vector<int> order(jobs_.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return jobs_[a].tileCost > jobs_[b].tileCost; });
    const uint64_t tilePix = 64ull * (uint64_t)tuning_.tile_waves;
    std::vector<uint64_t> tileEnd(jobs_.size());
    uint64_t t = 0;
    for (size_t i = 0; i < order.size(); ++i) {
        t += (jobs_[order[i]].npix + tilePix - 1) / tilePix;
        tileEnd[i] = t;
    }
    totalTiles_ = t;

    int occ = occupancyBlocksPerCU(3);
    if (occ <= 0) occ = 1;
    const int perCu = tuning_.blocks_per_cu > 0 ? std::min(tuning_.blocks_per_cu, occ) : occ;
    blocks_ = std::max(1, cus_ * perCu);
    const uint64_t lanes = (uint64_t)blocks_ * OMV_BLOCK;
    int lvCap = 1;
    for (const Job& j : jobs_) lvCap = std::max(lvCap, lvCapOf(j.p));

    if ((e = dFilm_.reserve(filmFloats * sizeof(float))) != hipSuccess) return hipFail(e, "film");
    if ((e = dLv_.reserve(lanes * (uint64_t)lvCap * 6 * sizeof(float))) != hipSuccess) return hipFail(e, "light vertex scratch");
    if ((e = dStack_.reserve(lanes * (OMV_STACK_TOTAL - OMV_STACK_LDS) * sizeof(uint32_t))) != hipSuccess) return hipFail(e, "stack spill");
    if ((e = dCounters_.reserve(16 * sizeof(unsigned long long))) != hipSuccess) return hipFail(e, "counters");

    const size_t rec = batchJobBytes();
    std::vector<uint8_t> table(rec * jobs_.size());
    views_.resize(jobs_.size());
    for (size_t i = 0; i < order.size(); ++i) {
        const Job& j = jobs_[order[i]];
        views_[order[i]] = j.scene->view();
        packBatchJob(table.data() + i * rec, j.scene->view(), j.cam, j.p, dFilm_.as<float>() + j.filmOff, j.npix);
    }
    if ((e = dJobs_.upload(table.data(), table.size())) != hipSuccess) return hipFail(e, "job table");
    if ((e = dTileEnd_.upload(tileEnd.data(), tileEnd.size() * sizeof(uint64_t))) != hipSuccess) return hipFail(e, "tile table");

    if (tuning_.use_graph) {
        if ((e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal)) != hipSuccess) return hipFail(e, "begin capture");
        hipError_t le = launch() == OMV_OK ? hipSuccess : hipErrorLaunchFailure;
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(stream_, &g);
        if (le != hipSuccess || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hipFail(e != hipSuccess ? e : le, "graph capture");
        }
        graph_ = g;
        if ((e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0)) != hipSuccess) { dropGraph(); return hipFail(e, "graph instantiate"); }
    }
    dirty_ = false;
    return OMV_OK;
}

int BatchRunner::launch() {
    hipError_t e = launchBatch(dJobs_.ptr, (int)jobs_.size(), dTileEnd_.as<uint64_t>(), totalTiles_, tuning_.tile_waves,
                               dCounters_.as<unsigned long long>(), dLv_.as<float>(), dStack_.as<uint32_t>(), blocks_, stream_);
    if (e != hipSuccess) return hipFail(e, "batch kernel launch");
    return OMV_OK;
}

int BatchRunner::run() {
    // a re-committed scene moves its device buffers: rebuild the job table
    if (!dirty_)
        for (size_t i = 0; i < jobs_.size(); ++i) {
            OmvSceneView v = jobs_[i].scene->view();
            if (std::memcmp(&v, &views_[i], sizeof(v)) != 0) { dirty_ = true; break; }
        }
    if (dirty_) {
        int rc = prepare();
        if (rc) return rc;
    }
    std::memset(&stats_, 0, sizeof(stats_));
    hipError_t e;
    (void)hipEventRecord(ev_[0], stream_);
    if (exec_) {
        if ((e = hipGraphLaunch(exec_, stream_)) != hipSuccess) return hipFail(e, "graph launch");
    } else {
        int rc = launch();
        if (rc) return rc;
    }
    (void)hipEventRecord(ev_[1], stream_);
    if ((e = hipEventSynchronize(ev_[1])) != hipSuccess) return hipFail(e, "batch kernel");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ev_[0], ev_[1]);
    unsigned long long c[8] = {0};
    if ((e = hipMemcpy(c, dCounters_.ptr, sizeof(c), hipMemcpyDeviceToHost)) != hipSuccess) return hipFail(e, "counter readback");
    for (const Job& j : jobs_) stats_.samples += j.npix * (uint64_t)j.p.samples_per_pixel;
    stats_.rays = c[0]; stats_.camera_rays = c[1]; stats_.light_rays = c[2]; stats_.shadow_rays = c[3];
    stats_.nodes_visited = c[4]; stats_.tris_tested = c[5]; stats_.stack_overflows = c[6];
    stats_.render_ms = ms;
    stats_.kernel_ms = ms;
    if (stats_.stack_overflows) { setError("batch: traversal stack overflow"); return OMV_E_INTERNAL; }
    return OMV_OK;
}

int BatchRunner::autotune(int reps, omv_batch_tuning* best) {
    if (jobs_.empty()) { setError("batch: no jobs"); return OMV_E_STATE; }
    reps = std::max(1, reps);
    const int occ = std::max(1, occupancyBlocksPerCU(3));
    const omv_batch_tuning start = tuning_;
    omv_batch_tuning bestT = start;
    double bestMs = 1e300;
    const int tiles[] = {1, 2, 4};
    for (int bpc = 1; bpc <= occ; ++bpc)
        for (int tw : tiles) {
            omv_batch_tuning t{bpc, tw, start.use_graph};
            int rc = setTuning(t);
            if (rc) return rc;
            double m = 1e300;
            for (int r = 0; r < reps; ++r) {
                rc = run();
                if (rc) return rc;
                m = std::min(m, stats_.render_ms);
            }
            if (m < bestMs) { bestMs = m; bestT = t; }
        }
    int rc = setTuning(bestT);
    if (rc) return rc;
    if (best) *best = bestT;
    return OMV_OK;
}

int BatchRunner::film(int job, float* out, size_t cap) const {
    if (job < 0 || job >= (int)jobs_.size()) { setError("batch: job id out of range"); return OMV_E_RANGE; }
    const Job& j = jobs_[job];
    if (!dFilm_.ptr || dirty_) { setError("batch: not run"); return OMV_E_STATE; }
    if (cap < j.npix * 4) { setError("batch: film buffer too small"); return OMV_E_RANGE; }
    hipError_t e = hipMemcpy(out, dFilm_.as<float>() + j.filmOff, j.npix * 4 * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipFail(e, "film readback");
    return OMV_OK;
}

}  // namespace omv
