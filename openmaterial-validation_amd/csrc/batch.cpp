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
    for (auto& r : sres_) {
        if (r.done) (void)hipEventDestroy(r.done);
        if (r.s) (void)hipStreamDestroy(r.s);
    }
    if (fork_) (void)hipEventDestroy(fork_);
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
    if ((p.extensions & OMV_EXT_LIGHT_SPLAT) && s->view().num_lights > 0) {
        setError("batch: OMV_EXT_LIGHT_SPLAT jobs are rendered with omv_renderer_render");
        return OMV_E_ARG;
    }
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
    if (t.blocks_per_cu < 0 || t.tile_waves < 1 || t.tile_waves > 64 || t.streams < 0 || t.streams > 16) {
        setError("batch: tuning out of range");
        return OMV_E_RANGE;
    }
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
    order_ = order;
    if (tuning_.streams > 0) {
        if ((e = dFilm_.reserve(filmFloats * sizeof(float))) != hipSuccess) return hipFail(e, "film");
        if ((e = dCounters_.reserve(16 * sizeof(unsigned long long))) != hipSuccess) return hipFail(e, "counters");
        views_.resize(jobs_.size());
        for (size_t i = 0; i < jobs_.size(); ++i) views_[i] = jobs_[i].scene->view();
        int rc = prepareWavefront(order);
        if (rc) return rc;
        dirty_ = false;
        return OMV_OK;
    }
    const uint64_t tilePix = 64ull * (uint64_t)tuning_.tile_waves;
    std::vector<uint64_t> tileEnd(jobs_.size());
    uint64_t t = 0;
    for (size_t i = 0; i < order.size(); ++i) {
        t += (jobs_[order[i]].npix + tilePix - 1) / tilePix;
        tileEnd[i] = t;
    }
    totalTiles_ = t;

    anyNee_ = false;
    for (const Job& j : jobs_) anyNee_ = anyNee_ || ((j.p.extensions & OMV_EXT_HDRI_NEE) && j.scene->view().env_tex_pdf);
    int occ = occupancyBlocksPerCU(anyNee_ ? 6 : 3);
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

// Wavefront jobs: per-stream state sized for the largest job chunk, the
// whole sequence optionally captured into one graph (fork/join over streams).
int BatchRunner::prepareWavefront(const std::vector<int>& order) {
    hipError_t e;
    const int ns = activeStreams();
    while ((int)sres_.size() < ns) {
        sres_.emplace_back();
        StreamRes& r = sres_.back();
        if ((e = hipStreamCreate(&r.s)) != hipSuccess) return hipFail(e, "hipStreamCreate");
        if ((e = hipEventCreateWithFlags(&r.done, hipEventDisableTiming)) != hipSuccess) return hipFail(e, "hipEventCreate");
    }
    // Events that a capture records belong to that capture's graph: every
    // capture gets its own fork / join events, so no event recorded into a
    // graph that has since been destroyed is recorded or waited on again.
    if (tuning_.use_graph) {
        if (fork_) (void)hipEventDestroy(fork_);
        fork_ = nullptr;
        for (int k = 0; k < ns; ++k) {
            if (sres_[k].done) (void)hipEventDestroy(sres_[k].done);
            sres_[k].done = nullptr;
            if ((e = hipEventCreateWithFlags(&sres_[k].done, hipEventDisableTiming)) != hipSuccess) return hipFail(e, "hipEventCreate");
        }
    }
    if (!fork_ && (e = hipEventCreateWithFlags(&fork_, hipEventDisableTiming)) != hipSuccess) return hipFail(e, "hipEventCreate");
    if (!wfPerCu_[0]) wavefrontBlocksPerCU(wfPerCu_);
    wfMaxSlots_ = 1ull << 22;
    const int blocksTrace = cus_ * wfPerCu_[0];
    size_t arena = 0, rad = 0;
    int ncnt = 0;
    for (int j : order) {
        const Job& J = jobs_[j];
        const int spp = std::max(J.p.samples_per_pixel, 0);
        if (spp == 0) continue;
        const bool lights = J.scene->view().num_lights > 0;
        const int lvCap = lvCapOf(J.p);
        const int lc = std::max(1, std::min(lvCap, std::max(J.p.max_path_length, 1)));
        const uint64_t chunkPix = std::max<uint64_t>(1, wfMaxSlots_ / (uint64_t)spp);
        const uint64_t slots = std::min(J.npix, chunkPix) * (uint64_t)spp;
        if (slots * (uint64_t)lc >= (1ull << 32)) { setError("batch: too many samples per pixel for one chunk"); return OMV_E_RANGE; }
        const bool nee = (J.p.extensions & OMV_EXT_HDRI_NEE) && J.scene->view().env_tex_pdf;
        arena = std::max(arena, wavefrontArenaBytes((uint32_t)slots, lvCap, lc, lights, nee, false, false,
                                                    cus_ * shadeCamBlocksPerCU(wfPerCu_, lights, nee), cus_ * wfPerCu_[3]));
        rad = std::max(rad, (size_t)slots * 3 * sizeof(float));
        ncnt = std::max(ncnt, wavefrontCounterCount(J.p.camera_bounces, lvCap));
    }
    for (int k = 0; k < ns; ++k) {
        StreamRes& r = sres_[k];
        if ((e = r.arena.reserve(std::max<size_t>(arena, 256))) != hipSuccess) return hipFail(e, "wavefront state");
        if ((e = r.rad.reserve(std::max<size_t>(rad, 16))) != hipSuccess) return hipFail(e, "sample buffer");
        if ((e = r.qcount.reserve(std::max(ncnt, 1) * sizeof(uint32_t))) != hipSuccess) return hipFail(e, "queue counters");
        if ((e = r.gstack.reserve((size_t)blocksTrace * OMV_BLOCK * (OMV_STACK_TOTAL - OMV_STACK_LDS) * sizeof(uint32_t))) != hipSuccess)
            return hipFail(e, "stack spill");
    }
    if (tuning_.use_graph) {
        if ((e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal)) != hipSuccess) return hipFail(e, "begin capture");
        int rc = launchWavefrontJobs();
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(stream_, &g);
        if (rc != OMV_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc != OMV_OK ? rc : hipFail(e, "graph capture");
        }
        graph_ = g;
        if ((e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0)) != hipSuccess) { dropGraph(); return hipFail(e, "graph instantiate"); }
    }
    return OMV_OK;
}

// The job sequence over activeStreams() streams: fork from stream_, each
// job's launches on its stream, join back into stream_.  A failed launch
// stops issuing work but the side streams are still joined (so a capture
// can end and a plain launch leaves no stream running ahead); the first
// error is returned.
int BatchRunner::launchWavefrontJobs() {
    hipError_t e;
    const int ns = activeStreams();
    if ((e = launchFillU64(dCounters_.as<unsigned long long>(), 16, stream_)) != hipSuccess) return hipFail(e, "counter fill");
    if ((e = hipEventRecord(fork_, stream_)) != hipSuccess) return hipFail(e, "fork");
    int forked = 0;
    int rc = OMV_OK;
    for (; forked < ns; ++forked)
        if ((e = hipStreamWaitEvent(sres_[forked].s, fork_, 0)) != hipSuccess) { rc = hipFail(e, "fork wait"); break; }
    for (size_t i = 0; rc == OMV_OK && i < order_.size(); ++i) {
        const Job& J = jobs_[order_[i]];
        StreamRes& r = sres_[i % ns];
        float* film = dFilm_.as<float>() + J.filmOff;
        const int spp = std::max(J.p.samples_per_pixel, 0);
        if (spp == 0) {   // PinholeCamera::stop with a zero count: rgb 0, alpha 1
            if ((e = launchFilm(r.rad.as<float>(), 0, 0, J.npix, film, r.s)) != hipSuccess) rc = hipFail(e, "film kernel launch");
            continue;
        }
        const int lvCap = lvCapOf(J.p);
        WfLaunch L;
        std::memset(&L, 0, sizeof(L));
        L.S = J.scene->view(); L.cam = J.cam; L.p = J.p; L.lvCap = lvCap;
        L.lc = std::max(1, std::min(lvCap, std::max(J.p.max_path_length, 1)));
        L.arena = r.arena.ptr; L.qcount = r.qcount.as<uint32_t>(); L.rad = r.rad.as<float>(); L.gstack = r.gstack.as<uint32_t>();
        L.counters = dCounters_.as<unsigned long long>();
        L.blocksTrace = cus_ * wfPerCu_[0];
        L.blocksShade = cus_ * shadeCamBlocksPerCU(wfPerCu_, J.scene->view().num_lights > 0,
                                                   (J.p.extensions & OMV_EXT_HDRI_NEE) && J.scene->view().env_tex_pdf);
        L.blocksGen = cus_ * wfPerCu_[3];
        const uint64_t chunkPix = std::max<uint64_t>(1, wfMaxSlots_ / (uint64_t)spp);
        for (uint64_t pix0 = 0; rc == OMV_OK && pix0 < J.npix; pix0 += chunkPix) {
            const uint64_t np = std::min(chunkPix, J.npix - pix0);
            L.sample0 = pix0 * (uint64_t)spp;
            L.nslots = (uint32_t)(np * (uint64_t)spp);
            if ((e = launchWavefront(L, r.s)) != hipSuccess) rc = hipFail(e, "wavefront launch");
            else if ((e = launchFilm(r.rad.as<float>(), spp, pix0, np, film, r.s)) != hipSuccess) rc = hipFail(e, "film kernel launch");
        }
    }
    for (int k = 0; k < forked; ++k) {
        if ((e = hipEventRecord(sres_[k].done, sres_[k].s)) != hipSuccess) { if (rc == OMV_OK) rc = hipFail(e, "join"); continue; }
        if ((e = hipStreamWaitEvent(stream_, sres_[k].done, 0)) != hipSuccess && rc == OMV_OK) rc = hipFail(e, "join wait");
    }
    return rc;
}

int BatchRunner::launch() {
    if (tuning_.streams > 0) return launchWavefrontJobs();
    hipError_t e = launchBatch(dJobs_.ptr, (int)jobs_.size(), dTileEnd_.as<uint64_t>(), totalTiles_, tuning_.tile_waves,
                               dCounters_.as<unsigned long long>(), dLv_.as<float>(), dStack_.as<uint32_t>(), blocks_, anyNee_, stream_);
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

// Measures every candidate with plain launches (no graph capture during the
// sweep), reps runs each, and keeps the fastest; the chosen configuration is
// captured once, at the next run, when the starting tuning asked for a graph.
// Candidates: the jobs' wavefront sequences over 1, 2, 4 and 8 streams (no
// more streams than jobs), and the persistent per-pixel kernel at every
// resident-blocks-per-CU count up to the occupancy limit x 1, 2 or 4 waves
// of pixels per work item.
int BatchRunner::autotune(int reps, omv_batch_tuning* best) {
    if (jobs_.empty()) { setError("batch: no jobs"); return OMV_E_STATE; }
    reps = std::max(1, reps);
    bool nee = false;
    for (const Job& j : jobs_) nee = nee || ((j.p.extensions & OMV_EXT_HDRI_NEE) && j.scene->view().env_tex_pdf);
    const int occ = std::max(1, occupancyBlocksPerCU(nee ? 6 : 3));
    const omv_batch_tuning start = tuning_;
    omv_batch_tuning bestT = start;
    double bestMs = 1e300;
    const int tiles[] = {1, 2, 4};
    std::vector<omv_batch_tuning> cands;
    for (int ns : {1, 2, 4, 8})
        if (ns == 1 || ns <= (int)jobs_.size()) cands.push_back(omv_batch_tuning{0, 1, 0, ns});
    for (int bpc = 1; bpc <= occ; ++bpc)
        for (int tw : tiles) cands.push_back(omv_batch_tuning{bpc, tw, 0, 0});
    tunedMs_.clear();
    for (const omv_batch_tuning& t : cands) {
        int rc = setTuning(t);
        if (rc) return rc;
        double m = 1e300;
        for (int r = 0; r < reps; ++r) {
            rc = run();
            if (rc) return rc;
            m = std::min(m, stats_.render_ms);
        }
        tunedMs_.push_back(m);
        if (m < bestMs) { bestMs = m; bestT = t; }
    }
    bestT.use_graph = start.use_graph;
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
