// hipstub.cpp -- CPU stand-in for the HIP runtime calls the host library makes, with stream
// capture bookkeeping and memory-range checks on every "kernel" (launcher)
// the library calls.  Device memory is host memory (so ASan sees it).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "omv_kernels.h"

static int g_errors = 0;
#define FAIL(...) do { std::fprintf(stderr, "STUB-ERROR: " __VA_ARGS__); std::fprintf(stderr, "\n"); ++g_errors; } while (0)
#define LOG(...) do { if (getenv("STUB_LOG")) { std::fprintf(stderr, "stub: " __VA_ARGS__); std::fprintf(stderr, "\n"); } } while (0)

extern "C" int stub_errors() { return g_errors; }

// ------------------------------------------------------------ memory
static std::map<uintptr_t, size_t> g_alloc;   // base -> size
static std::mutex g_mu;

static bool inRange(const void* p, size_t n, const char* what) {
    if (!p) { if (n) { FAIL("%s: null pointer for %zu bytes", what, n); return false; } return true; }
    uintptr_t a = (uintptr_t)p;
    auto it = g_alloc.upper_bound(a);
    if (it == g_alloc.begin()) { FAIL("%s: %p not in any live allocation", what, p); return false; }
    --it;
    if (a + n > it->first + it->second) {
        if (a >= it->first + it->second) FAIL("%s: %p not in any live allocation (nearest %p + %zu)", what, p, (void*)it->first, it->second);
        else FAIL("%s: range %p + %zu past the end of allocation %p + %zu", what, p, n, (void*)it->first, it->second);
        return false;
    }
    if (n) { volatile const char* c = (const char*)p; (void)c[0]; (void)c[n - 1]; }   // ASan sees freed memory
    return true;
}

// ------------------------------------------------------------ streams, events, graphs
struct ihipGraph;
struct ihipStream_t {
    int id;
    ihipGraph* cap = nullptr;
    long seq = 0;        // operations issued on this stream
    bool destroyed = false;
};
struct ihipEvent_t {
    ihipGraph* cap = nullptr;
    ihipStream_t* from = nullptr;
    long seq = 0;
    bool timing = true;
};
struct Node { std::string name; std::function<void()> touch; };
struct ihipGraph {
    ihipStream_t* origin = nullptr;
    std::set<ihipStream_t*> members;
    std::map<ihipStream_t*, long> joinedAt;   // side stream -> its seq when the origin waited on it
    std::vector<Node> nodes;
    bool invalid = false;
    bool ended = false;
};
struct hipGraphExec { std::vector<Node> nodes; };

static std::set<ihipStream_t*> g_streams;
static std::set<ihipGraph*> g_graphs;
static int g_nextStream = 1;
static ihipGraph* g_activeCapture = nullptr;   // this thread's capture (ThreadLocal mode)

// run or record one operation on a stream
static void op(hipStream_t s, const char* name, std::function<void()> touch) {
    if (s && !g_streams.count(s)) { FAIL("%s on an unknown / destroyed stream %p", name, (void*)s); return; }
    if (s) s->seq++;
    if (s && s->cap) {
        if (s->cap->invalid) FAIL("%s on stream %d whose capture is invalidated", name, s->id);
        s->cap->nodes.push_back(Node{name, touch});
        LOG("capture %s on stream %d", name, s->id);
    } else {
        if (!s && g_activeCapture) FAIL("%s on the null stream during a capture", name);
        touch();
    }
}

static void syncCall(const char* name) {
    if (g_activeCapture) FAIL("%s called while a stream capture is active", name);
}

extern "C" {

hipError_t hipMalloc(void** p, size_t n) {
    syncCall("hipMalloc");
    *p = std::calloc(1, n ? n : 1);
    g_alloc[(uintptr_t)*p] = n ? n : 1;
    return hipSuccess;
}
hipError_t hipFree(void* p) {
    syncCall("hipFree");
    if (!p) return hipSuccess;
    auto it = g_alloc.find((uintptr_t)p);
    if (it == g_alloc.end()) { FAIL("hipFree of %p, not a live allocation", p); return hipErrorInvalidValue; }
    // a graph still alive that names this allocation is a dangling reference on replay
    g_alloc.erase(it);
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
static bool isDev(const void* p) {
    uintptr_t a = (uintptr_t)p;
    auto it = g_alloc.upper_bound(a);
    if (it == g_alloc.begin()) return false;
    --it;
    return a < it->first + it->second;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    syncCall("hipMemcpy");
    if (isDev(d)) inRange(d, n, "hipMemcpy dst");
    if (isDev(s)) inRange(s, n, "hipMemcpy src");
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st) {
    op(st, "hipMemcpyAsync", [=]() {
        if (isDev(d)) inRange(d, n, "hipMemcpyAsync dst");
        if (isDev(s)) inRange(s, n, "hipMemcpyAsync src");
        if (n) std::memmove(d, s, n);
    });
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st) {
    if (st && st->cap) FAIL("hipMemsetAsync recorded into a captured graph (stream %d)", st->id);
    op(st, "hipMemsetAsync", [=]() { if (inRange(d, n, "hipMemsetAsync")) std::memset(d, v, n); });
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t* s) {
    *s = new ihipStream_t();
    (*s)->id = g_nextStream++;
    g_streams.insert(*s);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    if (!g_streams.count(s)) { FAIL("hipStreamDestroy of unknown stream"); return hipErrorInvalidValue; }
    if (s->cap) FAIL("hipStreamDestroy of stream %d while it is part of a capture", s->id);
    g_streams.erase(s);
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    syncCall("hipStreamSynchronize");
    if (s && s->cap) FAIL("hipStreamSynchronize of capturing stream %d", s->id);
    return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) { syncCall("hipDeviceSynchronize"); return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new ihipEvent_t(); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = new ihipEvent_t(); (*e)->timing = false; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    if (!e) { FAIL("hipEventRecord of null event"); return hipErrorInvalidValue; }
    if (s && !g_streams.count(s)) { FAIL("hipEventRecord on unknown stream"); return hipErrorInvalidValue; }
    e->cap = s ? s->cap : nullptr;
    e->from = s;
    e->seq = s ? s->seq : 0;
    if (s && s->cap && e->timing) LOG("timing event recorded in capture");
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
    if (!g_streams.count(s)) { FAIL("hipStreamWaitEvent on unknown stream"); return hipErrorInvalidValue; }
    if (e->cap) {
        ihipGraph* g = e->cap;
        if (g->ended) { FAIL("wait on an event captured in a capture that has ended"); return hipErrorInvalidValue; }
        if (!s->cap) {   // the stream joins the capture
            s->cap = g;
            g->members.insert(s);
            LOG("stream %d joins the capture of stream %d", s->id, g->origin->id);
        } else if (s->cap != g) {
            FAIL("stream %d waits on an event of another capture", s->id);
            return hipErrorInvalidValue;
        }
        if (s == g->origin && e->from && e->from != g->origin) g->joinedAt[e->from] = e->seq;
    } else if (s->cap) {
        // a capturing stream waiting on an event recorded outside the capture
        FAIL("capturing stream %d waits on an event recorded outside the capture", s->id);
        s->cap->invalid = true;
        return hipErrorStreamCaptureIsolation;
    }
    s->seq++;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    syncCall("hipEventSynchronize");
    if (e->cap) FAIL("hipEventSynchronize of an event recorded in a capture");
    return hipSuccess;
}
static float g_ms = 10.0f;
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    if (a->cap || b->cap) FAIL("hipEventElapsedTime of captured events");
    g_ms = g_ms * 0.97f + 0.1f;
    *ms = g_ms;
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    if (!g_streams.count(s)) { FAIL("begin capture on unknown stream"); return hipErrorInvalidValue; }
    if (s->cap) { FAIL("begin capture on stream %d already capturing", s->id); return hipErrorIllegalState; }
    if (g_activeCapture) FAIL("a second capture begins while one is active");
    ihipGraph* g = new ihipGraph();
    g->origin = s;
    g->members.insert(s);
    s->cap = g;
    g_graphs.insert(g);
    g_activeCapture = g;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* out) {
    *out = nullptr;
    ihipGraph* g = s->cap;
    if (!g || g->origin != s) { FAIL("end capture on stream %d that is not a capture origin", s->id); return hipErrorIllegalState; }
    bool unjoined = false;
    for (ihipStream_t* m : g->members) {
        if (m == s) continue;
        auto it = g->joinedAt.find(m);
        if (it == g->joinedAt.end() || it->second < m->seq) { unjoined = true; FAIL("stream %d not joined before end capture", m->id); }
    }
    for (ihipStream_t* m : g->members) m->cap = nullptr;
    g->ended = true;
    g_activeCapture = nullptr;
    if (unjoined || g->invalid) return unjoined ? hipErrorStreamCaptureUnjoined : hipErrorStreamCaptureInvalidated;
    *out = g;
    LOG("capture of stream %d ended: %zu nodes", s->id, g->nodes.size());
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    if (!g_graphs.count(g)) { FAIL("hipGraphDestroy of unknown graph"); return hipErrorInvalidValue; }
    g_graphs.erase(g);
    delete g;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    if (!g_graphs.count(g)) { FAIL("instantiate of unknown graph"); return hipErrorInvalidValue; }
    *e = new hipGraphExec();
    (*e)->nodes = g->nodes;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
    if (!e) { FAIL("launch of null graph exec"); return hipErrorInvalidValue; }
    op(s, "hipGraphLaunch", [e]() { for (auto& n : e->nodes) n.touch(); });
    return hipSuccess;
}
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
    *v = a == hipDeviceAttributeMultiprocessorCount ? 256 : 1;
    return hipSuccess;
}
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600* p, int) {
    std::memset(p, 0, sizeof(*p));
    std::strcpy(p->gcnArchName, "gfx950-stub");
    p->multiProcessorCount = 256;
    p->totalGlobalMem = (size_t)288 << 30;
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t e) {
    static char buf[64];
    std::snprintf(buf, sizeof(buf), "stub hip error %d", (int)e);
    return buf;
}
hipError_t hipGetLastError(void) { return hipSuccess; }

}  // extern "C"

// ------------------------------------------------------------ launchers (kernels)
namespace omv {

static size_t upb(size_t b) { return (b + 255) & ~(size_t)255; }

static uint32_t segCapFor(uint32_t nslots, int grid) {
    const uint64_t g = (uint64_t)(grid > 1 ? grid : 1), threads = g * OMV_BLOCK;
    const uint64_t perBlock = (uint64_t)OMV_BLOCK * ((nslots + threads - 1) / threads);
    return (uint32_t)(((g + OMV_QSEG - 1) / OMV_QSEG) * perBlock);
}
uint32_t wavefrontSegCap(uint32_t nslots, int gridShade, int gridGen) {
    const uint32_t a = segCapFor(nslots, gridShade), b = segCapFor(nslots, gridGen);
    return a > b ? a : b;
}
size_t wavefrontArenaBytes(uint32_t nslots, int lvCap, int lc, bool lights, bool nee, bool splat, bool sortRays, int gridShade, int gridGen) {
    const size_t N = nslots;
    const size_t Q = (size_t)OMV_QSEG * wavefrontSegCap(nslots, gridShade, gridGen);
    size_t b = 6 * upb(Q * 16) + upb(N * 16) + upb(N * 8) + 3 * upb(N * 16) + 3 * upb(N * 4);
    if (lights)
        b += upb(N * 16) + 3 * upb(N * 4) + upb(N * (size_t)lvCap * 24) + upb(N * (size_t)lc * 12) + upb(N * (size_t)lc) +
             3 * upb(Q * (size_t)lc * 16);
    if (nee) b += upb(N * 16) + upb(N) + 2 * upb(Q * 16);
    if (lights && splat) b += upb(N * 16);
    if (sortRays) b += upb(N * 4) + upb(4096 * 4);
    b += upb(64 * 16 * 8);   // traversal stat slots
    return b;
}
int wavefrontCounterCount(int B, int lvCap) { return OMV_QSEG * (2 * (lvCap + 2) + 3 * (B + 2)); }
void wavefrontBlocksPerCU(int out[9]) { out[0] = 8; out[1] = 6; out[2] = 6; out[3] = 8; out[4] = 4; out[5] = 8; out[6] = 6; out[7] = 6; out[8] = 8; }
int occupancyBlocksPerCU(int) { return 4; }
int adjointBlocksPerCU() { return 4; }

static void sceneCheck(const OmvSceneView& S, const char* what) {
    std::string w = std::string(what) + " scene";
    if (S.num_meshes > 0) inRange(S.meshes, sizeof(OmvMesh) * S.num_meshes, (w + " meshes").c_str());
    if (S.num_insts > 0) inRange(S.insts, sizeof(OmvInstance) * S.num_insts, (w + " instances").c_str());
    if (S.num_mats > 0) inRange(S.mats, sizeof(OmvMaterial) * S.num_mats, (w + " materials").c_str());
    inRange(S.nodes, 1, (w + " nodes").c_str());
    inRange(S.tris, 1, (w + " tris").c_str());
    inRange(S.vtx, 1, (w + " vertices").c_str());
}

hipError_t launchWavefront(const WfLaunch& L, hipStream_t stream) {
    WfLaunch c = L;
    op(stream, "wavefront", [c]() {
        const bool lights = c.S.num_lights > 0;
        const bool nee = (c.p.extensions & OMV_EXT_HDRI_NEE) && c.S.env_tex_pdf;
        const bool splat = lights && c.splat && (c.p.extensions & OMV_EXT_LIGHT_SPLAT);
        inRange(c.arena, wavefrontArenaBytes(c.nslots, c.lvCap, c.lc, lights, nee, splat, c.sortRays, c.blocksShade, c.blocksGen), "wavefront arena");
        inRange(c.qcount, wavefrontCounterCount(c.p.camera_bounces, c.lvCap) * 4, "wavefront qcount");
        inRange(c.rad, (size_t)c.nslots * 12, "wavefront rad");
        inRange(c.gstack, (size_t)c.blocksTrace * OMV_BLOCK * (OMV_STACK_TOTAL - OMV_STACK_LDS) * 4, "wavefront gstack");
        inRange(c.counters, 64, "wavefront counters");
        sceneCheck(c.S, "wavefront");
    });
    return hipSuccess;
}
hipError_t launchFilm(const float* rad, int spp, uint64_t pix0, uint64_t npix, float* film, hipStream_t stream) {
    op(stream, "film", [=]() {
        if (spp > 0) inRange(rad, npix * (uint64_t)spp * 12, "film rad");
        inRange(film + pix0 * 4, npix * 16, "film out");
    });
    return hipSuccess;
}
hipError_t launchSplatAdd(const float* splat, uint64_t npix, float, float* film, hipStream_t stream) {
    op(stream, "splat add", [=]() { inRange(splat, npix * 12, "splat"); inRange(film, npix * 16, "splat film"); });
    return hipSuccess;
}
hipError_t launchFillU64(unsigned long long* p, int n, hipStream_t stream) {
    op(stream, "fill", [=]() { if (inRange(p, (size_t)n * 8, "fill")) std::memset(p, 0, (size_t)n * 8); });
    return hipSuccess;
}
struct StubJob { OmvSceneView S; omv_camera_desc cam; omv_renderer_params p; float* film; uint64_t npix; };
size_t batchJobBytes() { return sizeof(StubJob); }
void packBatchJob(void* dst, const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, float* film, uint64_t npix) {
    StubJob j;
    std::memset(&j, 0, sizeof(j));
    j.S = S; j.cam = cam; j.p = p; j.film = film; j.npix = npix;
    std::memcpy(dst, &j, sizeof(j));
}
hipError_t launchBatch(const void* jobs, int njobs, const uint64_t* tileEnd, uint64_t totalTiles, int tileWaves,
                       unsigned long long* counters, float* lv, uint32_t* gstack, int blocks, bool, hipStream_t stream) {
    op(stream, "batch fill", [=]() { if (inRange(counters, 128, "batch counters")) std::memset(counters, 0, 128); });
    op(stream, "batch kernel", [=]() {
        if (!inRange(jobs, (size_t)njobs * sizeof(StubJob), "batch jobs")) return;
        inRange(tileEnd, (size_t)njobs * 8, "batch tileEnd");
        if (tileEnd[njobs - 1] != totalTiles) FAIL("batch: tileEnd[last] %llu != totalTiles %llu", (unsigned long long)tileEnd[njobs - 1],
                                                   (unsigned long long)totalTiles);
        const StubJob* J = static_cast<const StubJob*>(jobs);
        int lvMax = 1;
        for (int i = 0; i < njobs; ++i) {
            inRange(J[i].film, J[i].npix * 16, "batch job film");
            sceneCheck(J[i].S, "batch job");
            lvMax = std::max(lvMax, std::max(std::min(J[i].p.light_bounces + 1, std::max(J[i].p.max_path_length, 1)), 1));
        }
        inRange(lv, (size_t)blocks * OMV_BLOCK * lvMax * 24, "batch lv");
        inRange(gstack, (size_t)blocks * OMV_BLOCK * (OMV_STACK_TOTAL - OMV_STACK_LDS) * 4, "batch gstack");
        (void)tileWaves;
    });
    return hipSuccess;
}
hipError_t launchBdpt(const OmvSceneView& S, const omv_camera_desc&, const omv_renderer_params&, uint64_t, uint64_t nsamples, float* rad, float*,
                      uint32_t*, unsigned long long* counters, int, float*, hipStream_t stream) {
    op(stream, "bdpt", [=]() { inRange(rad, nsamples * 12, "bdpt rad"); inRange(counters, 64, "bdpt counters"); sceneCheck(S, "bdpt"); });
    return hipSuccess;
}
hipError_t launchFc(const OmvSceneView&, const omv_camera_desc&, uint32_t, float*, int32_t*, uint8_t*, uint32_t*, unsigned long long*, int,
                    hipStream_t stream) {
    op(stream, "fc", []() {});
    return hipSuccess;
}
hipError_t launchAdjoint(const OmvSceneView&, const omv_camera_desc&, const omv_renderer_params&, uint64_t, uint64_t, const float*, const int*,
                         float* const*, const int*, float*, float*, uint32_t*, unsigned long long*, int, hipStream_t stream) {
    op(stream, "adjoint", []() {});
    return hipSuccess;
}
int detachedBlocksPerCU() { return 2; }
hipError_t launchDetached(const OmvSceneView& S0, const OmvSceneView& S1, const omv_camera_desc&, const omv_renderer_params&, uint64_t npix,
                          double* film, uint32_t* gstack, unsigned long long* counters, int blocks, hipStream_t stream) {
    op(stream, "detached", [=]() {
        inRange(film, npix * 32, "detached film");
        inRange(gstack, (size_t)blocks * OMV_BLOCK * (OMV_STACK_TOTAL - OMV_STACK_LDS) * 4, "detached gstack");
        inRange(counters, 64, "detached counters");
        sceneCheck(S0, "detached");
        sceneCheck(S1, "detached eval");
    });
    return hipSuccess;
}
hipError_t launchPost(const float*, float*, float*, uint8_t*, int, int, float, float, bool, hipStream_t stream) {
    op(stream, "post", []() {});
    return hipSuccess;
}
size_t blasDeviceScratchBytes(uint32_t n) { return 64 + (size_t)n * 64; }
hipError_t buildBlasDevice(const float*, uint32_t, uint32_t, const uint32_t*, uint32_t, uint32_t, int, int, int, uint32_t, OmvNode4*, float*, void*,
                           hipStream_t, int* root, int* nodeCount, int* depth) {
    *root = 0; *nodeCount = 1; *depth = 1;
    return hipSuccess;
}
size_t meshDeviceScratchBytes(uint32_t vn, uint32_t fn) { return 64 + (size_t)(vn + fn) * 64; }
hipError_t preprocessMeshDevice(float*, uint32_t, uint32_t, const uint32_t*, uint32_t, bool, bool, int, void*, hipStream_t, uint32_t* bad) {
    if (bad) *bad = 0;
    return hipSuccess;
}
hipError_t launchEnvBuild(const float* bg, int W, int H, float* rowProb, int* rowAlias, float* colProb, int* colAlias, float* texPdf, void*,
                          hipStream_t stream) {
    op(stream, "env build", [=]() {
        inRange(bg, (size_t)W * H * 16, "env bg");
        inRange(rowProb, (size_t)H * 4, "env rowProb");
        inRange(rowAlias, (size_t)H * 4, "env rowAlias");
        inRange(colProb, (size_t)W * H * 4, "env colProb");
        inRange(colAlias, (size_t)W * H * 4, "env colAlias");
        inRange(texPdf, (size_t)W * H * 4, "env texPdf");
    });
    return hipSuccess;
}
int renderNoalloc(const OmvSceneView&, const omv_camera_desc&, const omv_renderer_params&, float*, hipStream_t, unsigned long long*, hipError_t& e,
                  const char*& w) {
    e = hipSuccess; w = nullptr;
    return 0;
}

}  // namespace omv
