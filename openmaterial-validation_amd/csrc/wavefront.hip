// wavefront.hip -- the render loop of BDPTRenderer (BDPTRenderer.cpp:128-301)
// as a sequence of small gfx950 kernels over compacted ray queues.
//
// Every sample keeps a slot of path state in HBM (structure of arrays, slot
// = sample - first sample of the chunk).  A chunk runs as
//
//   gen             camera ray + light-subpath start of every slot
//   light rounds    trace(light rays) -> shade_light   (vertex i = 1..)
//   camera rounds   trace(camera rays) -> shade_cam    (vertex i = 1..B)
//                   [ -> trace(shadow rays)  when there are lights ]
//                   [ -> trace(environment rays)  with OMV_EXT_HDRI_NEE ]
//                   [ -> connect  when either is there ]
//
// so each trace launch sees only rays that exist, whatever the lengths of
// the paths around them, and the traversal kernel stays small (registers,
// code) next to the material code.  Shading appends the next ray with one
// atomic per wave.  The per-sample arithmetic and the order in which every
// sum is formed are those of the per-sample kernel (kernels.hip), so the
// film is identical to it bit for bit; only the schedule differs.
//
// Philox stream positions persist between kernels as "draws consumed".
#include <algorithm>
#include <cstring>

#include "path_device.h"

// Traversal kernels are compiled for 5 waves per SIMD (<= 96 VGPRs; a few
// bytes of spill in the camera and fused kernels) with a 28-entry LDS stack,
// so 10 workgroups fit a CU's 160 KB of LDS.  Measured on MI355X against 4
// waves (<= 128 VGPRs, 32-entry LDS stack): 2M-triangle instanced scene
// 15.85 -> 15.23 ms, torus knot 8.93 -> 8.53 ms per render; 3 waves
// (6 workgroups per CU) was 16 % slower than 4.
#define OMV_TRACE_WAVES 5
// Resident waves per SIMD the camera-shading kernel is compiled for (the
// register budget it gets: 3 -> 168 VGPRs, 4 -> 128).  Measured on MI355X:
// with point lights 4 (24 bytes of spill) renders the 2M-triangle instanced
// scene 0.2 ms faster than 3; without lights or NEE 4 (68 bytes of spill)
// takes the torus knot's shading from 1.95 to 1.89 ms per render.
#define OMV_SHADE_WAVES(lights, nee) (!(nee) ? 4 : 3)
// Resident waves per SIMD the tail launch is compiled for (1: no bound, the
// register count the compiler picks).  Measured on MI355X: the light-free
// tail at 4 (128 VGPRs, 152 bytes of spill) instead of 3 (167 VGPRs) takes
// the torus knot 512^2 x 64 from 6.14 to 6.00 ms per render -- the tail
// holds few paths per wave and gains from more waves in flight; with point
// lights 3 waves (168 VGPRs) instead of 2 (221) is 0.03 ms slower on the
// instanced scene, and with NEE alone 3 instead of 2 (196 VGPRs) is equal on
// the torus knot (9.50 ms).
#define OMV_TAIL_WAVES(lights, nee) ((lights) || (nee) ? 1 : 4)

namespace {

// ----------------------------------------------------------- state layout
struct WfQueue {       // one ray queue
    float4* o;         // origin xyz, tnear
    float4* d;         // direction xyz, as_float(slot)
};
struct WfShadowQ {     // connection visibility rays (isConnected, BDPTRenderer.cpp:208-235)
    float4* o;         // A xyz (light vertex), tnear = safeD
    float4* d;         // unit direction xyz, tfar
    float4* b;         // B xyz (camera vertex), as_float(slot * lc + l)
};
struct WfHits {
    float4* h;         // t, u, v, tnear after the masked re-trace loop (written for hits only)
    int2* id;          // instance (-1: miss), primitive | tie << 31
};
struct WfState {
    uint32_t n;        // slots
    float4* thr;       // camera throughput xyz, MIS weight of an escape (OMV_EXT_HDRI_NEE)
    float4* radc;      // camera radiance xyz, as_float(vertex count)
    float4* conn;      // connection sum xyz
    uint32_t* csn;     // camera stream draws consumed
    uint32_t* flags;   // 0 running, 1 ends after this round's connections, 2 written
    uint32_t* nconn;   // connections staged this round
    float4* lthr;      // light throughput xyz, first-hit attenuation rD
    uint32_t* lsn;     // light stream draws consumed
    int* lidx;         // sampled light
    int* lplen;        // light subpath vertex count
    float* lv;         // light vertices [vertex][6][slot]
    float* cc;         // staged connection contributions [l][3][slot]
    uint8_t* vis;      // [l][slot] 0 none/blocked, 1 pending, 2 visible
    float4* lbeta;     // light-tracer particle weight for the next vertex (OMV_EXT_LIGHT_SPLAT)
    float4* neeC;      // staged next-event contribution xyz (OMV_EXT_HDRI_NEE)
    uint8_t* nvis;     // [slot] its visibility: 0 none/blocked, 1 pending, 2 visible
    int lvCap, lc;     // light vertex capacity, connections per camera vertex
};

DEV uint32_t rngConsumed(const Rng& r) { return 4u * r.c0 + (uint32_t)r.idx - 4u; }
DEV void rngRestore(Rng& r, uint32_t seed, uint32_t stream, uint64_t sample, uint32_t n) {
    r.init(seed, stream, sample);
    if (n == 0u) return;
    const uint32_t block = (n - 1u) >> 2;
    r.c0 = block;
    r.refill();
    // drop the draws already taken from this block (b0 is the next draw)
    const int k = (int)(n - 4u * block);
    for (int j = 0; j < k; ++j) { r.b0 = r.b1; r.b1 = r.b2; r.b2 = r.b3; }
    r.idx = k;
}

// Slot of the calling lane in a queue: one atomic per wave.  Lanes that do
// not append pass want = false; all active lanes must make the call.
DEV uint32_t waveAppend(uint32_t* counter, bool want) {
    const uint64_t m = __ballot(want);
    if (m == 0ull) return 0u;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

DEV void stackInit(Stack& st, uint32_t* gstack, uint32_t gtid, uint32_t gthreads) { st.init(gstack, gtid, gthreads); }

// The traversal launches add their ray and visit counts into one of
// OMV_STAT_SLOTS slots (128 bytes apart) by workgroup, for the same reason
// the queues have several append counters; omv_wf_stats_sum folds the slots
// into the renderer's counters once per chunk.
#define OMV_STAT_SLOTS 64
DEV unsigned long long* statSlot(unsigned long long* base) { return base + (blockIdx.x % OMV_STAT_SLOTS) * 16u; }

// A queue's entries in K regions of cap entries: region r holds the physical
// entries [r*cap, r*cap + cnt[r]).  A producer appends to the region of its
// workgroup (blockIdx mod K), so the append atomics of a launch are spread
// over K counters instead of serialising on one (a wave-aggregated atomic on
// one address costs ~11 ns of that address's time on MI355X; over 64
// addresses ~0.35 ns).  Consumers walk the logical index 0 .. sum(cnt) - 1
// region by region.  The host sizes cap so that no region can overflow (a
// region is fed by ceil(grid/K) workgroups, each handling at most
// OMV_BLOCK * ceil(N / threads) input entries, with at most `per` appends
// each); a full region still drops the append and counts it in `lost`
// rather than write out of bounds.  K = 1 is a plain compact queue.
struct WfSeg {
    uint32_t* cnt;    // K counters (null: an empty queue)
    uint32_t K, cap;
};
inline WfSeg segOf(uint32_t* cnt, uint32_t K, uint32_t cap) { WfSeg g; g.cnt = cnt; g.K = K; g.cap = cap; return g; }
// Region prefix of a segmented queue, per workgroup in LDS: pre[r] = the
// entries in the regions before r.  The first wave builds it (lane r loads
// region r's count, then a wave scan), so the logical -> physical map is a
// 6-step search in LDS rather than a dependent walk over the counters.
struct SegTable {
    uint32_t pre[OMV_QSEG + 1];
    uint32_t K, cap;
    DEV uint32_t total() const { return pre[K]; }
    DEV uint32_t phys(uint32_t L) const {
        uint32_t r = 0;
#pragma unroll
        for (uint32_t st = OMV_QSEG / 2; st > 0; st >>= 1)
            if (r + st < K && pre[r + st] <= L) r += st;
        return r * cap + (L - pre[r]);
    }
};
static_assert(OMV_QSEG <= 64, "one wave builds a queue's region prefix");
// All threads of the workgroup call it: it synchronises the workgroup.
DEV void segLoad(SegTable& t, const WfSeg& g) {
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x;
        uint32_t v = (g.cnt && lane < g.K) ? min(g.cnt[lane], g.cap) : 0u;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(v, d, 64);
            if ((int)lane >= d) v += y;
        }
        t.pre[lane + 1] = v;
        if (lane == 0) { t.pre[0] = 0u; t.K = g.cnt ? g.K : 1u; t.cap = g.cap; }
    }
    __syncthreads();
}
// Physical slot of the calling lane's entry, or OMV_SEG_NONE when it does not
// append (or its region is full); all active lanes of the wave make the call.
#define OMV_SEG_NONE 0xFFFFFFFFu
DEV uint32_t segAppend(const WfSeg& g, bool want, unsigned long long* lost) {
    const uint32_t r = blockIdx.x % g.K;
    const uint32_t k = waveAppend(g.cnt + r, want);
    if (!want) return OMV_SEG_NONE;
    if (k >= g.cap) { atomicAdd(lost, 1ull); return OMV_SEG_NONE; }
    return r * g.cap + k;
}

// segAppend with the wave's entries ordered by key (key < nkeys <= 8): one
// atomic per wave as before; a lane's slot in the wave's chunk is the count
// of appending lanes with a smaller key, then of lanes before it with the
// same key.  Connection rays use the target light as the key, so the rays
// a traversal wave takes from one chunk head to one light together.
DEV uint32_t segAppendKeyed(const WfSeg& g, bool want, uint32_t key, uint32_t nkeys, unsigned long long* lost) {
    const uint64_t m = __ballot(want);
    if (m == 0ull) return OMV_SEG_NONE;
    const int lane = threadIdx.x & 63;
    const uint64_t before = (1ull << lane) - 1ull;
    uint32_t rank = 0;
    for (uint32_t k = 0; k < nkeys; ++k) {
        const uint64_t mk = __ballot(want && key == k);
        rank += k < key ? (uint32_t)__popcll(mk) : (k == key ? (uint32_t)__popcll(mk & before) : 0u);
    }
    const uint32_t r = blockIdx.x % g.K;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(g.cnt + r, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (!want) return OMV_SEG_NONE;
    const uint32_t k = base + rank;
    if (k >= g.cap) { atomicAdd(lost, 1ull); return OMV_SEG_NONE; }
    return r * g.cap + k;
}

DEV f3 xyz(const float4& v) { return mk3(v.x, v.y, v.z); }
DEV float4 f4of(const f3& v, float w) { float4 r; r.x = v.x; r.y = v.y; r.z = v.z; r.w = w; return r; }

// Round-1 rays of sample s: the camera ray (PinholeCamera::getRay, 2 draws
// of the camera stream) offset from the camera position -- computed where it
// is used (the round-1 camera traversal, shading and connect kernels)
// instead of stored by the generation kernel -- and the light choice and
// emitted ray (tryComputeLightPath, BDPTRenderer.cpp:194-206) with the
// light-stream draws they consumed.
DEV void genCameraRay(const omv_camera_desc& cam, uint32_t seed, int spp, uint64_t s, f3& org, f3& dir) {
    int x, y;
    if (s < (1ull << 32)) {   // 32-bit division (the 64-bit one expands to a long sequence)
        const uint32_t pix = (uint32_t)s / (uint32_t)spp;
        x = (int)(pix % (uint32_t)cam.width); y = (int)(pix / (uint32_t)cam.width);
    } else {
        const uint64_t pix = s / (uint64_t)spp;
        x = (int)(pix % (uint64_t)cam.width); y = (int)(pix / (uint64_t)cam.width);
    }
    Rng cs; cs.init(seed, STREAM_CAMERA, s);
    const float u0 = cs.next(), u1 = cs.next();
    f3 O;
    cameraRay(cam, x, y, u0, u1, O, dir);
    const float delta = 3.0f * FLT_EPSILON * maxAbs(O);
    org = mk3(O.x + delta * dir.x, O.y + delta * dir.y, O.z + delta * dir.z);
}
DEV void genLightRay(Rng& ls, const f3& lp, f3& org, f3& dir) {   // after the light choice's draw
    dir = uniformSphere(ls);
    const float delta = 3.0f * FLT_EPSILON * maxAbs(lp);
    org = mk3(lp.x + delta * dir.x, lp.y + delta * dir.y, lp.z + delta * dir.z);
}
// True when every light ray from lp misses the children of the TLAS root,
// whatever direction it draws: each non-empty child lies below lp on some
// axis by 2e-30 or more.  The direction is in the closed positive octant
// (uniformSphere), so the origin lp + delta * dir is >= lp on every axis,
// and on that axis the far-plane distance (hi - o) * idir of the root-box
// test is at most about -2e-30 times idir (idir of about 1 or more: a
// negative normal float, or -inf for a zero component); its tmax is then
// below its tmin >= 0 and the child is missed.  The root-box test culls
// such a ray for every direction, so the direction need not be drawn.
DEV bool lightBelowRoot(const OmvSceneView& S, const f3& lp) {
    const char* base = reinterpret_cast<const char*>(S.nodes) + ((size_t)(uint32_t)S.tlas_root << 7);
    const float4 hx = ldRow(base, 16u), hy = ldRow(base, 48u), hz = ldRow(base, 80u);   // the +x, +y, +z far planes
    const uint4 en = *reinterpret_cast<const uint4*>(base + 112u);
    const float HX[4] = {hx.x, hx.y, hx.z, hx.w}, HY[4] = {hy.x, hy.y, hy.z, hy.w}, HZ[4] = {hz.x, hz.y, hz.z, hz.w};
    const uint32_t E[4] = {en.x, en.y, en.z, en.w};
    bool below = true;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        below = below && (E[k] == (uint32_t)E_NONE || lp.x - HX[k] >= 2e-30f || lp.y - HY[k] >= 2e-30f || lp.z - HZ[k] >= 2e-30f);
    return below;
}
// Where the round-1 camera kernels find their rays: the camera queue, or
// (gen) computed from the entry, which is the slot
struct WfGen {
    int gen;                 // 1: round-1 rays are computed, entry e is slot e
    omv_camera_desc cam;
    uint32_t seed;
    int spp;
    uint64_t sample0;
};

// Light vertices l >= 1 of one slot in the wavefront state ([vertex][6][slot])
struct WfLightVerts {
    const float* lv;
    uint64_t N;
    uint32_t slot;
    DEV f3 pos(int l) const {
        return mk3(lv[((uint64_t)l * 6 + 0) * N + slot], lv[((uint64_t)l * 6 + 1) * N + slot], lv[((uint64_t)l * 6 + 2) * N + slot]);
    }
    DEV f3 rad(int l) const {
        return mk3(lv[((uint64_t)l * 6 + 3) * N + slot], lv[((uint64_t)l * 6 + 4) * N + slot], lv[((uint64_t)l * 6 + 5) * N + slot]);
    }
};

// XCD-banded walk over a queue of `total` entries.  Workgroups are dispatched
// to the X XCDs round-robin (workgroup b on XCD b mod X), and each XCD has an
// L2 of its own.  The queue is cut into chunks of C entries (a power of two,
// 128 .. 32768, about total / 16X), dealt to the XCDs round-robin, and the
// workgroups of one XCD stride over that XCD's chunks as over one array, so
// the rays an XCD traces at a time come from a few contiguous stretches of
// the queue (for camera rays: of the image) and share their node and
// triangle lines in that XCD's L2.  Every entry below `total` is visited by
// exactly one thread: chunk g belongs to XCD g mod X as its chunk g div X,
// and within an XCD the offsets v = local thread + k * (threads of the XCD)
// cover 0, 1, 2, ... once each.  With X = 1, or a grid that is not a
// multiple of X, it is the plain grid stride.  Entries must stay below 2^31.
struct XcdWalk {
    uint32_t x, X, logC, v, stride, total, e;
    DEV void start(uint32_t tot, uint32_t nX) {
        total = tot;
        X = nX > 1u && gridDim.x % nX == 0u ? nX : 1u;
        x = blockIdx.x % X;
        v = (blockIdx.x / X) * blockDim.x + threadIdx.x;
        stride = (gridDim.x / X) * blockDim.x;
        const uint32_t want = tot / (16u * X);
        logC = 7u;
        while (logC < 15u && (2u << logC) <= want) ++logC;
        e = at(v);
    }
    DEV uint32_t at(uint32_t vv) const {
        return ((((vv >> logC) * X) + x) << logC) | (vv & ((1u << logC) - 1u));
    }
    DEV bool more() const { return e < total; }
    DEV void next() { v += stride; e = at(v); }
};

}  // namespace

// ------------------------------------------------------------------ gen
// Camera ray of every slot (PinholeCamera::getRay, 2 draws) and, with
// lights, the light-subpath start (tryComputeLightPath, BDPTRenderer.cpp:194-206).
// With writeQueues = 0 the round-1 camera rays are not stored: the round-1
// camera kernels compute them (WfGen).  Light rays are stored (their
// direction comes from a rejection loop, cheaper to keep than to redo) by
// appending to the segmented round-1 light queue (lightSeg, whose regions
// are sized for this launch's grid) -- except a ray that
// misses all four child boxes of the TLAS root: the traversal's first step
// is that same test (boxTest4Rows, tnear 0, tfar FLT_MAX), after which it
// ends without a hit, so such a ray is not traced.  Its trace is still
// counted (one ray, one light ray, the root's node visit), and light round
// 1 leaves its subpath at the light vertex, as for a traced miss.  A light
// that lies above every root child on some axis has all its rays culled
// (lightBelowRoot); their directions are not drawn.
template <bool kLights>
__global__ __launch_bounds__(OMV_BLOCK) void omv_wf_gen(OmvSceneView S, omv_camera_desc cam, RP P, int spp, uint64_t sample0,
                                                        WfState W, WfQueue camQ, WfQueue lightQ, uint32_t* __restrict__ camCount,
                                                        WfSeg lightSeg, float* __restrict__ rad, int writeQueues,
                                                        unsigned long long* __restrict__ counters, unsigned long long* __restrict__ lost) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = W.n;
    const bool camRays = P.camera_bounces >= 1;
    if (gtid == 0) *camCount = camRays ? n : 0u;
    uint32_t culled = 0;
    for (uint32_t slot = gtid; slot < n; slot += gthreads) {
        const uint64_t s = sample0 + slot;
        // the initial camera state (throughput 1, escape weight 1, radiance 0
        // after 1 vertex, connection sum 0, 2 draws consumed) is not stored:
        // camera round 1 and its connect kernel start from it
        W.flags[slot] = 0u;
        if (camRays) {
            if (writeQueues) {
                f3 org, D;
                genCameraRay(cam, P.seed, spp, s, org, D);
                camQ.o[slot] = f4of(org, 0.0f);
                camQ.d[slot] = f4of(D, __uint_as_float(slot));
            }
        } else {
            // no camera vertex: conn * 1 + radiance with both zero
            rad[(uint64_t)slot * 3 + 0] = 0.0f;
            rad[(uint64_t)slot * 3 + 1] = 0.0f;
            rad[(uint64_t)slot * 3 + 2] = 0.0f;
            W.flags[slot] = 2u;
        }
        if (kLights) {
            // light vertex 0 is the sampled light itself (read from the light
            // table where it is needed); the light throughput starts at 1
            // in light round 1
            // W.lsn (the light-stream draws consumed) is read only where a
            // traced light ray hits, so a culled ray need not store it
            Rng ls; ls.init(P.seed, STREAM_LIGHT, s);
            const int li = sampleLight(S, ls.next());
            W.lidx[slot] = li;
            if (!S.lights_leave) W.lplen[slot] = 1;   // with lights_leave every subpath keeps 1 vertex: not stored
            if (W.lvCap > 1) {
                const f3 lp = ld3(S.lights[li].pos);
                bool trace = !(S.tlas_root >= 0 && (S.lights_leave || lightBelowRoot(S, lp)));
                f3 org = lp, dir = mk3(0.0f, 0.0f, 1.0f);
                if (trace) {
                    genLightRay(ls, lp, org, dir);
                    W.lsn[slot] = rngConsumed(ls);
                    if (S.tlas_root >= 0) {
                        const f3 idir = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
                        float tent[4]; bool hit[4]; uint32_t ent[4];
                        boxTest4Rows(S.nodes, (uint32_t)S.tlas_root, nearRows(idir), org, idir, 0.0f, FLT_MAX, tent, hit, ent);
                        trace = false;
#pragma unroll
                        for (int k = 0; k < 4; ++k) trace = trace || (hit[k] && ent[k] != (uint32_t)E_NONE);
                    }
                }
                const uint32_t k = segAppend(lightSeg, trace, lost);
                if (k != OMV_SEG_NONE) {
                    lightQ.o[k] = f4of(org, 0.0f);
                    lightQ.d[k] = f4of(dir, __uint_as_float(slot));
                }
                culled += trace ? 0u : 1u;
            }
        }
    }
    if (kLights) {
        Counters c; c.nodes = culled; c.tris = 0; c.overflow = 0;
        flushCounters(statSlot(counters), culled, 0u, culled, 0u, c);
    }
}

// ---------------------------------------------------------------- trace
// sceneIntersect (BDPTRenderer.cpp:75-102) over a queue.  kKind 0: camera
// rays, 1: light rays (closest hit + the re-trace tnear), 2: connection
// rays (the visibility decision of isConnected, BDPTRenderer.cpp:221-234),
// 3: environment rays of OMV_EXT_HDRI_NEE (visible iff nothing is hit),
// 4: pinhole rays of OMV_EXT_LIGHT_SPLAT (visible iff nothing is hit; the
//    visible contributions are added to the splat film).
template <int kKind>
__global__ __launch_bounds__(OMV_BLOCK, OMV_TRACE_WAVES) void omv_wf_trace(OmvSceneView S, WfQueue q, WfShadowQ sq, WfSeg count,
                                                          WfHits H, uint8_t* __restrict__ vis, uint32_t nslots, int lc,
                                                          uint32_t* __restrict__ gstack, unsigned long long* __restrict__ counters,
                                                          float* __restrict__ splat, uint32_t xcds, WfGen G) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    stackInit(st, gstack, gtid, gthreads);
    Counters cnt; cnt.nodes = 0; cnt.tris = 0; cnt.overflow = 0;
    uint32_t rays = 0;
    XcdWalk w;
    __shared__ SegTable cur;
    segLoad(cur, count);
    for (w.start(cur.total(), xcds); w.more(); w.next()) {
        const uint32_t e = w.e;                                       // logical entry (hit records)
        const uint32_t p = kKind == 0 && G.gen ? e : cur.phys(e);   // its queue slot
        Hit h;
        if (kKind == 2) {
            const float4 a = sq.o[p], dd = sq.d[p], b = sq.b[p];
            float tnear = a.w;
            bool visible;
            if (!sceneIntersect(S, xyz(a), xyz(dd), tnear, dd.w, h, st, cnt, rays, connectStop(xyz(a), xyz(b)))) {
                visible = true;
            } else {
                f3 df = sub3(hitPosition(S, h), xyz(b));
                float sd = dot3(df, df);
                float safeD = 10.0f * FLT_EPSILON * maxAbs(xyz(b));
                visible = sd <= safeD * safeD;
            }
            const uint32_t id = __float_as_uint(b.w);
            const uint32_t slot = id / (uint32_t)lc, l = id - slot * (uint32_t)lc;
            vis[(uint64_t)l * nslots + slot] = visible ? 2 : 0;
        } else if (kKind == 3) {
            const float4 o = q.o[p], d = q.d[p];
            float tnear = o.w;
            const bool hit = sceneIntersect(S, xyz(o), xyz(d), tnear, FLT_MAX, h, st, cnt, rays, FLT_MAX);
            vis[__float_as_uint(d.w)] = hit ? 0 : 2;
        } else if (kKind == 4) {
            const float4 a = sq.o[p], dd = sq.d[p], b = sq.b[p];
            float tnear = a.w;
            const bool hit = sceneIntersect(S, xyz(a), xyz(dd), tnear, dd.w, h, st, cnt, rays, FLT_MAX);
            splatAdd(splat, !hit, __float_as_uint(b.w), xyz(b));
        } else {
            f3 org, dir;
            float tnear = 0.0f;
            if (kKind == 0 && G.gen) {   // round 1: the camera ray of slot e
                genCameraRay(G.cam, G.seed, G.spp, G.sample0 + e, org, dir);
            } else {
                const float4 o = q.o[p], d = q.d[p];
                org = xyz(o); dir = xyz(d); tnear = o.w;
            }
            const bool hit = sceneIntersect(S, org, dir, tnear, FLT_MAX, h, st, cnt, rays);
            float4 r; r.x = h.t; r.y = h.u; r.z = h.v; r.w = tnear;
            int2 id; id.x = hit ? h.inst : -1; id.y = h.prim | (h.tie ? (int)0x80000000 : 0);
            if (hit) H.h[e] = r;   // read by the shading kernels for hits only
            H.id[e] = id;
        }
    }
    // every entry below the total is traced once, so the ray kinds are
    // counted once per launch rather than per lane
    const uint32_t all = blockIdx.x == 0 && threadIdx.x == 0 ? cur.total() : 0u;
    flushCounters(statSlot(counters), rays, kKind == 0 ? all : 0u, kKind == 1 ? all : 0u, kKind >= 2 ? all : 0u, cnt);
}

// Fused traversal of one camera round boundary: the camera extension rays of
// bounce i + 1 (kind 0), then the connection rays (kind 2) and environment
// rays (kind 3) of bounce i, one grid over the three queues back to back.
// The three are independent given bounce i's shading, and each entry gets
// exactly the per-ray arithmetic of its own kind, so the outputs are those
// of three separate launches; one launch pays one latency tail instead of
// up to three.  A null count is an empty queue.
__global__ __launch_bounds__(OMV_BLOCK, OMV_TRACE_WAVES) void omv_wf_trace_fused(
    OmvSceneView S, WfQueue q, WfSeg qCount, WfShadowQ sq, WfSeg sqCount, WfQueue eq, WfSeg eqCount, WfHits H, uint8_t* __restrict__ vis,
    uint8_t* __restrict__ nvis, uint32_t nslots, int lc, uint32_t* __restrict__ gstack, unsigned long long* __restrict__ counters,
    uint32_t xcds) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    stackInit(st, gstack, gtid, gthreads);
    Counters cnt; cnt.nodes = 0; cnt.tris = 0; cnt.overflow = 0;
    uint32_t rays = 0;
    __shared__ SegTable cA, cB, cC;
    segLoad(cA, qCount); segLoad(cB, sqCount); segLoad(cC, eqCount);
    const uint32_t nA = cA.total(), nB = cB.total(), nC = cC.total();
    XcdWalk w;
    for (w.start(nA + nB + nC, xcds); w.more(); w.next()) {
        const uint32_t e = w.e;
        // the entry's ray: one traversal call site for the three kinds (the
        // arguments each kind's own launch would pass), then its epilogue
        const int kind = e < nA ? 0 : (e < nA + nB ? 2 : 3);
        float4 o, d;
        float tfar = FLT_MAX, tstop = -1.0f;
        if (kind == 0) {
            const uint32_t p = cA.phys(e);
            o = q.o[p]; d = q.d[p];
        } else if (kind == 2) {
            const uint32_t k = cB.phys(e - nA);
            o = sq.o[k]; d = sq.d[k];
            tfar = d.w;
            tstop = connectStop(xyz(o), xyz(sq.b[k]));   // the target is read again after the traversal
        } else {
            const uint32_t k = cC.phys(e - nA - nB);
            o = eq.o[k]; d = eq.d[k];
            tstop = FLT_MAX;
        }
        Hit h;
        float tnear = o.w;
        const bool hit = sceneIntersect(S, xyz(o), xyz(d), tnear, tfar, h, st, cnt, rays, tstop);
        if (kind == 0) {
            float4 r; r.x = h.t; r.y = h.u; r.z = h.v; r.w = tnear;
            int2 id; id.x = hit ? h.inst : -1; id.y = h.prim | (h.tie ? (int)0x80000000 : 0);
            if (hit) H.h[e] = r;   // read by the shading kernels for hits only
            H.id[e] = id;
        } else if (kind == 2) {
            const float4 b = sq.b[cB.phys(e - nA)];   // not held in registers across the traversal
            bool visible = true;
            if (hit) {
                f3 df = sub3(hitPosition(S, h), xyz(b));
                float sd = dot3(df, df);
                float safeD = 10.0f * FLT_EPSILON * maxAbs(xyz(b));
                visible = sd <= safeD * safeD;
            }
            const uint32_t id = __float_as_uint(b.w);
            const uint32_t slot = id / (uint32_t)lc, l = id - slot * (uint32_t)lc;
            vis[(uint64_t)l * nslots + slot] = visible ? 2 : 0;
        } else {
            nvis[__float_as_uint(eq.d[cC.phys(e - nA - nB)].w)] = hit ? 0 : 2;   // the slot, read again
        }
    }
    // every entry is traced once: the kinds are counted once per launch
    const bool first = blockIdx.x == 0 && threadIdx.x == 0;
    flushCounters(statSlot(counters), rays, first ? nA : 0u, 0u, first ? nB + nC : 0u, cnt);
}

// ------------------------------------------------ packet traversal (round 1)
// Round-1 camera rays are generated in sample order, so the 64 lanes of a
// wave hold 64 consecutive samples: a few neighbouring pixels (one at 64
// samples per pixel), nearly parallel rays that visit nearly the same nodes.
// omv_wf_trace_cam1 traverses them as one packet (masked SIMD packet
// traversal, as Wald, Slusallek, Benthin & Wagner 2001 and Garanzha & Loop
// 2010 with per-lane masks): the wave takes one entry at a time from a
// wave-uniform stack of (entry, lane mask) pairs in LDS; node, instance and
// triangle records are read once per wave with scalar loads; every lane of
// the entry's mask tests its own ray against the four boxes or the leaf's
// triangles, and a child is pushed with the mask of the lanes whose ray hits
// its box within their current closest t.  Children are visited nearest
// first by the entry distances of one lane (the first lane that hits any of
// them).  Each lane still sees every leaf its own traversal would (its box
// tests are the same conservative slab tests against its own closest t) and
// offer() keeps the (t, instance, primitive) minimum whatever the order, so
// hits equal those of the per-ray traversal.
typedef __attribute__((address_space(4))) const float ConstF;
DEV float4 ldUniform(const void* base, uint32_t i) {
    const ConstF* p = (const ConstF*)base + 4u * i;
    float4 r; r.x = p[0]; r.y = p[1]; r.z = p[2]; r.w = p[3];
    return r;
}
DEV uint32_t wuni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
DEV uint64_t wuni64(uint64_t v) { return ((uint64_t)wuni((uint32_t)(v >> 32)) << 32) | wuni((uint32_t)v); }

struct PacketRay {   // one lane's ray and closest hit
    f3 wo, wd, widir;   // world ray
    Pre op;             // object ray of the packet's current instance
    Hit h;
    float tnear, tfar;
    bool found;
};

// One traceClosest of every lane in M0 (closest hit, no occlusion early-out).
// stE / stM: this wave's LDS stack of OMV_STACK_TOTAL entries.
DEV void packetTrace(const OmvSceneView& S, PacketRay& R, uint64_t M0, uint32_t* stE, uint64_t* stM, Counters& c) {
    const uint32_t lane = threadIdx.x & 63u;
    int sp = 0;
    uint32_t e = (E_TLAS_NODE << E_KIND_SHIFT) | (uint32_t)S.tlas_root;
    uint64_t M = M0;
    bool has = M != 0ull;
    int curInst = -1;
    for (;;) {
        if (!has) {
            if (sp == 0) break;
            --sp;
            e = wuni(stE[sp]);
            M = wuni64(stM[sp]);
        }
        has = false;
        const bool act = (M >> lane) & 1ull;
        const uint32_t kind = e >> E_KIND_SHIFT, idx = e & 0x3FFFFFFFu;
        if (kind == E_BLAS_LEAF) {
            const uint32_t first = idx >> 3, cnt = (idx & 7u) + 1u;
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4 r0 = ldUniform(S.tris, 3u * (first + k)), r1 = ldUniform(S.tris, 3u * (first + k) + 1u),
                             r2 = ldUniform(S.tris, 3u * (first + k) + 2u);
                if (act) {
                    ++c.tris;
                    float t, u, v;
                    if (triIntersect(R.op, r0, r1, r2, R.tnear, R.found ? R.h.t : R.tfar, t, u, v))
                        offer(R.h, R.found, t, u, v, curInst, __float_as_int(r0.w));
                }
            }
            continue;
        }
        if (kind == E_INST) {
            // (instance end markers are not needed: the BLAS entries on the
            // stack always belong to the last instance entered)
            const int inst = (int)(idx & 0x1FFFFFFFu);
            const OmvInstance& in = S.insts[inst];
            const float* I = in.inv;
            bool ok = false;
            if (act) {
                const f3 o = R.wo, d = R.wd;
                f3 oo = mk3(I[0] * o.x + I[1] * o.y + I[2] * o.z + I[3], I[4] * o.x + I[5] * o.y + I[6] * o.z + I[7],
                            I[8] * o.x + I[9] * o.y + I[10] * o.z + I[11]);
                f3 od = mk3(I[0] * d.x + I[1] * d.y + I[2] * d.z, I[4] * d.x + I[5] * d.y + I[6] * d.z, I[8] * d.x + I[9] * d.y + I[10] * d.z);
                ok = rayPrep(oo, od, R.op);
            }
            const uint64_t M2 = __ballot(ok);
            if (M2 == 0ull) continue;
            curInst = inst;
            e = (E_BLAS_NODE << E_KIND_SHIFT) | wuni((uint32_t)in.blas_root);
            M = M2;
            has = true;
            continue;
        }
        // internal node (TLAS: world ray, BLAS: object ray)
        OmvNode4 nd;
        {
            float4* dst = reinterpret_cast<float4*>(&nd);
#pragma unroll
            for (int i = 0; i < 8; ++i) dst[i] = ldUniform(S.nodes + idx, (uint32_t)i);
        }
        const float tb = R.found ? R.h.t : R.tfar;
        float tent[4]; bool hit[4];
        if (kind == E_TLAS_NODE) boxTest4(nd, R.wo, R.widir, R.tnear, tb, tent, hit);
        else boxTest4(nd, R.op.o, R.op.idir, R.tnear, tb, tent, hit);
        if (act) ++c.nodes;
        uint64_t m[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) m[k] = nd.entry[k] != E_NONE ? __ballot(act && hit[k]) : 0ull;   // empty slots: never pushed
        const uint64_t any = m[0] | m[1] | m[2] | m[3];
        if (any == 0ull) continue;
        // order by the first hitting lane's entry distances; a child that
        // lane misses goes after the ones it hits, an empty mask first (skipped)
        const int leader = __ffsll((unsigned long long)any) - 1;
        uint32_t key[4], ent[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t kt = (uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(tent[k] + 0.0f), leader);
            key[k] = m[k] == 0ull ? 0xFFFFFFFFu : (((m[k] >> leader) & 1ull) ? kt : 0x7F800000u);
            ent[k] = nd.entry[k];
        }
#define OMV_PSWAP(a, b)                                                       \
        if (key[b] > key[a]) {                                                \
            const uint32_t tk = key[a]; key[a] = key[b]; key[b] = tk;         \
            const uint32_t te = ent[a]; ent[a] = ent[b]; ent[b] = te;         \
            const uint64_t tm = m[a]; m[a] = m[b]; m[b] = tm;                 \
        }
        OMV_PSWAP(0, 1) OMV_PSWAP(2, 3) OMV_PSWAP(0, 2) OMV_PSWAP(1, 3) OMV_PSWAP(1, 2)
#undef OMV_PSWAP
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (m[k] == 0ull) continue;
            if (sp < OMV_STACK_TOTAL) {
                stE[sp] = ent[k];
                stM[sp] = m[k];
                ++sp;
            } else if ((m[k] >> lane) & 1ull) {
                ++c.overflow;
            }
        }
        e = ent[3];
        M = m[3];
        has = true;
    }
}

// Round-1 camera rays (computed from the slot, as WfGen) traced as wave
// packets: the hit records of omv_wf_trace<0>, sceneIntersect's alpha-mask
// re-trace loop per lane (a wave re-traces the lanes whose hit was cut out).
// Compiled for 6 waves per SIMD (<= 80 VGPRs, 28 bytes of spill): on MI355X
// 0.1 ms per render faster than 5 waves on the 2M-triangle instanced scene,
// the same on the torus knot; 8 waves spills 120 bytes inside the loop.
__global__ __launch_bounds__(OMV_BLOCK, 6) void omv_wf_trace_cam1(OmvSceneView S, WfSeg count, WfHits H,
                                                               unsigned long long* __restrict__ counters, uint32_t xcds, WfGen G) {
    __shared__ uint32_t stE[OMV_BLOCK / 64][OMV_STACK_TOTAL];
    __shared__ uint64_t stM[OMV_BLOCK / 64][OMV_STACK_TOTAL];
    __shared__ SegTable cur;
    segLoad(cur, count);
    const uint32_t wave = threadIdx.x >> 6;
    Counters cnt; cnt.nodes = 0; cnt.tris = 0; cnt.overflow = 0;
    uint32_t rays = 0, nCam = 0;
    XcdWalk w;
    // the wave stays together: lanes past the end of the queue go on with
    // the loop (inactive in every packet) until no lane has an entry
    for (w.start(cur.total(), xcds);; w.next()) {
        const bool has = w.more();
        if (__ballot(has) == 0ull) break;
        const uint32_t e = w.e;
        f3 org = mk3(0.0f, 0.0f, 0.0f), dir = mk3(0.0f, 0.0f, 0.0f);
        if (has) {
            genCameraRay(G.cam, G.seed, G.spp, G.sample0 + e, org, dir);
            ++nCam;
        }
        const bool dirOk = fabsf(dir.x) > 0.0f || fabsf(dir.y) > 0.0f || fabsf(dir.z) > 0.0f;
        PacketRay R;
        R.wo = org; R.wd = dir;
        R.widir = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
        R.tnear = 0.0f;
        R.tfar = FLT_MAX;
        R.found = false;
        R.h.t = R.tfar; R.h.u = 0.0f; R.h.v = 0.0f; R.h.inst = -1; R.h.prim = 0; R.h.tie = false;
        bool need = has, hit = false;
        int itr = 10;
        for (;;) {
            const uint64_t M = __ballot(need && dirOk);
            if (need) {
                ++rays;
                R.found = false;
                R.h.t = R.tfar; R.h.u = 0.0f; R.h.v = 0.0f; R.h.inst = -1; R.h.prim = 0; R.h.tie = false;
            }
            if (M != 0ull) packetTrace(S, R, M, stE[wave], stM[wave], cnt);
            if (need) {
                --itr;
                if (!R.found) {
                    hit = false; need = false;
                } else {
                    const int ms = maskStep(S, org, dir, R.h, R.tnear, R.tfar, FLT_MAX);
                    if (ms == 0) { hit = true; need = false; }
                    else if (ms == 2 || itr == 0) { hit = false; need = false; }
                }
            }
            if (__ballot(need) == 0ull) break;
        }
        if (has) {
            float4 r; r.x = R.h.t; r.y = R.h.u; r.z = R.h.v; r.w = R.tnear;
            int2 id; id.x = hit ? R.h.inst : -1; id.y = R.h.prim | (R.h.tie ? (int)0x80000000 : 0);
            if (hit) H.h[e] = r;   // read by the shading kernels for hits only
            H.id[e] = id;
        }
    }
    flushCounters(statSlot(counters), rays, nCam, 0u, 0u, cnt);
}

// Arguments of the refilling traversal: the queues of one launch (kKind
// 0..4 as omv_wf_trace: q or sq with its count; kKind 5: the fused launch's
// camera, connection and environment queues back to back).
struct WfTraceArgs {
    WfQueue q; WfSeg qCount;
    WfShadowQ sq; WfSeg sqCount;
    WfQueue eq; WfSeg eqCount;
    WfHits H;
    uint8_t* vis; uint8_t* nvis;
    uint32_t nslots; int lc;
    uint32_t* gstack; unsigned long long* counters; float* splat;
    uint32_t xcds;
    int keep;
};

// Refilling form of the traversal (tuning trace_keep > 0): each lane holds
// one ray and steps its traversal entry by entry; once fewer than `keep`
// lanes of the wave are busy, every idle lane takes the next ray of its own
// list (the XCD-banded walk over the launch's entries, the same hand-out as
// omv_wf_trace: no atomics, each entry once), so lanes do not idle behind
// the longest traversal of their wave.  Same per-ray arithmetic (Trav + the
// masked re-trace loop of sceneIntersect, the occlusion early-out) and the
// same outputs as omv_wf_trace / omv_wf_trace_fused.
template <int kKind>
__global__ __launch_bounds__(OMV_BLOCK, 4) void omv_wf_trace_pt(OmvSceneView S, WfTraceArgs A) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    stackInit(st, A.gstack, gtid, gthreads);
    Counters cnt; cnt.nodes = 0; cnt.tris = 0; cnt.overflow = 0;
    uint32_t rays = 0, nCam = 0, nLight = 0, nShadow = 0;
    // entries [0, nA) camera / light rays of q, [nA, nA + nB) rays of sq,
    // [nA + nB, total) environment rays of eq (by kKind)
    __shared__ SegTable cA, cB, cC;
    segLoad(cA, A.qCount); segLoad(cB, A.sqCount); segLoad(cC, A.eqCount);
    const uint32_t nA = cA.total(), nB = cB.total(), nC = cC.total();
    XcdWalk w;
    w.start(nA + nB + nC, A.xcds);
    Trav T;
    bool busy = false;
    int kind = 0, itr = 0;   // kind of the held ray: 0 q (camera / light), 2 connection / pinhole, 3 environment
    uint32_t k = 0;          // its index in its own queue
    f3 org = mk3(0.0f, 0.0f, 0.0f), dir = org, B = org;
    float t1 = 0.0f;
    for (;;) {
        // idle lanes take the next entry of their list
        if (!busy && w.more()) {
            const uint32_t e = w.e;
            w.next();
            float tn, ts;
            if (e < nA) {
                k = e;                       // logical: the hit record's index
                kind = 0;
                const uint32_t p = cA.phys(e);
                const float4 o = A.q.o[p], d = A.q.d[p];
                org = xyz(o); dir = xyz(d); tn = o.w; t1 = FLT_MAX; ts = -1.0f;
                if (kKind == 1) ++nLight; else ++nCam;
            } else if (e < nA + nB) {
                k = cB.phys(e - nA);         // physical: re-read at the end
                kind = 2;
                const float4 a = A.sq.o[k], dd = A.sq.d[k], b = A.sq.b[k];
                org = xyz(a); dir = xyz(dd); B = xyz(b); tn = a.w; t1 = dd.w;
                ts = kKind == 4 ? FLT_MAX : connectStop(org, B);
                ++nShadow;
            } else {
                k = cC.phys(e - nA - nB);
                kind = 3;
                const float4 o = A.eq.o[k], d = A.eq.d[k];
                org = xyz(o); dir = xyz(d); tn = o.w; t1 = FLT_MAX; ts = FLT_MAX;
                ++nShadow;
            }
            itr = 10;
            ++rays;
            // a zero or NaN direction leaves nothing to traverse: its
            // result (no hit) is written at the next step
            T.start(S, org, dir, tn, t1, st, cnt);
            T.tstop = (S.any_masked || S.closest_only) ? -1.0f : ts;
            busy = true;
        }
        if (__ballot(busy) == 0ull) break;
        // step until fewer than keep lanes of the wave are busy
        for (;;) {
            if (busy) {
                if (T.more(st)) T.step(S, st, cnt);
                if (!T.more(st)) {
                    // one traceClosest call finished: sceneIntersect's masked re-trace loop
                    bool done = true, hit = T.found;
                    if (hit && S.any_masked) {
                        const int mesh = S.insts[T.h.inst].mesh_id;
                        const OmvMaterial& M = S.mats[S.meshes[mesh].mat_id];
                        if (M.type == OMV_MAT_PBR && M.alpha_mode == OMV_ALPHA_MASK &&
                            isMaskedHit(S, mesh, T.h.prim, mk3(1.0f - T.h.u - T.h.v, T.h.u, T.h.v))) {
                            f3 o2 = mk3(org.x + T.h.t * dir.x, org.y + T.h.t * dir.y, org.z + T.h.t * dir.z);
                            float v = maxAbs(o2);
                            v = v > T.h.t ? v : T.h.t;
                            float delta = 3.0f * FLT_EPSILON * v;
                            const float tn = T.h.t + delta;
                            hit = false;
                            T.tnear = tn;
                            if (!(t1 <= tn) && --itr > 0) {
                                ++rays;
                                T.start(S, org, dir, tn, t1, st, cnt);
                                done = false;
                            }
                        }
                    }
                    if (done) {
                        if (kind == 2 && kKind != 4) {
                            bool visible = true;
                            if (hit) {
                                f3 df = sub3(hitPosition(S, T.h), B);
                                float sd = dot3(df, df);
                                float safeD = 10.0f * FLT_EPSILON * maxAbs(B);
                                visible = sd <= safeD * safeD;
                            }
                            const uint32_t id = __float_as_uint(A.sq.b[k].w);
                            const uint32_t slot = id / (uint32_t)A.lc, l = id - slot * (uint32_t)A.lc;
                            A.vis[(uint64_t)l * A.nslots + slot] = visible ? 2 : 0;
                        } else if (kind == 2) {
                            splatAdd(A.splat, !hit, __float_as_uint(A.sq.b[k].w), B);
                        } else if (kind == 3) {
                            A.nvis[__float_as_uint(A.eq.d[k].w)] = hit ? 0 : 2;
                        } else {
                            float4 r; r.x = T.h.t; r.y = T.h.u; r.z = T.h.v; r.w = T.tnear;
                            int2 id; id.x = hit ? T.h.inst : -1; id.y = T.h.prim | (T.h.tie ? (int)0x80000000 : 0);
                            if (hit) A.H.h[k] = r;   // read by the shading kernels for hits only
                            A.H.id[k] = id;
                        }
                        busy = false;
                    }
                }
            }
            if (__popcll(__ballot(busy)) < A.keep) break;
        }
    }
    flushCounters(statSlot(A.counters), rays, nCam, nLight, nShadow, cnt);
}

// ---------------------------------------------------------- shade light
// Light-subpath vertex i (computePath with inv = false, BDPTRenderer.cpp:149-188)
// and, with OMV_EXT_LIGHT_SPLAT (kSplat), its pinhole connection staged for
// the visibility trace.
template <bool kSplat>
__global__ __launch_bounds__(OMV_BLOCK) void omv_wf_shade_light(OmvSceneView S, omv_camera_desc cam, RP P, uint64_t sample0, int i, WfState W,
                                                                WfQueue in, WfSeg inCount, WfHits H, WfQueue out, WfSeg outCount, WfShadowQ sq,
                                                                WfSeg sqCount, unsigned long long* __restrict__ lost) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ SegTable cur;
    segLoad(cur, inCount);
    const uint32_t total = cur.total();
    const uint64_t N = W.n;
    for (uint32_t e = gtid; e < total; e += gthreads) {
        const int2 id = H.id[e];
        bool emit = false, spl = false;
        f3 nPos = mk3(0.0f, 0.0f, 0.0f), nDir = nPos, sPos = nPos, sDir = nPos, ctb = nPos;
        float nTnear = 0.0f, sTn = 0.0f, sD = 0.0f;
        uint32_t pix = 0, slot = 0;
        if (id.x >= 0) {   // a missed light ray ends its subpath: nothing to read or write
            const float4 dd = in.d[cur.phys(e)];
            slot = __float_as_uint(dd.w);
            const f3 pDir = xyz(dd);
            const float4 hr = H.h[e];
            Hit h; h.t = hr.x; h.u = hr.y; h.v = hr.z; h.inst = id.x; h.prim = id.y & 0x7FFFFFFF; h.tie = id.y < 0;
            SP sp;
            computeSurfacePoint(S, h, sp, false);
            modifyFrame(S, sp);
            Rng ls;
            rngRestore(ls, P.seed, STREAM_LIGHT, sample0 + slot, W.lsn[slot]);
            const OmvLight Lt = S.lights[W.lidx[slot]];
            const f3 lightPos = ld3(Lt.pos);
            const f3 lRad = mk3(Lt.intensity[0] * P.light_scale, Lt.intensity[1] * P.light_scale, Lt.intensity[2] * P.light_scale);
            f3 outRay, brdf, erad;
            matGetRayAndBrdf(S, P.x, sp, pDir, outRay, brdf, erad, false, ls);
            const float4 lt = i == 1 ? f4of(mk3(1.0f, 1.0f, 1.0f), 1.0f) : W.lthr[slot];
            f3 of = xyz(lt);
            float rD = lt.w;
            float ct = fabsf(dot3(pDir, sp.n));
            of = mk3(of.x * brdf.x * ct, of.y * brdf.y * ct, of.z * brdf.z * ct);
            if (i == 1) rD = attenuation(P, length3(sub3(sp.pos, lightPos)), Lt.range);
            if (kSplat) {   // every hit of the walk, before the hard cut
                f3 beta = i == 1 ? splatBeta1(P, ld3(Lt.intensity), rD, length3(sub3(sp.pos, lightPos))) : xyz(W.lbeta[slot]);
                spl = splatPrepare(S, P.x, cam, sp, pDir, beta, ctb, pix, sDir, sD, sTn);
                sPos = sp.pos;
                const float co = fabsf(dot3(outRay, sp.n));
                W.lbeta[slot] = f4of(mk3(beta.x * brdf.x * co, beta.y * brdf.y * co, beta.z * brdf.z * co), 0.0f);
            }
            f3 outRad = mk3(of.x * lRad.x * rD, of.y * lRad.y * rD, of.z * lRad.z * rD);
            if (!(intensity(outRad) < P.cut)) {
                W.lv[((uint64_t)i * 6 + 0) * N + slot] = sp.pos.x;
                W.lv[((uint64_t)i * 6 + 1) * N + slot] = sp.pos.y;
                W.lv[((uint64_t)i * 6 + 2) * N + slot] = sp.pos.z;
                W.lv[((uint64_t)i * 6 + 3) * N + slot] = outRad.x;
                W.lv[((uint64_t)i * 6 + 4) * N + slot] = outRad.y;
                W.lv[((uint64_t)i * 6 + 5) * N + slot] = outRad.z;
                W.lplen[slot] = i + 1;
                W.lthr[slot] = f4of(of, rD);
                W.lsn[slot] = rngConsumed(ls);
                if (i + 1 < W.lvCap) {
                    emit = true;
                    float delta = 3.0f * FLT_EPSILON * maxAbs(sp.pos);
                    nPos = mk3(sp.pos.x + delta * outRay.x, sp.pos.y + delta * outRay.y, sp.pos.z + delta * outRay.z);
                    nDir = outRay;
                    nTnear = hr.w;
                }
            }
        }
        const uint32_t k = segAppend(outCount, emit, lost);
        if (k != OMV_SEG_NONE) {
            out.o[k] = f4of(nPos, nTnear);
            out.d[k] = f4of(nDir, __uint_as_float(slot));
        }
        if (kSplat) {
            const uint32_t ks = segAppend(sqCount, spl, lost);
            if (ks != OMV_SEG_NONE) {
                sq.o[ks] = f4of(sPos, sTn);
                sq.d[ks] = f4of(sDir, sD);
                sq.b[ks] = f4of(ctb, __uint_as_float(pix));
            }
        }
    }
}

// ---------------------------------------------------------- shade camera
// Camera vertex i (computePath with inv = true, BDPTRenderer.cpp:149-176),
// the connections of that vertex staged for the visibility trace
// (BDPTRenderer.cpp:266-297), the escape term (:251-258) and, with
// OMV_EXT_HDRI_NEE, the next-event estimate staged for its visibility trace.
template <bool kLights, bool kNee>
__global__ __launch_bounds__(OMV_BLOCK, OMV_SHADE_WAVES(kLights, kNee)) void omv_wf_shade_cam(OmvSceneView S, RP P, uint64_t sample0, int i, WfState W, WfQueue in,
                                                              WfSeg inCount, WfHits H, WfQueue out, WfSeg outCount, WfShadowQ sq,
                                                              WfSeg sqCount, WfQueue eq, WfSeg eqCount, float* __restrict__ rad, WfGen G,
                                                              unsigned long long* __restrict__ lost) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ SegTable cur;
    segLoad(cur, inCount);
    const uint32_t total = cur.total();
    const uint64_t N = W.n;
    const int q = P.q;
    const bool gen = G.gen && i == 1;
    for (uint32_t e = gtid; e < total; e += gthreads) {
        uint32_t slot;
        f3 pDir;
        if (gen) {   // round 1: the camera ray of slot e
            f3 org;
            slot = e;
            genCameraRay(G.cam, G.seed, G.spp, G.sample0 + e, org, pDir);
        } else {
            const float4 dd = in.d[cur.phys(e)];
            slot = __float_as_uint(dd.w);
            pDir = xyz(dd);
        }
        const int2 id = H.id[e];
        const bool first = i == 1;
        const float4 rc = first ? f4of(mk3(0.0f, 0.0f, 0.0f), __int_as_float(1)) : W.radc[slot];
        f3 radiance = xyz(rc);
        int n = __float_as_int(rc.w);
        const float4 th = first ? f4of(mk3(1.0f, 1.0f, 1.0f), 1.0f) : W.thr[slot];
        f3 of = xyz(th);
        bool emit = false, envRay = false;
        f3 eOrg = mk3(0.0f, 0.0f, 0.0f), eDir = eOrg;
        f3 nPos = mk3(0.0f, 0.0f, 0.0f), nDir = nPos;
        float nTnear = 0.0f;
        if (id.x < 0) {
            // escape (BDPTRenderer.cpp:251-258); no connections this round
            // (the path is written here, so the connect kernel reads nothing else of it)
            if (S.bg_type != OMV_BG_NONE && n <= q) {
                f3 bg = background(S, pDir);
                bg = mk3(bg.x * of.x, bg.y * of.y, bg.z * of.z);
                if (kNee) bg = mk3(bg.x * th.w, bg.y * th.w, bg.z * th.w);
                radiance = mk3(bg.x * 1.0f + radiance.x, bg.y * 1.0f + radiance.y, bg.z * 1.0f + radiance.z);
            }
            // the connection sum exists with point lights only (0 before round 1's connect)
            const f3 conn = (kLights && !first) ? xyz(W.conn[slot]) : mk3(0.0f, 0.0f, 0.0f);
            rad[(uint64_t)slot * 3 + 0] = conn.x * 1.0f + radiance.x;
            rad[(uint64_t)slot * 3 + 1] = conn.y * 1.0f + radiance.y;
            rad[(uint64_t)slot * 3 + 2] = conn.z * 1.0f + radiance.z;
            W.flags[slot] = 2u;
        } else {
            const float4 hr = H.h[e];
            Hit h; h.t = hr.x; h.u = hr.y; h.v = hr.z; h.inst = id.x; h.prim = id.y & 0x7FFFFFFF; h.tie = id.y < 0;
            SP sp;
            computeSurfacePoint(S, h, sp, false);
            modifyFrame(S, sp);
            MatIn mi;   // PBR texture inputs, read once for the connections and the bounce
            if (kLights) matInputs(S, sp, mi);
            if (kLights) {
                const int li = W.lidx[slot];
                const OmvLight Lt = S.lights[li];
                const f3 lightInt = ld3(Lt.intensity);
                const int lEnd = min(S.lights_leave ? 1 : W.lplen[slot], q - i + 1);
                const f3 outV = neg3(pDir);
                W.nconn[slot] = (uint32_t)max(lEnd, 0);
                for (int l = 0; l < lEnd; ++l) {
                    f3 lPos = l == 0 ? ld3(Lt.pos)
                                     : mk3(W.lv[((uint64_t)l * 6 + 0) * N + slot], W.lv[((uint64_t)l * 6 + 1) * N + slot],
                                           W.lv[((uint64_t)l * 6 + 2) * N + slot]);
                    f3 lRad;
                    f3 LC = normalize3(sub3(sp.pos, lPos));
                    if (l == 0) {
                        float lsc = attenuation(P, length3(sub3(sp.pos, lPos)), Lt.range);
                        float k = P.light_scale * lsc;
                        lRad = mk3(lightInt.x * k, lightInt.y * k, lightInt.z * k);
                    } else {
                        lRad = mk3(W.lv[((uint64_t)l * 6 + 3) * N + slot], W.lv[((uint64_t)l * 6 + 4) * N + slot],
                                   W.lv[((uint64_t)l * 6 + 5) * N + slot]);
                    }
                    float c = fabsf(dot3(LC, sp.n));
                    f3 lc = mk3(lRad.x * c * of.x, lRad.y * c * of.y, lRad.z * c * of.z);
                    uint8_t v = 0;
                    bool ray = false;
                    f3 A = lPos, B = sp.pos, dir = mk3(0.0f, 0.0f, 0.0f);
                    float tnear = 0.0f, tfar = 0.0f;
                    if (!(intensity(lc) < P.cut)) {
                        f3 brdf = matGetBrdf(S, P.x, sp, LC, outV, &mi);
                        brdf = mk3(brdf.x * c, brdf.y * c, brdf.z * c);
                        lc = mk3(brdf.x * of.x * lRad.x, brdf.y * of.y * lRad.y, brdf.z * of.z * lRad.z);
                        if (!(intensity(lc) < P.cut)) {
                            // isConnected(A = light vertex, B = camera vertex) up to the ray cast
                            dir = sub3(B, A);
                            float len = length3(dir);
                            float safeD = 2.0f * 3.0f * FLT_EPSILON * maxAbs(A);
                            if (len <= safeD) {
                                v = 2;
                            } else {
                                dir = mk3(dir.x / len, dir.y / len, dir.z / len);
                                A = mk3(A.x + 0.0f * dir.x, A.y + 0.0f * dir.y, A.z + 0.0f * dir.z);
                                tnear = safeD;
                                tfar = len + (0.1f * len > safeD ? 0.1f * len : safeD);
                                v = 1;
                                ray = true;
                            }
                            W.cc[((uint64_t)l * 3 + 0) * N + slot] = lc.x;
                            W.cc[((uint64_t)l * 3 + 1) * N + slot] = lc.y;
                            W.cc[((uint64_t)l * 3 + 2) * N + slot] = lc.z;
                        }
                    }
                    // connections to the light itself grouped by light (key = light index),
                    // to light subpath vertices after them (key = lights)
                    const uint32_t nk = S.num_lights < 8 ? (uint32_t)S.num_lights + 1u : 1u;
                    const uint32_t key = nk > 1u ? (l == 0 ? (uint32_t)li : (uint32_t)S.num_lights) : 0u;
                    const uint32_t k = segAppendKeyed(sqCount, ray, key, nk, lost);
                    if (ray && k == OMV_SEG_NONE) v = 0;   // a full region (cannot happen with the host's sizing)
                    W.vis[(uint64_t)l * N + slot] = v;
                    if (k != OMV_SEG_NONE) {
                        sq.o[k] = f4of(A, tnear);
                        sq.d[k] = f4of(dir, tfar);
                        sq.b[k] = f4of(B, __uint_as_float(slot * (uint32_t)W.lc + (uint32_t)l));
                    }
                }
            }
            Rng cs;
            rngRestore(cs, P.seed, STREAM_CAMERA, sample0 + slot, first ? 2u : W.csn[slot]);   // 2: the camera jitter's draws
            f3 outRay, brdf, erad;
            BounceInfo bi;
            matGetRayAndBrdf(S, P.x, sp, pDir, outRay, brdf, erad, true, cs, kNee ? &bi : nullptr, kLights ? &mi : nullptr);
            float ct = fabsf(dot3(outRay, sp.n));
            const f3 ofPrev = of;
            of = mk3(of.x * brdf.x * ct, of.y * brdf.y * ct, of.z * brdf.z * ct);
            radiance = mk3(radiance.x + of.x * erad.x, radiance.y + of.y * erad.y, radiance.z + of.z * erad.z);
            float escW = 1.0f;
            if (kNee) {
                f3 c;
                envRay = neePrepare(S, P, sp, bi, ofPrev, outRay, i, sample0 + slot, escW, c, eOrg, eDir);
                if (envRay) W.neeC[slot] = f4of(c, 0.0f);
                W.nvis[slot] = envRay ? 1 : 0;
            }
            ++n;
            W.csn[slot] = rngConsumed(cs);
            W.thr[slot] = f4of(of, escW);
            W.radc[slot] = f4of(radiance, __int_as_float(n));
            // continue unless the loop ends (i == B) or an exact shortcut stops it
            const bool zero = of.x == 0.0f && of.y == 0.0f && of.z == 0.0f;
            const bool cont = i < P.camera_bounces && !zero && (S.any_emissive || n <= q);
            if (cont) {
                emit = true;
                float delta = 3.0f * FLT_EPSILON * maxAbs(sp.pos);
                nPos = mk3(sp.pos.x + delta * outRay.x, sp.pos.y + delta * outRay.y, sp.pos.z + delta * outRay.z);
                nDir = outRay;
                nTnear = hr.w;
            } else if (kLights || kNee) {
                W.flags[slot] = 1u;   // written by the connect kernel of this round
            } else {
                const f3 conn = mk3(0.0f, 0.0f, 0.0f);   // no point lights: no connections
                rad[(uint64_t)slot * 3 + 0] = conn.x * 1.0f + radiance.x;
                rad[(uint64_t)slot * 3 + 1] = conn.y * 1.0f + radiance.y;
                rad[(uint64_t)slot * 3 + 2] = conn.z * 1.0f + radiance.z;
                W.flags[slot] = 2u;
            }
        }
        const uint32_t k = segAppend(outCount, emit, lost);
        if (k != OMV_SEG_NONE) {
            out.o[k] = f4of(nPos, nTnear);
            out.d[k] = f4of(nDir, __uint_as_float(slot));
        }
        if (kNee) {
            const uint32_t ke = segAppend(eqCount, envRay, lost);
            if (envRay && ke == OMV_SEG_NONE) W.nvis[slot] = 0;
            if (ke != OMV_SEG_NONE) {
                eq.o[ke] = f4of(eOrg, 0.0f);
                eq.d[ke] = f4of(eDir, __uint_as_float(slot));
            }
        }
    }
}

// -------------------------------------------------------------- connect
// Adds the visible connections of this round in light-vertex order and the
// visible next-event estimate (after the vertex's emission, as the
// per-sample function does), and writes the sample of paths that ended
// this round.
__global__ __launch_bounds__(OMV_BLOCK) void omv_wf_connect(WfState W, WfQueue in, WfSeg inCount, float* __restrict__ rad, int lights,
                                                            int nee, int round, int gen) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ SegTable cur;
    segLoad(cur, inCount);
    const uint32_t total = cur.total();
    const uint64_t N = W.n;
    const bool slotIsEntry = gen && round == 1;   // round-1 rays computed from the slot: entry e is slot e
    for (uint32_t e = gtid; e < total; e += gthreads) {
        const uint32_t slot = slotIsEntry ? e : __float_as_uint(in.d[cur.phys(e)].w);
        const uint32_t fl = W.flags[slot];
        if (fl == 2u) continue;
        f3 conn = (lights && round > 1) ? xyz(W.conn[slot]) : mk3(0.0f, 0.0f, 0.0f);
        if (lights) {
            const uint32_t nc = W.nconn[slot];
            for (uint32_t l = 0; l < nc; ++l) {
                if (W.vis[(uint64_t)l * N + slot] != 2) continue;
                f3 lc = mk3(W.cc[((uint64_t)l * 3 + 0) * N + slot], W.cc[((uint64_t)l * 3 + 1) * N + slot],
                            W.cc[((uint64_t)l * 3 + 2) * N + slot]);
                conn = mk3(lc.x + conn.x, lc.y + conn.y, lc.z + conn.z);
            }
            W.conn[slot] = f4of(conn, 0.0f);
        }
        // the radiance so far is needed by a visible next-event estimate and
        // by a path that ends this round
        const bool neeVis = nee && W.nvis[slot] == 2;
        if (!neeVis && fl != 1u) continue;
        float4 rc = W.radc[slot];
        if (neeVis) {
            const f3 c = xyz(W.neeC[slot]);
            rc.x = rc.x + c.x; rc.y = rc.y + c.y; rc.z = rc.z + c.z;
            W.radc[slot] = rc;
        }
        if (fl == 1u) {
            const f3 radiance = xyz(rc);
            rad[(uint64_t)slot * 3 + 0] = conn.x * 1.0f + radiance.x;
            rad[(uint64_t)slot * 3 + 1] = conn.y * 1.0f + radiance.y;
            rad[(uint64_t)slot * 3 + 2] = conn.z * 1.0f + radiance.z;
            W.flags[slot] = 2u;
        }
    }
}

// Paths that end when the camera loop runs out (i == B) are written by
// shade/connect above; nothing is left over after the last round.

// ----------------------------------------------------------------- tail
// Camera rounds i0 .. B of the paths still alive after round i0 - 1, in one
// launch: each lane takes one path of the round-i0 queue to its end with the
// per-sample function's camera loop (cameraRounds: trace, shade, its
// connections and next-event ray traced in place), from the state the
// wavefront kernels keep for it (ray in the queue, throughput / escape
// weight, radiance and vertex count, connection sum, camera stream position,
// light subpath), and writes the path's sample.  The late rounds hold few
// rays and cost one launch latency each for trace, shade and connect; here
// they cost the longest remaining path once.  Sums are formed in the same
// order, so the film is unchanged.
template <bool kLights, bool kNee>
__global__ __launch_bounds__(OMV_BLOCK, OMV_TAIL_WAVES(kLights, kNee)) void omv_wf_tail(OmvSceneView S, RP P, uint64_t sample0, int i0, WfState W, WfQueue in,
                                                         WfSeg inCount, float* __restrict__ rad, uint32_t* __restrict__ gstack,
                                                         unsigned long long* __restrict__ counters) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Lane L;
    laneInit(L, gstack, nullptr, gtid, gthreads);
    __shared__ SegTable cur;
    segLoad(cur, inCount);
    const uint32_t total = cur.total();
    const uint64_t N = W.n;
    for (uint32_t e = gtid; e < total; e += gthreads) {
        const uint32_t p = cur.phys(e);
        const float4 o = in.o[p], d = in.d[p];
        const uint32_t slot = __float_as_uint(d.w);
        const float4 rc = W.radc[slot], th = W.thr[slot];
        LightSub ls;
        ls.idx = -1; ls.len = 0; ls.pos = mk3(0.0f, 0.0f, 0.0f); ls.intensity = ls.pos; ls.range = FLT_MAX;
        f3 conn = mk3(0.0f, 0.0f, 0.0f);
        if (kLights) {
            ls.idx = W.lidx[slot];
            ls.len = S.lights_leave ? 1 : W.lplen[slot];
            const OmvLight Lt = S.lights[ls.idx];
            ls.pos = ld3(Lt.pos); ls.intensity = ld3(Lt.intensity); ls.range = Lt.range;
            conn = xyz(W.conn[slot]);
        }
        WfLightVerts lv;
        lv.lv = W.lv; lv.N = N; lv.slot = slot;
        Rng cs;
        rngRestore(cs, P.seed, STREAM_CAMERA, sample0 + slot, W.csn[slot]);
        const f3 r = cameraRounds<kLights, kNee>(S, P, sample0 + slot, i0, xyz(o), xyz(d), o.w, xyz(th), th.w, xyz(rc), conn,
                                                 __float_as_int(rc.w), cs, ls, lv, L);
        rad[(uint64_t)slot * 3 + 0] = r.x;
        rad[(uint64_t)slot * 3 + 1] = r.y;
        rad[(uint64_t)slot * 3 + 2] = r.z;
    }
    flushCounters(statSlot(counters), L.rays, L.camRays, 0u, L.shadowRays, L.cnt);
}

// ----------------------------------------------------------- ray sorting
// A round's secondary camera rays reordered by (direction octant, cell of a
// 16^3 grid over the scene box, Morton order) with a counting sort, so that
// a traversal wave holds rays that start close together and head the same
// way.  Order within a bucket is arbitrary; every ray's result is its own.
#define OMV_SORT_BUCKETS 32768
DEV uint32_t spread4(uint32_t v) {   // 4 bits -> every third bit
    v &= 0xFu;
    v = (v | (v << 4)) & 0x0C3u;
    v = (v | (v << 2)) & 0x249u;
    return v;
}
DEV uint32_t rayKey(const OmvSceneView& S, const float4& o, const float4& d) {
    const uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    const float p[3] = {o.x, o.y, o.z};
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int q = (int)((p[a] - S.scene_lo[a]) * S.scene_inv[a] * 16.0f);
        c[a] = (uint32_t)min(max(q, 0), 15);
    }
    return (oct << 12) | (spread4(c[0]) << 2) | (spread4(c[1]) << 1) | spread4(c[2]);
}
__global__ __launch_bounds__(OMV_BLOCK) void omv_wf_sort_keys(OmvSceneView S, WfQueue q, const uint32_t* __restrict__ count,
                                                              uint32_t* __restrict__ keys, uint32_t* __restrict__ hist) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t total = *count;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gthreads) {
        const uint32_t k = rayKey(S, q.o[e], q.d[e]);
        keys[e] = k;
        atomicAdd(hist + k, 1u);
    }
}
// exclusive offsets of the bucket counts, in place: one block of 1024 lanes,
// 32 consecutive buckets each
__global__ __launch_bounds__(1024) void omv_wf_sort_scan(uint32_t* __restrict__ hist) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    uint32_t v[OMV_SORT_BUCKETS / 1024];
    uint32_t s = 0u;
#pragma unroll
    for (int k = 0; k < OMV_SORT_BUCKETS / 1024; ++k) { v[k] = hist[t * (OMV_SORT_BUCKETS / 1024) + k]; s += v[k]; }
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t x = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t base = part[t] - s;
#pragma unroll
    for (int k = 0; k < OMV_SORT_BUCKETS / 1024; ++k) { hist[t * (OMV_SORT_BUCKETS / 1024) + k] = base; base += v[k]; }
}
__global__ __launch_bounds__(OMV_BLOCK) void omv_wf_sort_scatter(WfQueue q, const uint32_t* __restrict__ count,
                                                                 const uint32_t* __restrict__ keys, uint32_t* __restrict__ offs,
                                                                 WfQueue out) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t total = *count;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gthreads) {
        const uint32_t pos = atomicAdd(offs + keys[e], 1u);
        out.o[pos] = q.o[e];
        out.d[pos] = q.d[e];
    }
}

__global__ void omv_wf_fill_u32(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// the OMV_STAT_SLOTS stat slots of a chunk added into the renderer's counters
__global__ void omv_wf_stats_sum(const unsigned long long* __restrict__ slots, unsigned long long* __restrict__ out) {
    const int t = threadIdx.x;
    if (t < 7) {
        unsigned long long v = 0;
        for (int k = 0; k < OMV_STAT_SLOTS; ++k) v += slots[k * 16 + t];
        if (v) atomicAdd(out + t, v);
    }
}

// ------------------------------------------------------------ launchers
namespace omv {

namespace {
template <class K>
int occ(K k, size_t lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, OMV_BLOCK, lds) != hipSuccess || n < 1) n = 1;
    return n;
}
constexpr size_t kAlign = 256;
size_t up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }
}  // namespace

// Entries per append region of a queue fed by a launch of `grid`
// workgroups over at most nslots input entries with one append each: the
// workgroups of one region (ceil(grid / K)) handle at most OMV_BLOCK *
// ceil(nslots / threads) entries each.  The producers are the shading
// kernels (gridShade) and, for the round-1 light queue, generation
// (gridGen); regions are sized for the larger of the two.  A queue holds
// OMV_QSEG regions (>= nslots entries in all).
static uint32_t segCapFor(uint32_t nslots, int grid) {
    const uint64_t g = (uint64_t)std::max(grid, 1), threads = g * OMV_BLOCK;
    const uint64_t perBlock = (uint64_t)OMV_BLOCK * ((nslots + threads - 1) / threads);
    return (uint32_t)(((g + OMV_QSEG - 1) / OMV_QSEG) * perBlock);
}
uint32_t wavefrontSegCap(uint32_t nslots, int gridShade, int gridGen) {
    return std::max(segCapFor(nslots, gridShade), segCapFor(nslots, gridGen));
}

size_t wavefrontArenaBytes(uint32_t nslots, int lvCap, int lc, bool lights, bool nee, bool splat, bool sortRays, int gridShade, int gridGen) {
    const size_t N = nslots;
    const size_t Q = (size_t)OMV_QSEG * wavefrontSegCap(nslots, gridShade, gridGen);   // queue entries
    size_t b = 6 * up(Q * 16)                        // three ray queues
               + up(N * 16) + up(N * 8)              // hits
               + 3 * up(N * 16) + 3 * up(N * 4);     // thr, radc, conn, csn, flags, nconn
    if (lights)
        b += up(N * 16) + 3 * up(N * 4) + up(N * (size_t)lvCap * 24) + up(N * (size_t)lc * 12) + up(N * (size_t)lc) +
             3 * up(Q * (size_t)lc * 16);
    if (nee) b += up(N * 16) + up(N) + 2 * up(Q * 16);   // neeC, nvis, environment-ray queue
    if (lights && splat) b += up(N * 16);                 // lbeta
    if (sortRays) b += up(N * 4) + up(OMV_SORT_BUCKETS * 4);   // keys, bucket counts / offsets
    b += up(OMV_STAT_SLOTS * 16 * 8);                     // traversal stat slots
    return b;
}

// lightCnt, camCnt, shCnt, envCnt, splCnt, OMV_QSEG counters each (see launchWavefront)
int wavefrontCounterCount(int B, int lvCap) { return OMV_QSEG * (2 * (lvCap + 2) + 3 * (B + 2)); }

void wavefrontBlocksPerCU(int out[9]) {
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    int t = occ(omv_wf_trace<0>, lds);
    for (int k : {occ(omv_wf_trace<1>, lds), occ(omv_wf_trace<2>, lds), occ(omv_wf_trace<3>, lds), occ(omv_wf_trace<4>, lds),
                  occ(omv_wf_trace_fused, lds)})
        t = std::min(t, k);
    int tp = occ(omv_wf_trace_pt<0>, lds);
    for (int k : {occ(omv_wf_trace_pt<1>, lds), occ(omv_wf_trace_pt<2>, lds), occ(omv_wf_trace_pt<3>, lds), occ(omv_wf_trace_pt<4>, lds),
                  occ(omv_wf_trace_pt<5>, lds)})
        tp = std::min(tp, k);
    out[0] = t;
    out[4] = tp;
    out[1] = occ(omv_wf_shade_cam<true, true>, 0);
    out[5] = occ(omv_wf_shade_cam<true, false>, 0);
    out[6] = occ(omv_wf_shade_cam<false, true>, 0);
    out[7] = occ(omv_wf_shade_cam<false, false>, 0);
    out[2] = occ(omv_wf_shade_light<true>, 0);
    out[3] = occ(omv_wf_gen<true>, 0);
    out[8] = occ(omv_wf_trace_cam1, 0);
}

// One chunk of whole pixels: samples [sample0, sample0 + nslots).
hipError_t launchWavefront(const WfLaunch& L, hipStream_t stream) {
    const OmvSceneView& S = L.S;
    RP P = toRP(L.p);
    const bool lights = S.num_lights > 0;
    if (!S.env_tex_pdf) P.x.nee = false;   // no HDR environment: nothing to sample
    const bool nee = P.x.nee;
    if (!lights || !L.splat) P.x.splat = false;
    const bool splat = P.x.splat;
    const uint32_t n = L.nslots;
    const size_t N = n;
    // queues: OMV_QSEG append regions of segCap entries
    const uint32_t segCap = wavefrontSegCap(n, L.blocksShade, L.blocksGen);
    const size_t Q = (size_t)OMV_QSEG * segCap;
    uint8_t* cur = static_cast<uint8_t*>(L.arena);
    auto take = [&](size_t bytes) { uint8_t* p = cur; cur += up(bytes); return p; };
    auto queue = [&]() { WfQueue q; q.o = reinterpret_cast<float4*>(take(Q * 16)); q.d = reinterpret_cast<float4*>(take(Q * 16)); return q; };
    WfQueue qa = queue(), qb = queue(), qc = queue();
    WfHits H;
    H.h = reinterpret_cast<float4*>(take(N * 16));
    H.id = reinterpret_cast<int2*>(take(N * 8));
    WfState W;
    std::memset(&W, 0, sizeof(W));
    W.n = n; W.lvCap = L.lvCap; W.lc = L.lc;
    W.thr = reinterpret_cast<float4*>(take(N * 16));
    W.radc = reinterpret_cast<float4*>(take(N * 16));
    W.conn = reinterpret_cast<float4*>(take(N * 16));
    W.csn = reinterpret_cast<uint32_t*>(take(N * 4));
    W.flags = reinterpret_cast<uint32_t*>(take(N * 4));
    W.nconn = reinterpret_cast<uint32_t*>(take(N * 4));
    WfShadowQ sq;
    std::memset(&sq, 0, sizeof(sq));
    if (lights) {
        W.lthr = reinterpret_cast<float4*>(take(N * 16));
        W.lsn = reinterpret_cast<uint32_t*>(take(N * 4));
        W.lidx = reinterpret_cast<int*>(take(N * 4));
        W.lplen = reinterpret_cast<int*>(take(N * 4));
        W.lv = reinterpret_cast<float*>(take(N * (size_t)L.lvCap * 24));
        W.cc = reinterpret_cast<float*>(take(N * (size_t)L.lc * 12));
        W.vis = reinterpret_cast<uint8_t*>(take(N * (size_t)L.lc));
        sq.o = reinterpret_cast<float4*>(take(Q * (size_t)L.lc * 16));
        sq.d = reinterpret_cast<float4*>(take(Q * (size_t)L.lc * 16));
        sq.b = reinterpret_cast<float4*>(take(Q * (size_t)L.lc * 16));
    }
    if (splat) W.lbeta = reinterpret_cast<float4*>(take(N * 16));
    uint32_t* sortKeys = nullptr;
    uint32_t* sortHist = nullptr;
    unsigned long long* stats = reinterpret_cast<unsigned long long*>(take(OMV_STAT_SLOTS * 16 * 8));
    if (L.sortRays) {
        sortKeys = reinterpret_cast<uint32_t*>(take(N * 4));
        sortHist = reinterpret_cast<uint32_t*>(take(OMV_SORT_BUCKETS * 4));
    }
    WfQueue eq;
    std::memset(&eq, 0, sizeof(eq));
    if (nee) {
        W.neeC = reinterpret_cast<float4*>(take(N * 16));
        W.nvis = reinterpret_cast<uint8_t*>(take(N));
        eq = queue();
    }
    // queue counters: light round i reads lightCnt[i] (i = 1 .. lvCap-1) and
    // appends to lightCnt[i+1]; camera round i reads camCnt[i] (i = 1 .. B),
    // appends to camCnt[i+1], stages connection rays in shCnt[i] and
    // environment rays in envCnt[i]; light round i stages pinhole rays in
    // splCnt[i] (the connection-ray queue is free during the light rounds)
    const int B = L.p.camera_bounces;
    const int ncnt = wavefrontCounterCount(B, L.lvCap);
    const uint32_t K = OMV_QSEG;
    uint32_t* lightBase = L.qcount;
    uint32_t* camBase = lightBase + K * (L.lvCap + 2);
    uint32_t* shBase = camBase + K * (B + 2);
    uint32_t* envBase = shBase + K * (B + 2);
    uint32_t* splBase = envBase + K * (B + 2);
    // Segmented counters of round r's queues.  The round-1 queues (written by
    // gen at the slot's index) and camera queues that are reordered by the
    // ray sort are compact (one region).
    const uint32_t camK = L.sortRays ? 1u : K, camCap = L.sortRays ? (uint32_t)Q : segCap;
    auto lightCnt = [&](int r) { return segOf(lightBase + K * r, K, segCap); };
    auto camCnt = [&](int r) { return r == 1 ? segOf(camBase + K, 1u, (uint32_t)Q) : segOf(camBase + K * r, camK, camCap); };
    auto shCnt = [&](int r) { return segOf(shBase + K * r, K, segCap * (uint32_t)L.lc); };
    auto envCnt = [&](int r) { return segOf(envBase + K * r, K, segCap); };
    auto splCnt = [&](int r) { return segOf(splBase + K * r, K, segCap); };
    WfSeg none;
    std::memset(&none, 0, sizeof(none));
    unsigned long long* lost = L.counters + 6;   // the stats' overflow counter
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    const dim3 blk(OMV_BLOCK);
    const uint32_t xcds = L.xcdBands > 1 ? (uint32_t)L.xcdBands : 1u;
    // refilling traversal (trace_keep > 0): one argument block per launch
    auto ptArgs = [&](const WfQueue& q, const WfSeg& qc, const WfSeg& sqc, const WfQueue& e, const WfSeg& ec) {
        WfTraceArgs a;
        a.q = q; a.qCount = qc; a.sq = sq; a.sqCount = sqc; a.eq = e; a.eqCount = ec; a.H = H;
        a.vis = W.vis; a.nvis = W.nvis; a.nslots = n; a.lc = L.lc;
        a.gstack = L.gstack; a.counters = stats; a.splat = L.splat; a.xcds = xcds; a.keep = L.traceKeep;
        return a;
    };
    auto mark = [&]() {
        if (L.traceEv && L.traceEvUsed && *L.traceEvUsed < L.traceEvCap) (void)hipEventRecord(L.traceEv[(*L.traceEvUsed)++], stream);
    };
    hipLaunchKernelGGL(omv_wf_fill_u32, dim3((ncnt + 63) / 64), dim3(64), 0, stream, L.qcount, ncnt);
    hipLaunchKernelGGL(omv_wf_fill_u32, dim3(OMV_STAT_SLOTS * 32 / 64), dim3(64), 0, stream, reinterpret_cast<uint32_t*>(stats),
                       OMV_STAT_SLOTS * 32);
    // Round-1 camera rays: computed by the round-1 camera kernels from the
    // slot (Gen1), or with the refilling traversal stored by gen in qa;
    // round-1 light rays are stored in qb
    WfGen Gen1;
    std::memset(&Gen1, 0, sizeof(Gen1));
    Gen1.gen = L.traceKeep > 0 ? 0 : 1;
    Gen1.cam = L.cam; Gen1.seed = P.seed; Gen1.spp = L.p.samples_per_pixel; Gen1.sample0 = L.sample0;
    WfGen Gen0 = Gen1;
    Gen0.gen = 0;
    const int writeQueues = Gen1.gen ? 0 : 1;
    // with lights, gen appends the round-1 light rays (the queue regions are
    // sized for its grid too: wavefrontSegCap)
    if (lights)
        hipLaunchKernelGGL(omv_wf_gen<true>, dim3(L.blocksGen), blk, 0, stream, S, L.cam, P, L.p.samples_per_pixel, L.sample0, W, qa, qb,
                           camCnt(1).cnt, lightCnt(1), L.rad, writeQueues, stats, lost);
    else
        hipLaunchKernelGGL(omv_wf_gen<false>, dim3(L.blocksGen), blk, 0, stream, S, L.cam, P, L.p.samples_per_pixel, L.sample0, W, qa, qb,
                           camCnt(1).cnt, lightCnt(1), L.rad, writeQueues, stats, lost);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // with S.lights_leave generation queues no light ray, and the light
    // rounds would only find empty queues
    if (lights && !S.lights_leave) {
        WfQueue in = qb, out = qc;
        for (int i = 1; i < L.lvCap; ++i) {
            mark();
            if (L.traceKeep > 0)
                hipLaunchKernelGGL(omv_wf_trace_pt<1>, dim3(L.blocksTrace), blk, lds, stream, S,
                                   ptArgs(in, lightCnt(i), none, eq, none));
            else
                hipLaunchKernelGGL(omv_wf_trace<1>, dim3(L.blocksTrace), blk, lds, stream, S, in, sq, lightCnt(i), H, W.vis, n, L.lc, L.gstack,
                                   stats, L.splat, xcds, Gen0);
            mark();
            if (splat) {
                hipLaunchKernelGGL(omv_wf_shade_light<true>, dim3(L.blocksShade), blk, 0, stream, S, L.cam, P, L.sample0, i, W, in, lightCnt(i),
                                   H, out, lightCnt(i + 1), sq, splCnt(i), lost);
                mark();
                if (L.traceKeep > 0)
                    hipLaunchKernelGGL(omv_wf_trace_pt<4>, dim3(L.blocksTrace), blk, lds, stream, S,
                                       ptArgs(in, none, splCnt(i), eq, none));
                else
                    hipLaunchKernelGGL(omv_wf_trace<4>, dim3(L.blocksTrace), blk, lds, stream, S, in, sq, splCnt(i), H, W.vis, n, L.lc, L.gstack,
                                       stats, L.splat, xcds, Gen0);
                mark();
            } else {
                hipLaunchKernelGGL(omv_wf_shade_light<false>, dim3(L.blocksShade), blk, 0, stream, S, L.cam, P, L.sample0, i, W, in, lightCnt(i),
                                   H, out, lightCnt(i + 1), sq, splCnt(i), lost);
            }
            std::swap(in, out);
        }
    }
    WfQueue in = qa, out = qb, spare = qc;   // qc is free once the light rounds are done
    // With connection or environment rays (and the one-pass traversal), the
    // traversal of bounce i's visibility rays is fused with that of bounce
    // i + 1's camera rays: shade(i) -> fused trace -> connect(i) -> shade(i + 1).
    // connect(i) still runs before shade(i + 1) (which reads the connection
    // sum of escaping paths and overwrites the staged connections), so every
    // sum is formed in the same order as before.
    const bool fuse = lights || nee;
    // rounds i0 .. B in one launch (0: none); the round before i0 leaves
    // its camera rays untraced for it
    const int i0 = L.tailRound >= 2 && L.tailRound <= B ? L.tailRound : 0;
    bool traced = false;   // the camera rays of round i were traced by the previous fused launch
    bool roundEvDone = false;
    for (int i = 1; i <= B; ++i) {
        if (i == i0) {
            mark();
            if (lights) {
                if (nee)
                    hipLaunchKernelGGL((omv_wf_tail<true, true>), dim3(L.blocksTrace), blk, lds, stream, S, P, L.sample0, i, W, in, camCnt(i),
                                       L.rad, L.gstack, stats);
                else
                    hipLaunchKernelGGL((omv_wf_tail<true, false>), dim3(L.blocksTrace), blk, lds, stream, S, P, L.sample0, i, W, in, camCnt(i),
                                       L.rad, L.gstack, stats);
            } else {
                if (nee)
                    hipLaunchKernelGGL((omv_wf_tail<false, true>), dim3(L.blocksTrace), blk, lds, stream, S, P, L.sample0, i, W, in, camCnt(i),
                                       L.rad, L.gstack, stats);
                else
                    hipLaunchKernelGGL((omv_wf_tail<false, false>), dim3(L.blocksTrace), blk, lds, stream, S, P, L.sample0, i, W, in, camCnt(i),
                                       L.rad, L.gstack, stats);
            }
            mark();
            break;
        }
        if (!traced) {
            mark();
            if (L.traceKeep > 0)
                hipLaunchKernelGGL(omv_wf_trace_pt<0>, dim3(L.blocksTrace), blk, lds, stream, S,
                                   ptArgs(in, camCnt(i), none, eq, none));
            else if (i == 1 && Gen1.gen && L.blocksPacket > 0)
                hipLaunchKernelGGL(omv_wf_trace_cam1, dim3(L.blocksPacket), blk, 0, stream, S, camCnt(1), H, stats, xcds, Gen1);
            else
                hipLaunchKernelGGL(omv_wf_trace<0>, dim3(L.blocksTrace), blk, lds, stream, S, in, sq, camCnt(i), H, W.vis, n, L.lc,
                                   L.gstack, stats, L.splat, xcds, i == 1 ? Gen1 : Gen0);
            mark();
        }
        traced = false;
        if (lights) {
            if (nee)
                hipLaunchKernelGGL((omv_wf_shade_cam<true, true>), dim3(L.blocksShade), blk, 0, stream, S, P, L.sample0, i, W, in, camCnt(i), H,
                                   out, camCnt(i + 1), sq, shCnt(i), eq, envCnt(i), L.rad, Gen1, lost);
            else
                hipLaunchKernelGGL((omv_wf_shade_cam<true, false>), dim3(L.blocksShade), blk, 0, stream, S, P, L.sample0, i, W, in, camCnt(i),
                                   H, out, camCnt(i + 1), sq, shCnt(i), eq, envCnt(i), L.rad, Gen1, lost);
        } else {
            if (nee)
                hipLaunchKernelGGL((omv_wf_shade_cam<false, true>), dim3(L.blocksShade), blk, 0, stream, S, P, L.sample0, i, W, in, camCnt(i),
                                   H, out, camCnt(i + 1), sq, shCnt(i), eq, envCnt(i), L.rad, Gen1, lost);
            else
                hipLaunchKernelGGL((omv_wf_shade_cam<false, false>), dim3(L.blocksShade), blk, 0, stream, S, P, L.sample0, i, W, in, camCnt(i),
                                   H, out, camCnt(i + 1), sq, shCnt(i), eq, envCnt(i), L.rad, Gen1, lost);
        }
        bool sorted = false;
        const bool nextInWavefront = i < B && i + 1 != i0;   // round i + 1 is traced by this schedule's round kernels
        if (L.sortRays && nextInWavefront) {
            // reorder the next round's rays (secondary rays) into the spare queue
            hipLaunchKernelGGL(omv_wf_fill_u32, dim3(OMV_SORT_BUCKETS / 256), dim3(256), 0, stream, sortHist, OMV_SORT_BUCKETS);
            hipLaunchKernelGGL(omv_wf_sort_keys, dim3(L.blocksShade), blk, 0, stream, S, out, camCnt(i + 1).cnt, sortKeys, sortHist);
            hipLaunchKernelGGL(omv_wf_sort_scan, dim3(1), dim3(1024), 0, stream, sortHist);
            hipLaunchKernelGGL(omv_wf_sort_scatter, dim3(L.blocksShade), blk, 0, stream, out, camCnt(i + 1).cnt, sortKeys, sortHist, spare);
            sorted = true;
        }
        const WfQueue next = sorted ? spare : out;
        // the next round's hits go to H, which this round's connect does not read
        if (fuse) {
            mark();
            if (L.traceKeep > 0)
                hipLaunchKernelGGL(omv_wf_trace_pt<5>, dim3(L.blocksTrace), blk, lds, stream, S,
                                   ptArgs(next, nextInWavefront ? camCnt(i + 1) : none, lights ? shCnt(i) : none, eq,
                                          nee ? envCnt(i) : none));
            else
                hipLaunchKernelGGL(omv_wf_trace_fused, dim3(L.blocksTrace), blk, lds, stream, S, next, nextInWavefront ? camCnt(i + 1) : none,
                                   sq, lights ? shCnt(i) : none, eq, nee ? envCnt(i) : none, H, W.vis, W.nvis, n, L.lc, L.gstack,
                                   stats, xcds);
            mark();
            traced = nextInWavefront;
        } else {
            if (lights) {
                mark();
                if (L.traceKeep > 0)
                    hipLaunchKernelGGL(omv_wf_trace_pt<2>, dim3(L.blocksTrace), blk, lds, stream, S,
                                       ptArgs(in, none, shCnt(i), eq, none));
                else
                    hipLaunchKernelGGL(omv_wf_trace<2>, dim3(L.blocksTrace), blk, lds, stream, S, in, sq, shCnt(i), H, W.vis, n, L.lc,
                                       L.gstack, stats, L.splat, xcds, Gen0);
                mark();
            }
            if (nee) {
                mark();
                if (L.traceKeep > 0)
                    hipLaunchKernelGGL(omv_wf_trace_pt<3>, dim3(L.blocksTrace), blk, lds, stream, S,
                                       ptArgs(in, none, none, eq, envCnt(i)));
                else
                    hipLaunchKernelGGL(omv_wf_trace<3>, dim3(L.blocksTrace), blk, lds, stream, S, eq, sq, envCnt(i), H, W.nvis, n, L.lc,
                                       L.gstack, stats, L.splat, xcds, Gen0);
                mark();
            }
        }
        if (lights || nee)
            hipLaunchKernelGGL(omv_wf_connect, dim3(L.blocksShade), blk, 0, stream, W, in, camCnt(i), L.rad, lights ? 1 : 0, nee ? 1 : 0, i,
                               Gen1.gen);
        if (sorted) {
            const WfQueue freed = in;
            in = spare;
            spare = out;
            out = freed;
        } else {
            std::swap(in, out);
        }
        if (L.roundEv && i == std::min(L.roundEvAt, B)) { (void)hipEventRecord(L.roundEv, stream); roundEvDone = true; }
    }
    // fewer rounds, or the rounds from roundEvAt on ran in the tail launch
    if (L.roundEv && !roundEvDone) (void)hipEventRecord(L.roundEv, stream);
    hipLaunchKernelGGL(omv_wf_stats_sum, dim3(1), dim3(64), 0, stream, stats, L.counters);
    return hipGetLastError();
}

}  // namespace omv
