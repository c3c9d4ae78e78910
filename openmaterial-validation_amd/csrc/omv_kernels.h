// omv_kernels.h -- host-side entry points of the gfx950 kernels (kernels.hip).
#ifndef OMV_KERNELS_H
#define OMV_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/omv.h"
#include "omv_types.h"

#define OMV_BLOCK 128   // two wave64s per workgroup

namespace omv {

// Sample kernel over samples [sample0, sample0 + nsamples) of a chunk that
// covers whole pixels; rad receives 3 floats per sample (chunk-relative).
hipError_t launchBdpt(const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, uint64_t sample0,
                      uint64_t nsamples, float* rad, float* lv, uint32_t* gstack, unsigned long long* counters, int blocks, float* splat,
                      hipStream_t stream);
// Ordered per-pixel reduction + PinholeCamera::stop for pixels [pix0, pix0 + npix).
// OMV_EXT_LIGHT_SPLAT: film (float4 per pixel) += splat (3 floats per pixel) * inv
hipError_t launchSplatAdd(const float* splat, uint64_t npix, float inv, float* film, hipStream_t stream);
hipError_t launchFilm(const float* rad, int spp, uint64_t pix0, uint64_t npix, float* film, hipStream_t stream);
hipError_t launchFc(const OmvSceneView& S, const omv_camera_desc& cam, uint32_t seed, float* film, int32_t* ids, uint8_t* tie,
                    uint32_t* gstack, unsigned long long* counters, int blocks, hipStream_t stream);
// Batched jobs: the device job record is opaque to the host; packBatchJob
// writes one record of batchJobBytes() bytes.  tileEnd[j] is the exclusive
// end of job j's range of work tiles (64 * tileWaves pixels each).
// counters: 16 u64 (8 statistics + the work counter), cleared by the launch.
hipError_t launchFillU64(unsigned long long* p, int n, hipStream_t stream);   // zero-fill, graph-capturable
size_t batchJobBytes();
void packBatchJob(void* dst, const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, float* film,
                  uint64_t npix);
hipError_t launchBatch(const void* jobs, int njobs, const uint64_t* tileEnd, uint64_t totalTiles, int tileWaves,
                       unsigned long long* counters, float* lv, uint32_t* gstack, int blocks, bool nee, hipStream_t stream);
// Allocation-free render (noalloc.hip): one persistent per-pixel kernel over
// the file-scope scratch array; statsOut (may be null: no wait) receives the
// 8 counters after the stream has finished.  On failure err / what say why.
int renderNoalloc(const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, float* film, hipStream_t stream,
                  unsigned long long statsOut[8], hipError_t& err, const char*& what);
// Wavefront render of one chunk of whole pixels (wavefront.hip).
struct WfLaunch {
    OmvSceneView S;
    omv_camera_desc cam;
    omv_renderer_params p;
    uint64_t sample0;          // first sample of the chunk
    uint32_t nslots;           // samples in the chunk
    int lvCap, lc;             // light vertices kept, connections per camera vertex
    void* arena;               // wavefrontArenaBytes(nslots, lvCap, lc, lights, nee, splat, sortRays, blocksShade, blocksGen)
    uint32_t* qcount;          // wavefrontCounterCount(B, lvCap) counters
    float* rad;                // 3 floats per sample of the chunk
    uint32_t* gstack;          // spill stacks of blocksTrace * OMV_BLOCK lanes
    unsigned long long* counters;
    int blocksGen, blocksTrace, blocksShade;
    int traceKeep;             // > 0: persistent traversal, refill below this many busy lanes
    hipEvent_t* traceEv;       // optional: an event pair around every traversal launch
    int traceEvCap;            // events available from traceEv
    int* traceEvUsed;          // events recorded so far (in/out)
    float* splat;              // OMV_EXT_LIGHT_SPLAT sums (3 floats per pixel), or null
    bool sortRays;             // reorder secondary camera rays before each traversal
    hipEvent_t roundEv;        // optional: recorded once camera round roundEvAt is
    int roundEvAt;             // done (or at the end when there are fewer rounds)
    int xcdBands;              // > 1: traversal queues dealt to this many XCDs in chunks (L2 locality); else plain stride
    int tailRound;             // >= 2: camera rounds tailRound .. B run as one launch (omv_wf_tail); else round by round
    int blocksPacket;          // > 0: round-1 camera rays traced as wave packets (omv_wf_trace_cam1) on this many workgroups
};
// Ray queues are appended to through OMV_QSEG counters each (one per group of
// producing workgroups); wavefrontSegCap is the entries of one group's region
// for a producing launch of gridShade workgroups (blocksShade).
#define OMV_QSEG 64
uint32_t wavefrontSegCap(uint32_t nslots, int gridShade, int gridGen);
size_t wavefrontArenaBytes(uint32_t nslots, int lvCap, int lc, bool lights, bool nee, bool splat, bool sortRays, int gridShade, int gridGen);
int wavefrontCounterCount(int B, int lvCap);
// resident blocks per CU: [0] trace, [1] camera shading with lights + NEE,
// [2] light shading, [3] gen, [4] persistent trace, [5] camera shading with
// lights, [6] with NEE, [7] with neither, [8] round-1 packet trace
void wavefrontBlocksPerCU(int out[9]);
inline int shadeCamBlocksPerCU(const int occ[9], bool lights, bool nee) { return lights ? (nee ? occ[1] : occ[5]) : (nee ? occ[6] : occ[7]); }
hipError_t launchWavefront(const WfLaunch& L, hipStream_t stream);
// Path-replay adjoint over samples [sample0, sample0 + nsamples) (adjoint.hip).
int adjointBlocksPerCU();
hipError_t launchAdjoint(const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, uint64_t sample0,
                         uint64_t nsamples, const float* dfilm, const int* texSlot, float* const* texGrad, const int* omSlot,
                         float* omGrad, float* lv, uint32_t* gstack, unsigned long long* counters, int blocks, hipStream_t stream);
// FP64 detached film (adjoint.hip): directions from S0, weights with S1's
// parameters, W*H*4 doubles; gstack holds blocks * OMV_BLOCK spill stacks.
int detachedBlocksPerCU();
hipError_t launchDetached(const OmvSceneView& S0, const OmvSceneView& S1, const omv_camera_desc& cam, const omv_renderer_params& p,
                          uint64_t npix, double* film, uint32_t* gstack, unsigned long long* counters, int blocks, hipStream_t stream);
// Film post pass (post.hip): gamma (+ 3x3 median when denoise) into outF, and
// the 8-bit truncation of outF into outU8 when non-null; tmp holds w*h float4.
hipError_t launchPost(const float* film, float* tmp, float* outF, uint8_t* outU8, int w, int h, float A, float gamma, bool denoise,
                      hipStream_t stream);
// Environment alias tables (env.hip); scratch holds 4*W*H + 2*H + 2 words.
// bvh_device.hip: one mesh's BLAS built on the device (linear BVH collapsed
// into OmvNode4) into nodes[nodeBase ...] and records[12 * recBase ...]
size_t blasDeviceScratchBytes(uint32_t n);
hipError_t buildBlasDevice(const float* vtx, uint32_t vtxOff, uint32_t stride, const uint32_t* faces, uint32_t faceOff, uint32_t n,
                           int leafMax, int recBase, int nodeBase, uint32_t nodeCap, OmvNode4* nodes, float* records, void* scratch,
                           hipStream_t stream, int* root, int* nodeCount, int* depth);
// mesh_device.hip: commitMesh's normals / tangents of one mesh on the device
size_t meshDeviceScratchBytes(uint32_t vn, uint32_t fn);
hipError_t preprocessMeshDevice(float* vertices, uint32_t stride, uint32_t vn, const uint32_t* faces, uint32_t fn, bool normals,
                                bool tangents, int uvOff, void* scratch, hipStream_t stream, uint32_t* badIndices);
hipError_t launchEnvBuild(const float* bg, int W, int H, float* rowProb, int* rowAlias, float* colProb, int* colAlias, float* texPdf,
                          void* scratch, hipStream_t stream);
// 0: sample kernel (scene with lights), 1: FC kernel, 2: sample kernel (no lights), 3: batch kernel;
// 4, 5, 6: the sample kernels and the batch kernel with OMV_EXT_HDRI_NEE code
int occupancyBlocksPerCU(int which);

}  // namespace omv
#endif
