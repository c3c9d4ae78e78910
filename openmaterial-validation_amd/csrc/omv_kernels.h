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
                      uint64_t nsamples, float* rad, float* lv, uint32_t* gstack, unsigned long long* counters, int blocks,
                      hipStream_t stream);
// Ordered per-pixel reduction + PinholeCamera::stop for pixels [pix0, pix0 + npix).
hipError_t launchFilm(const float* rad, int spp, uint64_t pix0, uint64_t npix, float* film, hipStream_t stream);
hipError_t launchFc(const OmvSceneView& S, const omv_camera_desc& cam, uint32_t seed, float* film, int32_t* ids, uint8_t* tie,
                    uint32_t* gstack, unsigned long long* counters, int blocks, hipStream_t stream);
// Batched jobs: the device job record is opaque to the host; packBatchJob
// writes one record of batchJobBytes() bytes.  tileEnd[j] is the exclusive
// end of job j's range of work tiles (64 * tileWaves pixels each).
// counters: 16 u64 (8 statistics + the work counter), cleared by the launch.
size_t batchJobBytes();
void packBatchJob(void* dst, const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, float* film,
                  uint64_t npix);
hipError_t launchBatch(const void* jobs, int njobs, const uint64_t* tileEnd, uint64_t totalTiles, int tileWaves,
                       unsigned long long* counters, float* lv, uint32_t* gstack, int blocks, hipStream_t stream);
// 0: sample kernel (scene with lights), 1: FC kernel, 2: sample kernel (no lights), 3: batch kernel
int occupancyBlocksPerCU(int which);

}  // namespace omv
#endif
