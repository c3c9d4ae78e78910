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
// 0: sample kernel (scene with lights), 1: FC kernel, 2: sample kernel (no lights)
int occupancyBlocksPerCU(int which);

}  // namespace omv
#endif
