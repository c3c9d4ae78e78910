// noalloc.hip -- the allocation-free render entry point's device side.
//
// Some callers cannot allocate device memory at render time (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).  For
// them the render runs as one persistent per-pixel kernel whose only device
// memory beyond the committed scene (input) and the caller's film (output)
// is the file-scope array `scratch` below: the counters, the traversal spill
// stacks and the light-vertex scratch of every resident lane.  Each GPU the
// library runs on holds its own copy of the array until the process ends.
//
// In host code `scratch` names a host-side stand-in, not a device address;
// the device address is fetched once per device with hipGetSymbolAddress
// (the device current) and only that address is handed to kernels.
//
// Every call on a GPU shares that GPU's array; calls that overlap on one GPU
// race on it.
#include <algorithm>
#include <functional>
#include <mutex>

#include "path_device.h"
#include "omv_kernels.h"

namespace {

// Scratch layout (bytes): [0, 128) 16 u64 counters (0..6 statistics, 8 the
// work counter); [256, 256 + S) traversal spill stacks, S = lanes * 96 * 4;
// then, 256-aligned, the light vertices, lanes * lvCap * 6 * 4.
constexpr size_t kCounterBytes = 256;
constexpr size_t kSpillPerLane = (size_t)(OMV_STACK_TOTAL - OMV_STACK_LDS) * sizeof(uint32_t);
constexpr int kTileWaves = 2;                // pixels per work item = 64 * kTileWaves

static_assert(OMV_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size must be a multiple of 256");
// the documented light-vertex limit is exactly what the scratch holds for
// OMV_NOALLOC_MIN_BLOCKS blocks (the check in renderNoalloc)
static_assert(kCounterBytes + 2 * 256 + (size_t)OMV_NOALLOC_MIN_BLOCKS * OMV_BLOCK *
                                            (kSpillPerLane + (size_t)OMV_NOALLOC_MAX_LIGHT_VERTICES * 24) <= OMV_NOALLOC_SCRATCH_BYTES,
              "OMV_NOALLOC_MAX_LIGHT_VERTICES does not fit the scratch");
static_assert(kCounterBytes + 2 * 256 + (size_t)OMV_NOALLOC_MIN_BLOCKS * OMV_BLOCK *
                                            (kSpillPerLane + (size_t)(OMV_NOALLOC_MAX_LIGHT_VERTICES + 1) * 24) > OMV_NOALLOC_SCRATCH_BYTES,
              "OMV_NOALLOC_MAX_LIGHT_VERTICES is below what the scratch holds");
static __device__ __align__(256) unsigned char scratch[OMV_NOALLOC_SCRATCH_BYTES];

// One pass over the film: a wave claims kTileWaves * 64 consecutive pixels
// with one atomic on the work counter; each lane renders its pixel's spp
// samples in sample order and writes the normalised pixel.  Every wave leaves
// the loop once the counter has passed the last tile.
template <bool kLights, bool kNee>
__global__ __launch_bounds__(OMV_BLOCK) void omv_pixel_kernel(OmvSceneView S, omv_camera_desc cam, RP P, int spp, uint64_t npix,
                                                              uint64_t tiles, unsigned long long* __restrict__ counters,
                                                              uint32_t* __restrict__ gstack, float* __restrict__ lv,
                                                              float* __restrict__ film) {
    const uint32_t gthreads = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    Lane L;
    laneInit(L, gstack, lv, gtid, gthreads);
    for (;;) {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(counters + 8, 1ull);
        t = __shfl(t, 0, 64);
        if (t >= tiles) break;
        const uint64_t pix0 = t * 64ull * (uint64_t)kTileWaves;
        for (int w = 0; w < kTileWaves; ++w) {
            const uint64_t p = pix0 + (uint64_t)w * 64ull + (uint64_t)lane;
            if (p >= npix) continue;
            reinterpret_cast<float4*>(film)[p] = renderPixel<kLights, kNee>(S, cam, P, spp, p, L);
        }
    }
    flushCounters(counters, L.rays, L.camRays, L.lightRays, L.shadowRays, L.cnt);
}

__global__ void omv_noalloc_clear(unsigned long long* c) {
    if (threadIdx.x < 16) c[threadIdx.x] = 0ull;
}

// What one device needs before its first call: the scratch address and the
// launch limits, fetched once with the device current.
struct DeviceSetup {
    std::once_flag once;
    hipError_t err = hipSuccess;
    unsigned char* ptr = nullptr;   // device address of this device's `scratch`
    int cus = 0;
    int occ[4] = {0, 0, 0, 0};      // resident blocks per CU: [lights][nee]
};
constexpr int kMaxDevices = 64;
DeviceSetup g_setup[kMaxDevices];

void setUp(DeviceSetup& d, int dev) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) { d.err = e; return; }
    d.ptr = static_cast<unsigned char*>(p);
    if ((e = hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) { d.err = e; return; }
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    hipError_t o[4];
    o[0] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&d.occ[0], omv_pixel_kernel<false, false>, OMV_BLOCK, lds);
    o[1] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&d.occ[1], omv_pixel_kernel<false, true>, OMV_BLOCK, lds);
    o[2] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&d.occ[2], omv_pixel_kernel<true, false>, OMV_BLOCK, lds);
    o[3] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&d.occ[3], omv_pixel_kernel<true, true>, OMV_BLOCK, lds);
    for (int i = 0; i < 4; ++i) {
        if (o[i] != hipSuccess) { d.err = o[i]; return; }
        if (d.occ[i] < 1) d.occ[i] = 1;
    }
}

size_t alignUp(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

namespace omv {

size_t noallocLaneBytes(int lvCap, bool lights) {
    return kSpillPerLane + (lights ? (size_t)lvCap * 6 * sizeof(float) : 0);
}

int renderNoalloc(const OmvSceneView& S, const omv_camera_desc& cam, const omv_renderer_params& p, float* film, hipStream_t stream,
                  unsigned long long statsOut[8], hipError_t& err, const char*& what) {
    err = hipSuccess;
    what = nullptr;
    int dev = 0;
    if ((err = hipGetDevice(&dev)) != hipSuccess) { what = "hipGetDevice"; return OMV_E_DEVICE; }
    if (dev < 0 || dev >= kMaxDevices) { what = "render_noalloc: device index beyond the supported 64"; return OMV_E_RANGE; }
    DeviceSetup& d = g_setup[dev];
    std::call_once(d.once, setUp, std::ref(d), dev);
    if (d.err != hipSuccess) { err = d.err; what = "render_noalloc: scratch set-up (hipGetSymbolAddress / occupancy)"; return OMV_E_DEVICE; }

    const bool lights = S.num_lights > 0;
    RP P = toRP(p);
    if (!S.env_tex_pdf) P.x.nee = false;   // no HDR environment: nothing to sample
    P.x.splat = false;
    const int lvCap = std::max(std::min(p.light_bounces + 1, std::max(p.max_path_length, 1)), 1);
    const size_t perLane = noallocLaneBytes(lvCap, lights);
    const uint64_t minLanes = (uint64_t)OMV_NOALLOC_MIN_BLOCKS * OMV_BLOCK;
    if (kCounterBytes + 2 * 256 + minLanes * perLane > (size_t)OMV_NOALLOC_SCRATCH_BYTES) {
        what = "render_noalloc: the light subpath keeps more vertices than the fixed scratch holds "
               "(min(light_bounces + 1, max_path_length) must be <= OMV_NOALLOC_MAX_LIGHT_VERTICES)";
        return OMV_E_RANGE;
    }
    const uint64_t npix = (uint64_t)cam.width * (uint64_t)cam.height;
    const uint64_t tiles = (npix + 64ull * kTileWaves - 1) / (64ull * kTileWaves);
    // resident blocks: the occupancy limit, as many as the scratch holds, and
    // no more than there are work items for (two waves per block)
    const int occ = d.occ[(lights ? 2 : 0) + (P.x.nee ? 1 : 0)];
    uint64_t blocks = (uint64_t)d.cus * (uint64_t)occ;
    const uint64_t fit = ((size_t)OMV_NOALLOC_SCRATCH_BYTES - kCounterBytes - 2 * 256) / (perLane * OMV_BLOCK);
    blocks = std::min(blocks, fit);
    blocks = std::min<uint64_t>(blocks, std::max<uint64_t>(1, (tiles + 1) / 2));
    blocks = std::max<uint64_t>(blocks, 1);
    const uint64_t lanes = blocks * OMV_BLOCK;
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(d.ptr);
    uint32_t* gstack = reinterpret_cast<uint32_t*>(d.ptr + kCounterBytes);
    float* lv = reinterpret_cast<float*>(d.ptr + alignUp(kCounterBytes + lanes * kSpillPerLane));
    // the layout ends inside the array (checked on the host before any launch)
    const size_t end = alignUp(kCounterBytes + lanes * kSpillPerLane) + (lights ? lanes * (size_t)lvCap * 6 * sizeof(float) : 0);
    if (end > (size_t)OMV_NOALLOC_SCRATCH_BYTES) { what = "render_noalloc: scratch layout overflow"; return OMV_E_INTERNAL; }

    hipLaunchKernelGGL(omv_noalloc_clear, dim3(1), dim3(64), 0, stream, counters);
    if ((err = hipGetLastError()) != hipSuccess) { what = "render_noalloc: clear launch"; return OMV_E_DEVICE; }
    const size_t lds = (size_t)OMV_STACK_LDS * OMV_BLOCK * sizeof(uint32_t);
    const dim3 g((unsigned)blocks), b(OMV_BLOCK);
    const int spp = p.samples_per_pixel;
    if (lights && P.x.nee)
        hipLaunchKernelGGL((omv_pixel_kernel<true, true>), g, b, lds, stream, S, cam, P, spp, npix, tiles, counters, gstack, lv, film);
    else if (lights)
        hipLaunchKernelGGL((omv_pixel_kernel<true, false>), g, b, lds, stream, S, cam, P, spp, npix, tiles, counters, gstack, lv, film);
    else if (P.x.nee)
        hipLaunchKernelGGL((omv_pixel_kernel<false, true>), g, b, lds, stream, S, cam, P, spp, npix, tiles, counters, gstack, lv, film);
    else
        hipLaunchKernelGGL((omv_pixel_kernel<false, false>), g, b, lds, stream, S, cam, P, spp, npix, tiles, counters, gstack, lv, film);
    if ((err = hipGetLastError()) != hipSuccess) { what = "render_noalloc: kernel launch"; return OMV_E_DEVICE; }
    if (statsOut) {
        if ((err = hipMemcpyAsync(statsOut, counters, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream)) != hipSuccess) {
            what = "render_noalloc: counter read-back";
            return OMV_E_DEVICE;
        }
        if ((err = hipStreamSynchronize(stream)) != hipSuccess) { what = "render_noalloc"; return OMV_E_DEVICE; }
    }
    return OMV_OK;
}

}  // namespace omv
