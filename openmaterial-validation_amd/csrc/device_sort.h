// device_sort.h -- stable LSD radix sort of (uint32 key, uint32 value)
// pairs for the commit-time device kernels (bvh_device.hip,
// mesh_device.hip).  Four 8-bit passes; each lane owns a contiguous chunk of
// kSortChunk pairs and walks it in order (digit counts per chunk, one
// exclusive scan over [digit][chunk], scatter), so equal keys keep their
// input order.  Commit-time sizes (up to a few million pairs) make the
// simple per-chunk walk adequate.
#ifndef OMV_DEVICE_SORT_H
#define OMV_DEVICE_SORT_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

namespace {

constexpr int kSortChunk = 256;
constexpr int kSortBlock = 256;

// digit counts per chunk, digit-major [digit][chunk]
__global__ void sortHist(const uint32_t* __restrict__ key, uint32_t n, int shift, uint32_t nchunks, uint32_t* __restrict__ hist) {
    const uint32_t c = blockIdx.x * kSortBlock + threadIdx.x;
    if (c >= nchunks) return;
    const uint32_t end = min(n, (c + 1) * (uint32_t)kSortChunk);
    for (uint32_t i = c * kSortChunk; i < end; ++i) {
        const uint32_t d = (key[i] >> shift) & 0xFFu;
        hist[(size_t)d * nchunks + c] += 1u;
    }
}

// exclusive scan of m counts in one block (each lane scans a contiguous span)
__global__ void sortScan(uint32_t* __restrict__ a, uint32_t m) {
    __shared__ uint32_t sums[kSortBlock];
    const uint32_t per = (m + kSortBlock - 1) / kSortBlock;
    const uint32_t b = threadIdx.x * per, e = min(m, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += a[i];
    sums[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int t = 0; t < kSortBlock; ++t) { const uint32_t x = sums[t]; sums[t] = acc; acc += x; }
    }
    __syncthreads();
    uint32_t acc = sums[threadIdx.x];
    for (uint32_t i = b; i < e; ++i) { const uint32_t x = a[i]; a[i] = acc; acc += x; }
}

// stable scatter: each lane walks its chunk in order
__global__ void sortScatter(const uint32_t* __restrict__ key, const uint32_t* __restrict__ val, uint32_t n, int shift, uint32_t nchunks,
                            uint32_t* __restrict__ offs, uint32_t* __restrict__ keyOut, uint32_t* __restrict__ valOut) {
    const uint32_t c = blockIdx.x * kSortBlock + threadIdx.x;
    if (c >= nchunks) return;
    const uint32_t end = min(n, (c + 1) * (uint32_t)kSortChunk);
    for (uint32_t i = c * kSortChunk; i < end; ++i) {
        const uint32_t d = (key[i] >> shift) & 0xFFu;
        const uint32_t o = offs[(size_t)d * nchunks + c];
        offs[(size_t)d * nchunks + c] = o + 1u;
        keyOut[o] = key[i];
        valOut[o] = val[i];
    }
}

inline uint32_t sortChunks(uint32_t n) { return (n + kSortChunk - 1) / kSortChunk; }

// sorts (keyA, valA) in place using (keyB, valB) and hist (256 * sortChunks(n) words)
inline hipError_t sortPairs32(uint32_t* keyA, uint32_t* valA, uint32_t* keyB, uint32_t* valB, uint32_t* hist, uint32_t n,
                              hipStream_t stream) {
    const uint32_t nc = sortChunks(n);
    const uint32_t nb = (nc + kSortBlock - 1) / kSortBlock;
    uint32_t* kIn = keyA; uint32_t* vIn = valA; uint32_t* kOut = keyB; uint32_t* vOut = valB;
    for (int shift = 0; shift < 32; shift += 8) {
        hipError_t e = hipMemsetAsync(hist, 0, (size_t)256 * nc * 4, stream);
        if (e != hipSuccess) return e;
        sortHist<<<nb, kSortBlock, 0, stream>>>(kIn, n, shift, nc, hist);
        sortScan<<<1, kSortBlock, 0, stream>>>(hist, 256u * nc);
        sortScatter<<<nb, kSortBlock, 0, stream>>>(kIn, vIn, n, shift, nc, hist, kOut, vOut);
        std::swap(kIn, kOut);
        std::swap(vIn, vOut);
    }
    return hipGetLastError();   // four passes: the result is back in (keyA, valA)
}

}  // namespace

#endif
