// bvh_device.hip -- bottom-level BVH build on the device, the alternative to
// the host binned-SAH builder (bvh.cpp) when commit time matters more than
// the last node visit (batched validation jobs).
//
//   1. one lane per triangle: box and centroid from the vertex pool;
//      centroid bounds by a block reduction;
//   2. 30-bit Morton code of each centroid in those bounds; stable LSD radix
//      sort of (code, triangle) in four 8-bit passes -- each sort lane owns
//      a contiguous chunk of 256 keys and walks it in order, so equal codes
//      keep triangle order;
//   3. the binary radix tree over the sorted codes (Karras, "Maximizing
//      parallelism in the construction of BVHs, octrees, and k-d trees",
//      2012), codes made unique by appending the sorted position;
//   4. boxes bottom-up: one lane per leaf climbs while it is the second
//      child to arrive at a node (atomic counter per node);
//   5. collapse into the 128-byte 4-wide nodes the traversal reads, one
//      level per launch: a node takes its binary node's two children and
//      keeps replacing the largest-area one by its children until it has
//      four; subtrees of <= leafMax triangles become leaves (their triangles
//      are contiguous in sorted order); child boxes widened two ulps outward,
//      as the host builder does;
//   6. triangle records (v0 + id, v1, v2) in sorted order.
//
// The closest hit does not depend on the tree (watertight test, ties broken
// by (t, instance, primitive)), so films equal the host-built scene's bit
// for bit; only node and triangle counts change.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_sort.h"
#include "omv_kernels.h"

namespace {

constexpr int kChunk = 256;          // keys per sort lane
constexpr int kBlock = 256;
constexpr uint32_t kLeafBit = 0x80000000u;

__device__ inline float padDown(float v) { return nextafterf(nextafterf(v, -INFINITY), -INFINITY); }
__device__ inline float padUp(float v) { return nextafterf(nextafterf(v, INFINITY), INFINITY); }

// 1. triangle boxes and centroids; block partials of the centroid bounds
__global__ void lbBoxes(const float* __restrict__ vtx, uint32_t vtxOff, uint32_t stride, const uint32_t* __restrict__ faces,
                        uint32_t faceOff, uint32_t n, float* __restrict__ box, float* __restrict__ cen, float* __restrict__ part) {
    __shared__ float red[6][kBlock];
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float c[3] = {INFINITY, INFINITY, INFINITY}, cm[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (f < n) {
        for (int k = 0; k < 3; ++k) {
            const float* p = vtx + vtxOff + (size_t)faces[faceOff + 3 * f + k] * stride;
            for (int a = 0; a < 3; ++a) { lo[a] = fminf(lo[a], p[a]); hi[a] = fmaxf(hi[a], p[a]); }
        }
        for (int a = 0; a < 3; ++a) {
            box[(size_t)6 * f + a] = lo[a];
            box[(size_t)6 * f + 3 + a] = hi[a];
            c[a] = cm[a] = 0.5f * lo[a] + 0.5f * hi[a];
            cen[(size_t)3 * f + a] = c[a];
        }
    }
    for (int a = 0; a < 3; ++a) { red[a][threadIdx.x] = c[a]; red[3 + a][threadIdx.x] = cm[a]; }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 6; ++a) part[(size_t)6 * blockIdx.x + a] = red[a][0];
}

// one block folds the partials
__global__ void lbBounds(const float* __restrict__ part, uint32_t nb, float* __restrict__ out) {
    __shared__ float red[6][kBlock];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < nb; b += kBlock)
        for (int a = 0; a < 3; ++a) {
            v[a] = fminf(v[a], part[(size_t)6 * b + a]);
            v[3 + a] = fmaxf(v[3 + a], part[(size_t)6 * b + 3 + a]);
        }
    for (int a = 0; a < 6; ++a) red[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 6; ++a) out[a] = red[a][0];
}

__device__ inline uint32_t spread10(uint32_t v) {   // 10 bits -> every third bit
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 2a. Morton codes
__global__ void lbMorton(const float* __restrict__ cen, const float* __restrict__ bounds, uint32_t n, uint32_t* __restrict__ key,
                         uint32_t* __restrict__ val) {
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= n) return;
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = bounds[3 + a] - bounds[a];
        const float t = ext > 0.0f ? (cen[(size_t)3 * f + a] - bounds[a]) / ext : 0.0f;
        const float s = fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f);
        q[a] = (uint32_t)s;
    }
    key[f] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    val[f] = f;
}

// 3. radix tree; codes extended by the sorted position make them unique
__device__ inline int prefixLen(const uint32_t* __restrict__ key, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = ((uint64_t)key[i] << 32) | (uint32_t)i, b = ((uint64_t)key[j] << 32) | (uint32_t)j;
    return __clzll(a ^ b);
}

__global__ void lbTree(const uint32_t* __restrict__ key, int n, uint32_t* __restrict__ left, uint32_t* __restrict__ right,
                       int* __restrict__ first, int* __restrict__ last, int* __restrict__ parentNode, int* __restrict__ parentLeaf) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    const int d = prefixLen(key, n, i, i + 1) > prefixLen(key, n, i, i - 1) ? 1 : -1;
    const int dmin = prefixLen(key, n, i, i - d);
    int lmax = 2;
    while (prefixLen(key, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (prefixLen(key, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = prefixLen(key, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (prefixLen(key, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    first[i] = lo;
    last[i] = hi;
    if (lo == gamma) { left[i] = kLeafBit | (uint32_t)gamma; parentLeaf[gamma] = i; }
    else { left[i] = (uint32_t)gamma; parentNode[gamma] = i; }
    if (hi == gamma + 1) { right[i] = kLeafBit | (uint32_t)(gamma + 1); parentLeaf[gamma + 1] = i; }
    else { right[i] = (uint32_t)(gamma + 1); parentNode[gamma + 1] = i; }
}

__device__ inline void loadBox(const float* __restrict__ box, uint32_t idx, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = __hip_atomic_load(&box[(size_t)6 * idx + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hi[a] = __hip_atomic_load(&box[(size_t)6 * idx + 3 + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// 4. node boxes bottom-up (the second child to arrive builds the parent)
__global__ void lbRefit(const uint32_t* __restrict__ val, const float* __restrict__ triBox, int n, const uint32_t* __restrict__ left,
                        const uint32_t* __restrict__ right, const int* __restrict__ parentNode, const int* __restrict__ parentLeaf,
                        float* __restrict__ leafBox, float* __restrict__ nodeBox, uint32_t* __restrict__ arrived) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t prim = val[i];
    for (int a = 0; a < 6; ++a)
        __hip_atomic_store(&leafBox[(size_t)6 * i + a], triBox[(size_t)6 * prim + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int p = n > 1 ? parentLeaf[i] : -1;
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(&arrived[p], 1u) == 0u) return;   // the sibling is not done yet
        __threadfence();
        float lo[3], hi[3], lo2[3], hi2[3];
        const uint32_t l = left[p], r = right[p];
        loadBox((l & kLeafBit) ? leafBox : nodeBox, l & ~kLeafBit, lo, hi);
        loadBox((r & kLeafBit) ? leafBox : nodeBox, r & ~kLeafBit, lo2, hi2);
        for (int a = 0; a < 3; ++a) {
            __hip_atomic_store(&nodeBox[(size_t)6 * p + a], fminf(lo[a], lo2[a]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&nodeBox[(size_t)6 * p + 3 + a], fmaxf(hi[a], hi2[a]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        p = parentNode[p];
    }
}

struct BinRef {   // a binary tree node: internal index, or a leaf with kLeafBit
    uint32_t ref;
};

__device__ inline int countOf(uint32_t ref, const int* first, const int* last) {
    return (ref & kLeafBit) ? 1 : last[ref] - first[ref] + 1;
}
__device__ inline int firstOf(uint32_t ref, const int* first) { return (ref & kLeafBit) ? (int)(ref & ~kLeafBit) : first[ref]; }
__device__ inline float halfArea(const float lo[3], const float hi[3]) {
    const float d0 = fmaxf(0.0f, hi[0] - lo[0]), d1 = fmaxf(0.0f, hi[1] - lo[1]), d2 = fmaxf(0.0f, hi[2] - lo[2]);
    return d0 * d1 + d1 * d2 + d2 * d0;
}

// 5. one level of the collapse: jobs are (binary node, output node)
__global__ void lbCollapse(const uint2* __restrict__ jobs, uint32_t njobs, const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                           const int* __restrict__ first, const int* __restrict__ last, const float* __restrict__ leafBox,
                           const float* __restrict__ nodeBox, int leafMax, int recBase, int nodeBase, uint32_t nodeCap,
                           OmvNode4* __restrict__ out, uint32_t* __restrict__ nodeCount, uint2* __restrict__ next, uint32_t* __restrict__ nextCount,
                           uint32_t* __restrict__ overflow) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= njobs) return;
    const uint32_t b = jobs[t].x, o = jobs[t].y;
    uint32_t kids[4];
    int nk = 0;
    if ((b & kLeafBit) || countOf(b, first, last) <= leafMax) {
        kids[nk++] = b;   // a small root: one-child node holding a leaf
    } else {
        kids[nk++] = left[b];
        kids[nk++] = right[b];
        while (nk < 4) {
            int best = -1;
            float bestA = -1.0f;
            for (int k = 0; k < nk; ++k) {
                if ((kids[k] & kLeafBit) || countOf(kids[k], first, last) <= leafMax) continue;
                float lo[3], hi[3];
                for (int a = 0; a < 3; ++a) { lo[a] = nodeBox[(size_t)6 * kids[k] + a]; hi[a] = nodeBox[(size_t)6 * kids[k] + 3 + a]; }
                const float ar = halfArea(lo, hi);
                if (ar > bestA) { bestA = ar; best = k; }
            }
            if (best < 0) break;
            const uint32_t inner = kids[best];
            kids[best] = left[inner];
            kids[nk++] = right[inner];
        }
    }
    OmvNode4 nd;
    for (int k = 0; k < 4; ++k) {
        if (k >= nk) {
            nd.child[k] = OMV_EMPTY_CHILD;
            nd.entry[k] = OMV_ENT_NONE;
            nd.lox[k] = nd.loy[k] = nd.loz[k] = nd.hix[k] = nd.hiy[k] = nd.hiz[k] = INFINITY;   // no ray hits it
            continue;
        }
        const uint32_t c = kids[k];
        const float* bx = (c & kLeafBit) ? leafBox + (size_t)6 * (c & ~kLeafBit) : nodeBox + (size_t)6 * c;
        nd.lox[k] = padDown(bx[0]); nd.loy[k] = padDown(bx[1]); nd.loz[k] = padDown(bx[2]);
        nd.hix[k] = padUp(bx[3]); nd.hiy[k] = padUp(bx[4]); nd.hiz[k] = padUp(bx[5]);
        const int cnt = countOf(c, first, last);
        if (cnt <= leafMax) {
            nd.child[k] = -(recBase + firstOf(c, first) + 1);
            nd.entry[k] = OMV_ENT(OMV_ENT_BLAS_LEAF, ((uint32_t)(recBase + firstOf(c, first)) << 3) | ((uint32_t)(cnt - 1) & 7u));
        } else {
            const uint32_t idx = atomicAdd(nodeCount, 1u);
            if (idx >= nodeCap) {   // cannot happen (one output node per binary node); keep the launch in bounds
                atomicAdd(overflow, 1u);
                nd.child[k] = OMV_EMPTY_CHILD;
                nd.entry[k] = OMV_ENT_NONE;
                continue;
            }
            nd.child[k] = nodeBase + (int)idx;
            nd.entry[k] = OMV_ENT(OMV_ENT_BLAS_NODE, nodeBase + (int)idx);
            const uint32_t q = atomicAdd(nextCount, 1u);
            next[q] = make_uint2(c, idx);
        }
    }
    out[o] = nd;
}

// 6. triangle records in sorted order
__global__ void lbRecords(const float* __restrict__ vtx, uint32_t vtxOff, uint32_t stride, const uint32_t* __restrict__ faces,
                          uint32_t faceOff, const uint32_t* __restrict__ val, uint32_t n, float* __restrict__ rec) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t prim = val[i];
    float* r = rec + (size_t)12 * i;
    for (int k = 0; k < 3; ++k) {
        const float* p = vtx + vtxOff + (size_t)faces[faceOff + 3 * prim + k] * stride;
        r[4 * k + 0] = p[0]; r[4 * k + 1] = p[1]; r[4 * k + 2] = p[2];
        r[4 * k + 3] = k == 0 ? __uint_as_float(prim) : 0.0f;
    }
}

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace

namespace omv {

// Scratch sizes for n triangles, in bytes (one allocation, carved below).
size_t blasDeviceScratchBytes(uint32_t n) {
    const uint64_t nc = (n + kChunk - 1) / kChunk;
    // 22 arrays, each rounded up to an even number of words
    return 4 * ((uint64_t)6 * n + 3 * n + 6 * blocks(n) + 8 + 4 * (uint64_t)n + 256 * nc + 4 * (uint64_t)n + 2 * (uint64_t)n +
                6 * (uint64_t)n + 6 * (uint64_t)n + (uint64_t)n + 4 * (uint64_t)n + 8 + 2 * 22) + 256;
}

hipError_t buildBlasDevice(const float* vtx, uint32_t vtxOff, uint32_t stride, const uint32_t* faces, uint32_t faceOff, uint32_t n,
                           int leafMax, int recBase, int nodeBase, uint32_t nodeCap, OmvNode4* nodes, float* records, void* scratch,
                           hipStream_t stream, int* root, int* nodeCount, int* depth) {
    *root = -1; *nodeCount = 0; *depth = 0;
    if (n == 0 || nodeCap == 0) return hipSuccess;
    if (nodeCap > n) nodeCap = n;   // one output node per binary node at most: the job queues hold n entries
    const uint32_t nb = blocks(n), nc = (n + kChunk - 1) / kChunk;
    uint32_t* p = static_cast<uint32_t*>(scratch);
    auto take = [&p](uint64_t words) { uint32_t* q = p; p += (words + 1) & ~(uint64_t)1; return q; };   // 8-byte aligned
    float* triBox = reinterpret_cast<float*>(take((uint64_t)6 * n));
    float* cen = reinterpret_cast<float*>(take((uint64_t)3 * n));
    float* part = reinterpret_cast<float*>(take((uint64_t)6 * nb));
    float* bounds = reinterpret_cast<float*>(take(8));
    uint32_t* keyA = take(n);
    uint32_t* valA = take(n);
    uint32_t* keyB = take(n);
    uint32_t* valB = take(n);
    uint32_t* hist = take((uint64_t)256 * nc);
    uint32_t* left = take(n);
    uint32_t* right = take(n);
    int* first = reinterpret_cast<int*>(take(n));
    int* last = reinterpret_cast<int*>(take(n));
    int* parentNode = reinterpret_cast<int*>(take(n));
    int* parentLeaf = reinterpret_cast<int*>(take(n));
    float* leafBox = reinterpret_cast<float*>(take((uint64_t)6 * n));
    float* nodeBox = reinterpret_cast<float*>(take((uint64_t)6 * n));
    uint32_t* arrived = take(n);
    uint2* jobsA = reinterpret_cast<uint2*>(take((uint64_t)2 * n));
    uint2* jobsB = reinterpret_cast<uint2*>(take((uint64_t)2 * n));
    uint32_t* counters = take(8);   // [0] node count, [1] next jobs, [2] overflow

    hipError_t e;
    lbBoxes<<<nb, kBlock, 0, stream>>>(vtx, vtxOff, stride, faces, faceOff, n, triBox, cen, part);
    lbBounds<<<1, kBlock, 0, stream>>>(part, nb, bounds);
    lbMorton<<<nb, kBlock, 0, stream>>>(cen, bounds, n, keyA, valA);
    if ((e = sortPairs32(keyA, valA, keyB, valB, hist, n, stream)) != hipSuccess) return e;
    uint32_t* kIn = keyA; uint32_t* vIn = valA;
    // the root has no parent
    if ((e = hipMemsetAsync(parentNode, 0xFF, (size_t)n * 4, stream)) != hipSuccess) return e;
    if (n > 1) lbTree<<<blocks(n - 1), kBlock, 0, stream>>>(kIn, (int)n, left, right, first, last, parentNode, parentLeaf);
    if ((e = hipMemsetAsync(arrived, 0, (size_t)n * 4, stream)) != hipSuccess) return e;
    lbRefit<<<nb, kBlock, 0, stream>>>(vIn, triBox, (int)n, left, right, parentNode, parentLeaf, leafBox, nodeBox, arrived);
    lbRecords<<<nb, kBlock, 0, stream>>>(vtx, vtxOff, stride, faces, faceOff, vIn, n, records + (size_t)12 * recBase);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    // collapse, level by level from the root (binary node 0, or the single leaf)
    const uint2 rootJob = make_uint2(n > 1 ? 0u : kLeafBit, 0u);
    uint32_t init[3] = {1u, 0u, 0u};
    if ((e = hipMemcpyAsync(jobsA, &rootJob, sizeof(rootJob), hipMemcpyHostToDevice, stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(counters, init, sizeof(init), hipMemcpyHostToDevice, stream)) != hipSuccess) return e;
    uint32_t njobs = 1;
    int levels = 0;
    uint2* cur = jobsA; uint2* nxt = jobsB;
    while (njobs > 0) {
        ++levels;
        if ((e = hipMemsetAsync(counters + 1, 0, 4, stream)) != hipSuccess) return e;
        lbCollapse<<<blocks(njobs), kBlock, 0, stream>>>(cur, njobs, left, right, first, last, leafBox, nodeBox, leafMax, recBase, nodeBase,
                                                        nodeCap, nodes + nodeBase, counters, nxt, counters + 1, counters + 2);
        uint32_t c[3];
        if ((e = hipMemcpyAsync(c, counters, sizeof(c), hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        if (c[2]) return hipErrorInvalidValue;
        njobs = c[1];
        std::swap(cur, nxt);
        *nodeCount = (int)c[0];
    }
    *root = nodeBase;
    *depth = levels;
    return hipGetLastError();
}

}  // namespace omv
