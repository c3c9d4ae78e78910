// mesh_device.hip -- commitMesh's normals and tangents on the device
// (RenderMesh::computeNormals / computeTangents / randomTangent,
// RenderMesh.cpp:300-432), used by the device build mode.
//
// The reference sums face contributions into vertices in one sequential
// loop over faces.  Here the corners (3 * face + k) are sorted by vertex
// with a stable radix sort, so every vertex visits its corners in ascending
// face order and adds the same float values in the same order as the host
// loop: normals and tangents equal the host's bit for bit.
//
//   face pass     unit face normal (cross of the edges p1-p0, p2-p1,
//                 normalised when its square exceeds FLT_EPSILON); for uv
//                 tangents the uv-derivative frame (skipped when the uv
//                 determinant is below FLT_EPSILON)
//   vertex pass   normal = normalize((0, 0, eps) + sum of its faces'
//                 normals); tangent = Gram-Schmidt of the summed u frame
//                 against the normal (axis tangent when it vanishes),
//                 w = sign of (n x t) . summed v frame; or the axis tangent.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "device_sort.h"
#include "omv_kernels.h"

namespace {

constexpr int kBlk = 256;

struct f3 { float x, y, z; };
__device__ inline f3 mk3(float x, float y, float z) { return {x, y, z}; }
__device__ inline f3 crossf(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ inline float dotf(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float lenf(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ inline f3 nrmz(f3 a) {   // normalize3: multiply by 1 / |a|, (1, 0, 0) for a zero vector
    float L = lenf(a);
    if (L != 0.0f) { L = 1.0f / L; return mk3(a.x * L, a.y * L, a.z * L); }
    return mk3(1.0f, 0.0f, 0.0f);
}
__device__ inline void axisTangent(f3 n, float* t) {   // randomTangent, scale 1
    f3 r;
    if (n.z < 0.95f && n.z > -0.95f) r = crossf(mk3(0.0f, 0.0f, 1.0f), n);
    else if (n.y < 0.95f && n.y > -0.95f) r = crossf(mk3(0.0f, 1.0f, 0.0f), n);
    else r = crossf(mk3(1.0f, 0.0f, 0.0f), n);
    r = nrmz(r);
    t[0] = r.x * 1.0f; t[1] = r.y * 1.0f; t[2] = r.z * 1.0f; t[3] = 1.0f;
}

__global__ void mdCorners(const uint32_t* __restrict__ faces, uint32_t corners, uint32_t vn, uint32_t* __restrict__ key,
                          uint32_t* __restrict__ val, uint32_t* __restrict__ bad) {
    const uint32_t c = blockIdx.x * kBlk + threadIdx.x;
    if (c >= corners) return;
    const uint32_t v = faces[c];
    if (v >= vn) atomicAdd(bad, 1u);
    key[c] = v < vn ? v : vn;   // out-of-range ids sort past the last vertex and are never visited
    val[c] = c;
}

// corner range of every vertex in the sorted order
__global__ void mdRanges(const uint32_t* __restrict__ key, uint32_t corners, uint32_t vn, uint32_t* __restrict__ begin,
                         uint32_t* __restrict__ end) {
    const uint32_t c = blockIdx.x * kBlk + threadIdx.x;
    if (c >= corners) return;
    const uint32_t v = key[c];
    if (v >= vn) return;
    if (c == 0 || key[c - 1] != v) begin[v] = c;
    if (c + 1 == corners || key[c + 1] != v) end[v] = c + 1;
}

__global__ void mdFaceNormals(const float* __restrict__ V, uint32_t stride, const uint32_t* __restrict__ faces, uint32_t fn, uint32_t vn,
                              float* __restrict__ fnrm) {
    const uint32_t f = blockIdx.x * kBlk + threadIdx.x;
    if (f >= fn) return;
    const uint32_t i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
    f3 N = mk3(0.0f, 0.0f, 0.0f);
    if (i0 < vn && i1 < vn && i2 < vn) {
        const float* p0 = V + (size_t)i0 * stride;
        const float* p1 = V + (size_t)i1 * stride;
        const float* p2 = V + (size_t)i2 * stride;
        N = crossf(mk3(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]), mk3(p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]));
        if (N.x * N.x + N.y * N.y + N.z * N.z > FLT_EPSILON) N = nrmz(N);
    }
    fnrm[3 * f] = N.x; fnrm[3 * f + 1] = N.y; fnrm[3 * f + 2] = N.z;
}

__global__ void mdVertexNormals(float* __restrict__ V, uint32_t stride, uint32_t vn, const uint32_t* __restrict__ corner,
                                const uint32_t* __restrict__ begin, const uint32_t* __restrict__ end, const float* __restrict__ fnrm) {
    const uint32_t v = blockIdx.x * kBlk + threadIdx.x;
    if (v >= vn) return;
    float n0 = 0.0f, n1 = 0.0f, n2 = FLT_EPSILON;
    for (uint32_t c = begin[v]; c < end[v]; ++c) {
        const uint32_t f = corner[c] / 3u;
        n0 += fnrm[3 * f]; n1 += fnrm[3 * f + 1]; n2 += fnrm[3 * f + 2];
    }
    const f3 u = nrmz(mk3(n0, n1, n2));
    float* n = V + (size_t)v * stride + 3;
    n[0] = u.x; n[1] = u.y; n[2] = u.z;
}

// per-face uv frame (ud, vd) and a valid flag
__global__ void mdFaceFrames(const float* __restrict__ V, uint32_t stride, int uvOff, const uint32_t* __restrict__ faces, uint32_t fn,
                             uint32_t vn, float* __restrict__ frame) {
    const uint32_t f = blockIdx.x * kBlk + threadIdx.x;
    if (f >= fn) return;
    float* o = frame + (size_t)7 * f;
    o[6] = 0.0f;
    const uint32_t i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
    if (i0 >= vn || i1 >= vn || i2 >= vn) return;
    const float* X0 = V + (size_t)i0 * stride;
    const float* X1 = V + (size_t)i1 * stride;
    const float* X2 = V + (size_t)i2 * stride;
    const float d1[3] = {X1[0] - X0[0], X1[1] - X0[1], X1[2] - X0[2]};
    const float d2[3] = {X2[0] - X0[0], X2[1] - X0[1], X2[2] - X0[2]};
    const float s1[2] = {X1[uvOff] - X0[uvOff], X1[uvOff + 1] - X0[uvOff + 1]};
    const float s2[2] = {X2[uvOff] - X0[uvOff], X2[uvOff + 1] - X0[uvOff + 1]};
    const float det = s1[0] * s2[1] - s2[0] * s1[1];
    if (det < FLT_EPSILON) return;
    const float r = 1.0f / det;
    const float a00 = r * s2[1], a01 = -r * s2[0], a10 = -r * s1[1], a11 = r * s1[0];
    for (int c = 0; c < 3; ++c) {
        o[c] = d1[c] * a00 + d2[c] * a10;
        o[3 + c] = d1[c] * a01 + d2[c] * a11;
    }
    o[6] = 1.0f;
}

__global__ void mdVertexTangents(float* __restrict__ V, uint32_t stride, uint32_t vn, const uint32_t* __restrict__ corner,
                                 const uint32_t* __restrict__ begin, const uint32_t* __restrict__ end, const float* __restrict__ frame) {
    const uint32_t v = blockIdx.x * kBlk + threadIdx.x;
    if (v >= vn) return;
    float tv[3] = {0.0f, 0.0f, 0.0f}, bv[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t c = begin[v]; c < end[v]; ++c) {
        const float* o = frame + (size_t)7 * (corner[c] / 3u);
        if (o[6] == 0.0f) continue;
        for (int k = 0; k < 3; ++k) { tv[k] += o[k]; bv[k] += o[3 + k]; }
    }
    float* t = V + (size_t)v * stride + 6;
    const float* n = V + (size_t)v * stride + 3;
    const f3 nv = mk3(n[0], n[1], n[2]);
    const float proj = dotf(nv, mk3(tv[0], tv[1], tv[2]));
    t[0] = tv[0] - n[0] * proj;
    t[1] = tv[1] - n[1] * proj;
    t[2] = tv[2] - n[2] * proj;
    const float l = lenf(mk3(t[0], t[1], t[2]));
    if (l <= FLT_EPSILON) axisTangent(nv, t);
    else { t[0] /= l; t[1] /= l; t[2] /= l; }
    const f3 nt = crossf(nv, mk3(t[0], t[1], t[2]));
    t[3] = dotf(nt, mk3(bv[0], bv[1], bv[2])) < 0.0f ? -1.0f : 1.0f;
}

__global__ void mdAxisTangents(float* __restrict__ V, uint32_t stride, uint32_t vn) {
    const uint32_t v = blockIdx.x * kBlk + threadIdx.x;
    if (v >= vn) return;
    const float* n = V + (size_t)v * stride + 3;
    axisTangent(mk3(n[0], n[1], n[2]), V + (size_t)v * stride + 6);
}

inline uint32_t grid(uint64_t n) { return (uint32_t)((n + kBlk - 1) / kBlk); }

}  // namespace

namespace omv {

size_t meshDeviceScratchBytes(uint32_t vn, uint32_t fn) {
    const uint64_t corners = 3ull * fn;
    const uint64_t words = 4 * corners + 256ull * sortChunks((uint32_t)corners) + 2ull * vn + 7ull * fn + 8 + 2 * 9;
    return words * 4 + 256;
}

// vertices: this mesh's vertex pool slice (stride floats per vertex,
// position at 0, normal at 3, tangent at 6); uvOff >= 0 selects uv tangents
// from that float offset, < 0 axis tangents
hipError_t preprocessMeshDevice(float* vertices, uint32_t stride, uint32_t vn, const uint32_t* faces, uint32_t fn, bool normals,
                                bool tangents, int uvOff, void* scratch, hipStream_t stream, uint32_t* badIndices) {
    *badIndices = 0;
    if (vn == 0 || fn == 0 || (!normals && !tangents)) return hipSuccess;
    const uint32_t corners = 3 * fn;
    uint32_t* p = static_cast<uint32_t*>(scratch);
    auto take = [&p](uint64_t words) { uint32_t* q = p; p += (words + 1) & ~(uint64_t)1; return q; };
    uint32_t* keyA = take(corners);
    uint32_t* valA = take(corners);
    uint32_t* keyB = take(corners);
    uint32_t* valB = take(corners);
    uint32_t* hist = take(256ull * sortChunks(corners));
    uint32_t* begin = take(vn);
    uint32_t* end = take(vn);
    float* faceData = reinterpret_cast<float*>(take(7ull * fn));
    uint32_t* bad = take(2);
    hipError_t e;
    if ((e = hipMemsetAsync(bad, 0, 4, stream)) != hipSuccess) return e;
    mdCorners<<<grid(corners), kBlk, 0, stream>>>(faces, corners, vn, keyA, valA, bad);
    if ((e = sortPairs32(keyA, valA, keyB, valB, hist, corners, stream)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(begin, 0, (size_t)vn * 4, stream)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(end, 0, (size_t)vn * 4, stream)) != hipSuccess) return e;
    mdRanges<<<grid(corners), kBlk, 0, stream>>>(keyA, corners, vn, begin, end);
    if (normals) {
        mdFaceNormals<<<grid(fn), kBlk, 0, stream>>>(vertices, stride, faces, fn, vn, faceData);
        mdVertexNormals<<<grid(vn), kBlk, 0, stream>>>(vertices, stride, vn, valA, begin, end, faceData);
    }
    if (tangents) {
        if (uvOff >= 0) {
            mdFaceFrames<<<grid(fn), kBlk, 0, stream>>>(vertices, stride, uvOff, faces, fn, vn, faceData);
            mdVertexTangents<<<grid(vn), kBlk, 0, stream>>>(vertices, stride, vn, valA, begin, end, faceData);
        } else {
            mdAxisTangents<<<grid(vn), kBlk, 0, stream>>>(vertices, stride, vn);
        }
    }
    if ((e = hipMemcpyAsync(badIndices, bad, 4, hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
    return hipGetLastError();
}

}  // namespace omv
