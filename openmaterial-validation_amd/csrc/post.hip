// post.hip -- the film post pass of main.cpp:513-519 on the device:
// gammaCorrection(img, W, H, A, gamma) (ToneMapping.cpp:24-54), the 3x3
// median filter SimpleDenoiserMedian<1> (Denoiser.h:24-68) and the 8-bit
// truncation of saveImage (main.cpp:88-93).  One thread per pixel; the
// median reads a 3x3 neighbourhood of the gamma-corrected image.
#include "path_device.h"

// gammaCorrection: luminance-preserving power curve, then clamp to [0, 1]
__global__ void omv_gamma_kernel(const float4* __restrict__ in, float4* __restrict__ out, uint32_t npix, float A, float gamma) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    float4 p = in[i];
    const float inI = 0.2126f * p.x + 0.7152f * p.y + 0.0722f * p.z;
    const float outI = A * powf(inI, gamma);
    const float f = inI > FLT_EPSILON ? outI / inI : 1.0f;
    p.x = clamp01(f * p.x); p.y = clamp01(f * p.y); p.z = clamp01(f * p.z); p.w = 1.0f;
    out[i] = p;
}

namespace {
// median of 9 by a 19-exchange network (Paeth, Graphics Gems): the value of
// element 4 of the sorted window, as std::sort + [4] gives
DEV void cswap(float& a, float& b) { const float lo = fminf(a, b), hi = fmaxf(a, b); a = lo; b = hi; }
DEV float median9(float v[9]) {
    cswap(v[1], v[2]); cswap(v[4], v[5]); cswap(v[7], v[8]);
    cswap(v[0], v[1]); cswap(v[3], v[4]); cswap(v[6], v[7]);
    cswap(v[1], v[2]); cswap(v[4], v[5]); cswap(v[7], v[8]);
    cswap(v[0], v[3]); cswap(v[5], v[8]); cswap(v[4], v[7]);
    cswap(v[3], v[6]); cswap(v[1], v[4]); cswap(v[2], v[5]);
    cswap(v[4], v[7]); cswap(v[4], v[2]); cswap(v[6], v[4]);
    cswap(v[4], v[2]);
    return v[4];
}
}  // namespace

// SimpleDenoiserMedian<1>: interior pixels take the per-channel median of
// their 3x3 window, the one-pixel border keeps its value
__global__ void omv_median3_kernel(const float4* __restrict__ in, float4* __restrict__ out, int w, int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const int n = y * w + x;
    if (x < 1 || y < 1 || x >= w - 1 || y >= h - 1) { out[n] = in[n]; return; }
    float r[9], g[9], b[9];
    int k = 0;
    for (int yy = y - 1; yy <= y + 1; ++yy)
        for (int xx = x - 1; xx <= x + 1; ++xx) {
            const float4 p = in[yy * w + xx];
            r[k] = p.x; g[k] = p.y; b[k] = p.z;
            ++k;
        }
    float4 o;
    o.x = median9(r); o.y = median9(g); o.z = median9(b); o.w = 1.0f;
    out[n] = o;
}

// saveImage: static_cast<unsigned char>(255.0f * f) per channel
__global__ void omv_to_u8_kernel(const float4* __restrict__ in, uchar4* __restrict__ out, uint32_t npix) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const float4 p = in[i];
    uchar4 o;
    o.x = (unsigned char)(255.0f * p.x); o.y = (unsigned char)(255.0f * p.y);
    o.z = (unsigned char)(255.0f * p.z); o.w = (unsigned char)(255.0f * p.w);
    out[i] = o;
}

namespace omv {

hipError_t launchPost(const float* film, float* tmp, float* outF, uint8_t* outU8, int w, int h, float A, float gamma, bool denoise,
                      hipStream_t stream) {
    const uint32_t npix = (uint32_t)w * (uint32_t)h;
    const int tb = 256;
    const dim3 grid1((npix + tb - 1) / tb);
    float4* g = reinterpret_cast<float4*>(denoise ? tmp : outF);
    hipLaunchKernelGGL(omv_gamma_kernel, grid1, dim3(tb), 0, stream, reinterpret_cast<const float4*>(film), g, npix, A, gamma);
    if (denoise)
        hipLaunchKernelGGL(omv_median3_kernel, dim3((w + 63) / 64, h), dim3(64), 0, stream, reinterpret_cast<const float4*>(tmp),
                           reinterpret_cast<float4*>(outF), w, h);
    if (outU8)
        hipLaunchKernelGGL(omv_to_u8_kernel, grid1, dim3(tb), 0, stream, reinterpret_cast<const float4*>(outF),
                           reinterpret_cast<uchar4*>(outU8), npix);
    return hipGetLastError();
}

}  // namespace omv
