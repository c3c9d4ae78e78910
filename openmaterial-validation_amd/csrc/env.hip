// env.hip -- environment importance sampling tables (OMV_EXT_HDRI_NEE),
// built on the device at scene commit: texel weights (Rec.709 luminance x
// sine of the polar angle), one Vose alias table per row over its texels,
// one over the rows, and the texel probabilities.  Each table follows the
// sequential Vose method (sum in index order, LIFO work lists), so the tables
// equal, bit for bit, those the CPU oracle builds with the same operations.
// Tables of up to kWaveMax entries are built by one wave in LDS: the lanes
// scale the weights and split them into the small / large lists in index
// order (ballot + prefix count), lane 0 runs the sum and the pairing loop,
// and the lanes write the result; longer tables by one lane in global memory.
// The rows are built in parallel.
#include "path_device.h"

namespace {

// sin(polar angle) at row centre i by Bhaskara I's form, + - * / only
DEV float envRowSine(int i, int H) {
    const float pi = 3.14159265358979323846f;
    const float x = ((float)i + 0.5f) * pi / (float)H;
    const float a = x * (pi - x);
    return 16.0f * a / (5.0f * pi * pi - 4.0f * a);
}

DEV void vose(const float* w, int n, float* prob, int* alias, float* p, int* small, int* large) {
    float S = 0.0f;
    for (int k = 0; k < n; ++k) S = S + w[k];
    if (!(S > 0.0f)) {
        for (int k = 0; k < n; ++k) { prob[k] = 1.0f; alias[k] = k; }
        return;
    }
    const float scale = (float)n / S;
    int ns = 0, nl = 0;
    for (int k = 0; k < n; ++k) {
        p[k] = w[k] * scale;
        if (p[k] < 1.0f) small[ns++] = k; else large[nl++] = k;
    }
    while (ns > 0 && nl > 0) {
        const int l = small[--ns], g = large[--nl];
        prob[l] = p[l]; alias[l] = g;
        p[g] = (p[g] + p[l]) - 1.0f;
        if (p[g] < 1.0f) small[ns++] = g; else large[nl++] = g;
    }
    while (nl > 0) { const int g = large[--nl]; prob[g] = 1.0f; alias[g] = g; }
    while (ns > 0) { const int l = small[--ns]; prob[l] = 1.0f; alias[l] = l; }
}

constexpr int kWaveMax = 3264;   // 20 bytes of LDS per entry: <= 64 KB

// The Vose table of the n weights in p (LDS, overwritten) with total S, one
// wave; small / large / prob / alias are LDS arrays of n entries.
DEV void voseWave(float* p, int n, float S, int* small, int* large, float* prob, int* alias) {
    __shared__ int left[2];
    const int lane = threadIdx.x;
    if (!(S > 0.0f)) {
        for (int k = lane; k < n; k += 64) { prob[k] = 1.0f; alias[k] = k; }
        return;
    }
    const float scale = (float)n / S;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int ns = 0, nl = 0;   // the same on every lane
    for (int k0 = 0; k0 < n; k0 += 64) {
        const int k = k0 + lane;
        const bool valid = k < n;
        float v = 0.0f;
        if (valid) { v = p[k] * scale; p[k] = v; }
        const bool sm = valid && v < 1.0f, lg = valid && !(v < 1.0f);
        const uint64_t ms = __ballot(sm), ml = __ballot(lg);
        if (sm) small[ns + __popcll(ms & below)] = k;
        if (lg) large[nl + __popcll(ml & below)] = k;
        ns += __popcll(ms);
        nl += __popcll(ml);
    }
    __syncthreads();
    if (lane == 0) {
        while (ns > 0 && nl > 0) {
            const int l = small[--ns], g = large[--nl];
            prob[l] = p[l]; alias[l] = g;
            p[g] = (p[g] + p[l]) - 1.0f;
            if (p[g] < 1.0f) small[ns++] = g; else large[nl++] = g;
        }
        left[0] = ns; left[1] = nl;
    }
    __syncthreads();
    ns = left[0]; nl = left[1];
    // what is left gets probability 1 (each index is on one list at most once)
    for (int k = lane; k < nl; k += 64) { const int g = large[k]; prob[g] = 1.0f; alias[g] = g; }
    for (int k = lane; k < ns; k += 64) { const int l = small[k]; prob[l] = 1.0f; alias[l] = l; }
}

}  // namespace

// weights and column tables, W <= kWaveMax: one wave per row
__global__ __launch_bounds__(64) void omv_env_rows_wave_kernel(const float4* __restrict__ bg, int W, int H, float* __restrict__ w,
                                                               float* __restrict__ rowW, float* __restrict__ colProb,
                                                               int* __restrict__ colAlias) {
    extern __shared__ float lds[];
    __shared__ float sum;
    float* p = lds;
    int* small = reinterpret_cast<int*>(p + W);
    int* large = small + W;
    float* prob = reinterpret_cast<float*>(large + W);
    int* alias = reinterpret_cast<int*>(prob + W);
    const int i = blockIdx.x, lane = threadIdx.x;
    const float sn = envRowSine(i, H);
    const size_t o = (size_t)i * W;
    for (int j = lane; j < W; j += 64) {
        const float4 t = bg[o + j];
        const float lum = 0.2126f * t.x + 0.7152f * t.y + 0.0722f * t.z;
        const float v = (lum > 0.0f ? lum : 0.0f) * sn;
        w[o + j] = v;
        p[j] = v;
    }
    __syncthreads();
    if (lane == 0) {
        float rs = 0.0f;
        for (int j = 0; j < W; ++j) rs = rs + p[j];
        sum = rs;
        rowW[i] = rs;
    }
    __syncthreads();
    voseWave(p, W, sum, small, large, prob, alias);
    __syncthreads();
    for (int j = lane; j < W; j += 64) { colProb[o + j] = prob[j]; colAlias[o + j] = alias[j]; }
}

// the row table and the total, H <= kWaveMax: one wave
__global__ __launch_bounds__(64) void omv_env_marginal_wave_kernel(const float* __restrict__ rowW, int H, float* __restrict__ rowProb,
                                                                   int* __restrict__ rowAlias, float* __restrict__ total) {
    extern __shared__ float lds[];
    __shared__ float sum;
    float* p = lds;
    int* small = reinterpret_cast<int*>(p + H);
    int* large = small + H;
    float* prob = reinterpret_cast<float*>(large + H);
    int* alias = reinterpret_cast<int*>(prob + H);
    const int lane = threadIdx.x;
    for (int i = lane; i < H; i += 64) p[i] = rowW[i];
    __syncthreads();
    if (lane == 0) {
        float t = 0.0f;
        for (int i = 0; i < H; ++i) t = t + p[i];
        sum = t;
        *total = t;
    }
    __syncthreads();
    voseWave(p, H, sum, small, large, prob, alias);
    __syncthreads();
    for (int i = lane; i < H; i += 64) { rowProb[i] = prob[i]; rowAlias[i] = alias[i]; }
}

// weights and column tables, W > kWaveMax: one lane per row
__global__ void omv_env_rows_kernel(const float4* __restrict__ bg, int W, int H, float* __restrict__ w, float* __restrict__ rowW,
                                    float* __restrict__ colProb, int* __restrict__ colAlias, float* __restrict__ p,
                                    int* __restrict__ small, int* __restrict__ large) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H) return;
    const float sn = envRowSine(i, H);
    const size_t o = (size_t)i * W;
    float rs = 0.0f;
    for (int j = 0; j < W; ++j) {
        const float4 t = bg[o + j];
        const float lum = 0.2126f * t.x + 0.7152f * t.y + 0.0722f * t.z;
        const float v = (lum > 0.0f ? lum : 0.0f) * sn;
        w[o + j] = v;
        rs = rs + v;
    }
    rowW[i] = rs;
    vose(w + o, W, colProb + o, colAlias + o, p + o, small + o, large + o);
}

// the row table and the total, H > kWaveMax: one lane
__global__ void omv_env_marginal_kernel(const float* __restrict__ rowW, int H, float* __restrict__ rowProb, int* __restrict__ rowAlias,
                                        float* __restrict__ p, int* __restrict__ small, int* __restrict__ large, float* __restrict__ total) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float t = 0.0f;
    for (int i = 0; i < H; ++i) t = t + rowW[i];
    *total = t;
    vose(rowW, H, rowProb, rowAlias, p, small, large);
}

// texel probabilities w / total
__global__ void omv_env_pdf_kernel(const float* __restrict__ w, const float* __restrict__ total, uint32_t n, float* __restrict__ pdf) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float t = *total;
    pdf[k] = t > 0.0f ? w[k] / t : 0.0f;
}

namespace omv {

// scratch: 4 * W * H + 2 * H + 2 words; tables: rowProb/rowAlias (H), colProb/colAlias/texPdf (W*H)
hipError_t launchEnvBuild(const float* bg, int W, int H, float* rowProb, int* rowAlias, float* colProb, int* colAlias, float* texPdf,
                          void* scratch, hipStream_t stream) {
    const size_t n = (size_t)W * H;
    float* w = static_cast<float*>(scratch);
    float* p = w + n;
    int* small = reinterpret_cast<int*>(p + n);
    int* large = small + n;
    float* rowW = reinterpret_cast<float*>(large + n);
    float* total = rowW + H;
    if (W <= kWaveMax)
        hipLaunchKernelGGL(omv_env_rows_wave_kernel, dim3(H), dim3(64), (size_t)20 * W, stream, reinterpret_cast<const float4*>(bg), W, H, w,
                           rowW, colProb, colAlias);
    else
        hipLaunchKernelGGL(omv_env_rows_kernel, dim3((H + 63) / 64), dim3(64), 0, stream, reinterpret_cast<const float4*>(bg), W, H, w, rowW,
                           colProb, colAlias, p, small, large);
    if (H <= kWaveMax)
        hipLaunchKernelGGL(omv_env_marginal_wave_kernel, dim3(1), dim3(64), (size_t)20 * H, stream, rowW, H, rowProb, rowAlias, total);
    else
        hipLaunchKernelGGL(omv_env_marginal_kernel, dim3(1), dim3(64), 0, stream, rowW, H, rowProb, rowAlias, p, small, large, total);
    hipLaunchKernelGGL(omv_env_pdf_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, w, total, (uint32_t)n, texPdf);
    return hipGetLastError();
}

}  // namespace omv
