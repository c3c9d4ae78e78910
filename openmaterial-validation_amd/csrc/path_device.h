// path_device.h -- device-side building blocks of the gfx950 render core,
// shared by the sample kernels (kernels.hip) and the wavefront kernels
// (wavefront.hip): float math in the reference's operation order, Philox
// streams, warps, BSDFs, complex-IOR Fresnel, textures, surface points,
// materials, two-level BVH4 traversal with the watertight triangle test,
// and the path helpers (attenuation, visibility, background, lights, camera).
//
// Arithmetic follows the reference's operation order with FMA contraction
// off and correctly rounded division and square root, so the ray/hit/id path
// is bit-identical to the CPU oracle.
#ifndef OMV_PATH_DEVICE_H
#define OMV_PATH_DEVICE_H

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "../../include/omv.h"
#include "omv_types.h"
#include "omv_kernels.h"

#pragma clang fp contract(off)

#define DEV __device__ __forceinline__

// Dynamic LDS of every path kernel: the traversal stacks, [entry][lane].
extern __shared__ uint32_t omv_lds_stack[];

namespace {

// ------------------------------------------------------------------ math
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };
DEV f3 mk3(float a, float b, float c) { f3 r; r.x = a; r.y = b; r.z = c; return r; }
// UtilityFunctions.h:128-247, same operation order
DEV float dot3(const f3& a, const f3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEV f3 cross3(const f3& a, const f3& b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
DEV f3 sub3(const f3& a, const f3& b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
DEV f3 neg3(const f3& a) { return mk3(-a.x, -a.y, -a.z); }
DEV float length3(const f3& a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
DEV f3 normalize3(const f3& a) {
    float L = length3(a);
    if (L != 0.0f) { L = 1.0f / L; return mk3(a.x * L, a.y * L, a.z * L); }
    return mk3(1.0f, 0.0f, 0.0f);
}
DEV f3 normalizeIfNeeded3(const f3& a) {
    const float eps = 10.0f * FLT_EPSILON;
    float sL = dot3(a, a);
    if (sL < 1.0f + eps && sL > 1.0f - eps) return a;
    float L = sqrtf(sL);
    if (L != 0.0f) { L = 1.0f / L; return mk3(a.x * L, a.y * L, a.z * L); }
    return mk3(1.0f, 0.0f, 0.0f);
}
DEV float mnf(float a, float b) { return a < b ? a : b; }
DEV float mxf(float a, float b) { return a > b ? a : b; }
DEV float clampr(float v, float lo, float hi) { return mxf(lo, mnf(v, hi)); }  // reference clamp()
DEV float clamp01(float v) { return clampr(v, 0.0f, 1.0f); }
DEV float lerpr(float lo, float hi, float t) { return (hi - lo) * t + lo; }    // reference lerp()
DEV float stdmax(float a, float b) { return (a < b) ? b : a; }                 // std::max
DEV f3 reflect3(const f3& I, const f3& N) {
    const float w = 2.0f * dot3(I, N);
    return mk3(I.x - w * N.x, I.y - w * N.y, I.z - w * N.z);
}
DEV bool refract3(const f3& I, const f3& N, float eta, f3& out) {
    float NI = dot3(N, I);
    float k = 1.0f - eta * eta * (1.0f - NI * NI);
    if (k < 0.0f) { out = neg3(I); return false; }
    k = eta * NI + sqrtf(k);
    out = mk3(eta * I.x - k * N.x, eta * I.y - k * N.y, eta * I.z - k * N.z);
    return true;
}
DEV f3 mul33(const float* A, const f3& v) {
    return mk3(A[0] * v.x + A[1] * v.y + A[2] * v.z, A[3] * v.x + A[4] * v.y + A[5] * v.z, A[6] * v.x + A[7] * v.y + A[8] * v.z);
}
DEV f3 mul34(const float* A, const f3& v) {
    return mk3(A[0] * v.x + A[1] * v.y + A[2] * v.z + A[3], A[4] * v.x + A[5] * v.y + A[6] * v.z + A[7],
               A[8] * v.x + A[9] * v.y + A[10] * v.z + A[11]);
}
DEV f3 mulT33(const f3& v, const float* A) {
    return mk3(A[0] * v.x + A[3] * v.y + A[6] * v.z, A[1] * v.x + A[4] * v.y + A[7] * v.z, A[2] * v.x + A[5] * v.y + A[8] * v.z);
}
DEV f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
DEV float intensity(const f3& r) { return 0.299f * r.x + 0.587f * r.y + 0.114f * r.z; }  // BDPTRenderer.cpp:52-55
DEV float maxAbs(const f3& v) {                                                          // BDPTRenderer.cpp:35-40
    float a = fabsf(v.x), b = fabsf(v.y), c = fabsf(v.z);
    float m = (a > b && a > c) ? a : (b > c ? b : c);
    return m < 1.0f ? 1.0f : m;
}
constexpr float kPi = 3.14159265358979323846f;

// ------------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon et al., SC'11).  Key (seed, 0); counter
// (call, sub-stream, sample lo, sample hi); 4 uniforms per call.
struct Rng {
    uint32_t k0, c0, c1, c2, c3;
    uint32_t b0, b1, b2, b3;
    int idx;
    DEV void init(uint32_t seed, uint32_t stream, uint64_t sample) {
        k0 = seed; c0 = 0; c1 = stream; c2 = (uint32_t)sample; c3 = (uint32_t)(sample >> 32); idx = 4;
        b0 = b1 = b2 = b3 = 0;
    }
    DEV void refill() {
        uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3, ka = k0, kb = 0;
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            if (r) { ka += 0x9E3779B9u; kb += 0xBB67AE85u; }
            // one 32x32 -> 64 product each (v_mad_u64_u32) instead of a
            // separate low and high multiply
            const uint64_t p0 = (uint64_t)0xD2511F53u * x0, p1 = (uint64_t)0xCD9E8D57u * x2;
            uint32_t n0 = (uint32_t)(p1 >> 32) ^ x1 ^ ka, n2 = (uint32_t)(p0 >> 32) ^ x3 ^ kb;
            x0 = n0; x1 = (uint32_t)p1; x2 = n2; x3 = (uint32_t)p0;
        }
        b0 = x0; b1 = x1; b2 = x2; b3 = x3;
        ++c0; idx = 0;
    }
    // The block is kept rotated so that b0 is always the next draw (idx
    // counts the draws taken from it): selecting b[idx] by a per-lane idx
    // would compile to a branchy switch.
    DEV float next() {
        if (idx == 4) refill();
        const uint32_t v = b0;
        b0 = b1; b1 = b2; b2 = b3;
        ++idx;
        return (float)(v >> 8) * (1.0f / 16777216.0f);
    }
};
enum { STREAM_CAMERA = 0, STREAM_LIGHT = 1, STREAM_NEE = 2 };   // NEE at camera vertex i: 2 + 16 * i

// ----------------------------------------------------------------- warps
// RandomSampler.h:66-103
DEV f3 uniformHemisphere(Rng& r) {
    float phi = lerpr(0.0f, 2.0f * kPi, r.next());
    float cth = r.next();
    float sth = sqrtf(stdmax(1.0f - cth * cth, 0.0f));
    return mk3(sth * cosf(phi), sth * sinf(phi), cth);
}
DEV f3 uniformSphere(Rng& r) {   // positive octant only, bounded at 256 rounds
    f3 v; float sl = 0.0f; int tries = 0;
    do {
        if (++tries > 256) { v = mk3(1.0f, 1.0f, 1.0f); sl = 3.0f; break; }
        v.x = r.next(); v.y = r.next(); v.z = r.next();
        sl = dot3(v, v);
    } while (sl > 1.0f || sl < FLT_EPSILON);
    float s = 1.0f / sqrtf(sl);
    return mk3(s * v.x, s * v.y, s * v.z);
}
DEV f3 ggxHemisphere(Rng& r, float alpha2) {
    float phi = lerpr(0.0f, 2.0f * kPi, r.next());
    float rnd = r.next();
    float cth = sqrtf((1.0f - rnd) / (1.0f - (1.0f - alpha2) * rnd));
    float sth = sqrtf(stdmax(1.0f - cth * cth, 0.0f));
    return mk3(sth * cosf(phi), sth * sinf(phi), cth);
}
DEV f3 cosineSqrtHemisphere(Rng& r) {
    float phi = lerpr(0.0f, 2.0f * kPi, r.next());
    float rnd = r.next();
    float sth = sqrtf(rnd);
    return mk3(sth * cosf(phi), sth * sinf(phi), sqrtf(1.0f - rnd));
}
// transformToNormalSpace (BRDFModel.h:25-34)
DEV f3 toBasis(const f3& b0, const f3& b1, const f3& b2, const f3& c) {
    f3 r = mk3(b0.x * c.x + 0.0f, b0.y * c.x + 0.0f, b0.z * c.x + 0.0f);
    r = mk3(b1.x * c.y + r.x, b1.y * c.y + r.y, b1.z * c.y + r.z);
    return mk3(b2.x * c.z + r.x, b2.y * c.z + r.y, b2.z * c.z + r.z);
}

// ---------------------------------------------------------------- BSDF
// BRDFModel.h:40-80
DEV float ggxPdf(float c, float alpha) {
    float c2 = c * c, a2 = alpha * alpha;
    return a2 / (kPi * powf(lerpr(1.0f, a2, c2), 2.0f));
}
DEV float ggxLambda(float a) { return 0.5f * (sqrtf(1.0f + powf(a, -2.0f)) - 1.0f); }
DEV float smithG1(float c, float alpha) {
    float a = c / (alpha * sqrtf(1.0f - c * c));
    return (alpha > 0.0f || c < 0.99f) ? 1.0f / (1.0f + ggxLambda(a)) : 1.0f;
}
DEV float smithG2(float s, float m) { float p = s * m; return p / (s + m - p); }

struct Bsdf {
    f3 I, O, N, H, T;
    float c0, c1, c2, c3;   // base colour rgba
    bool transmissive;
    float metal, rough, alpha, eta;
    int lobe;               // sampled lobe: 0 none, 1 diffuse, 2 reflection, 3 transmission
};
DEV void bsdfInit(Bsdf& b) {
    b.I = b.O = b.N = b.H = b.T = mk3(0.0f, 0.0f, 0.0f);
    b.c0 = b.c1 = b.c2 = b.c3 = 1.0f; b.transmissive = false;
    b.metal = 0.5f; b.rough = 0.0f; b.alpha = 0.25f; b.eta = 1.0f; b.lobe = 0;
}
// BSDFModel::evaluate (BRDFModel.h:139-207)
DEV f3 bsdfEvaluate(const f3& vIn, const f3& vOut, const f3& N, const Bsdf& b, bool specular, bool transmitted, const f3& Hin, float P) {
    f3 H = transmitted ? Hin : normalize3(sub3(vOut, vIn));
    float VdotN = clampr(fabsf(dot3(vIn, N)), 0.0001f, 1.0f);
    float LdotN = clampr(fabsf(dot3(vOut, N)), 0.0001f, 1.0f);
    float NdotH = fabsf(dot3(N, H));
    float VdotH = fabsf(dot3(vIn, H));
    float LdotH = fabsf(dot3(vOut, H));
    float p5 = powf(clamp01(1.0f - VdotH), 5.0f);
    float col[3] = {b.c0, b.c1, b.c2};
    float F[3], fd[3], fs[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float cd = (1.0f - b.metal) * col[i];
        float F0 = lerpr(0.04f, col[i], b.metal);
        F[i] = lerpr(p5, 1.0f, F0);
        fd[i] = (1.0f - F[i]) * cd / kPi;
    }
    float G = smithG2(smithG1(VdotN, b.alpha), smithG1(LdotN, b.alpha));
#pragma unroll
    for (int i = 0; i < 3; ++i) fs[i] = clampr(0.25f * (transmitted ? 1.0f - F[i] : F[i]) * G / (VdotN * LdotN), 0.0f, FLT_MAX);
    if (specular) { float k = 4.0f * LdotH * VdotN / NdotH / LdotN / (1.0f - P); return mk3(k * fs[0], k * fs[1], k * fs[2]); }
    if (transmitted) { float k = 4.0f * LdotH * VdotN / NdotH / LdotN / P; return mk3(k * fs[0], k * fs[1], k * fs[2]); }
    float k = kPi / LdotN / P;
    return mk3(k * fd[0], k * fd[1], k * fd[2]);
}
// BSDFModel::evaluateDirect (BRDFModel.h:211-261)
DEV f3 evaluateDirect(const Bsdf& d) {
    float VdotN = -dot3(d.I, d.N), LdotN = dot3(d.O, d.N);
    f3 H = normalize3(sub3(d.O, d.I));
    float VdotH = fabsf(dot3(d.I, H)), NdotH = fabsf(dot3(d.N, H));
    if (!(VdotN > 0.0f && LdotN > 0.0f)) return mk3(0.0f, 0.0f, 0.0f);
    float p5 = powf(clamp01(1.0f - VdotH), 5.0f);
    float col[3] = {d.c0, d.c1, d.c2};
    float D = ggxPdf(clamp01(NdotH), d.alpha);
    float G = smithG2(smithG1(VdotN, d.alpha), smithG1(LdotN, d.alpha));
    float out[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float F0 = lerpr(0.04f, col[i], d.metal);
        float F = lerpr(p5, 1.0f, F0);
        float cd = (1.0f - d.metal) * col[i];
        float fdv = (1.0f - F) * cd / kPi;
        float fsv = clampr(0.25f * F * D * G / (VdotN * LdotN), 0.0f, FLT_MAX);
        out[i] = fdv + fsv;
    }
    return mk3(out[0], out[1], out[2]);
}
// BSDFModel::sampleIndirect (BRDFModel.h:265-337)
DEV f3 sampleIndirect(Bsdf& d, Rng& r) {
    float VdotN = -dot3(d.I, d.N);
    f3 binorm = cross3(d.N, d.T);
    f3 b0 = d.T, b1 = binorm;
    if (VdotN < 0.0f) {
        if (d.transmissive) { f3 t = b0; b0 = b1; b1 = t; d.N = neg3(d.N); }
        else return mk3(0.0f, 0.0f, 0.0f);
    }
    const float P = 0.8f * fabsf(VdotN) + 0.1f;
    float rnd = r.next();
    if (rnd < P) {
        if (rnd < P * (1.0f - d.c3)) {
            if (VdotN < 0.0f) d.eta = 1.0f / d.eta;
            f3 rd = ggxHemisphere(r, powf(d.alpha, 2.0f));
            d.H = normalize3(toBasis(b0, b1, d.N, rd));
            if (refract3(d.I, d.H, d.eta, d.O) && dot3(d.N, d.O) < 0.0f) {
                d.O = normalize3(d.O);
                d.lobe = 3;
                return bsdfEvaluate(d.I, d.O, d.N, d, false, true, d.H, P);
            }
            return mk3(0.0f, 0.0f, 0.0f);
        }
        f3 rd = cosineSqrtHemisphere(r);
        d.O = normalize3(toBasis(b0, b1, d.N, rd));
        d.lobe = 1;
        return bsdfEvaluate(d.I, d.O, d.N, d, false, false, d.H, P);
    }
    f3 rd = ggxHemisphere(r, powf(d.alpha, 2.0f));
    d.H = normalize3(toBasis(b0, b1, d.N, rd));
    d.O = reflect3(d.I, d.H);
    if (dot3(d.N, d.O) > 0.0f) {
        d.O = normalize3(d.O);
        d.lobe = 2;
        return bsdfEvaluate(d.I, d.O, d.N, d, true, false, d.H, P);
    }
    return mk3(0.0f, 0.0f, 0.0f);
}

// ----------------------------------------------- detached lobe derivatives
// With the sampled lobe and direction held fixed, a PBR bounce weighs
// w_ref(theta0) f(theta) / f(theta0), f being the sampled lobe's integrand up
// to material-independent factors: diffuse (1-F)(1-metal)color, reflection
// F*D*G, transmission (1-F)*D*G, on bsdfEvaluate's geometry.  These give
// d ln f / d(color_c, metal, alpha) for the path-replay adjoint.
struct LobeGeo { int lobe; float p5, NH, VN, LN; };
DEV LobeGeo lobeGeo(const Bsdf& d) {
    LobeGeo g;
    g.lobe = d.lobe;
    f3 H = d.lobe == 3 ? d.H : normalize3(sub3(d.O, d.I));
    g.VN = clampr(fabsf(dot3(d.I, d.N)), 0.0001f, 1.0f);
    g.LN = clampr(fabsf(dot3(d.O, d.N)), 0.0001f, 1.0f);
    g.NH = fabsf(dot3(d.N, H));
    float VH = fabsf(dot3(d.I, H));
    g.p5 = powf(clamp01(1.0f - VH), 5.0f);
    return g;
}
DEV float dlnSmithG1(float c, float alpha) {
    if (!(alpha > 0.0f || c < 0.99f)) return 0.0f;
    float a = c / (alpha * sqrtf(1.0f - c * c));
    float ai2 = powf(a, -2.0f);
    float lam = 0.5f * (sqrtf(1.0f + ai2) - 1.0f);
    float dlam = ai2 / (2.0f * alpha * sqrtf(1.0f + ai2));
    return -dlam / (1.0f + lam);
}
DEV void lobeDlog(const LobeGeo& g, const float color[3], float metal, float alpha, float dC[3], float dM[3], float dA[3]) {
    float dlnDG = 0.0f, DG = 1.0f;
    if (g.lobe >= 2) {
        float c2 = g.NH * g.NH;
        float L = lerpr(1.0f, alpha * alpha, c2);
        float dlnD = 2.0f / alpha - 4.0f * alpha * c2 / L;
        float x = smithG1(g.VN, alpha), y = smithG1(g.LN, alpha);
        float dx = x * dlnSmithG1(g.VN, alpha), dy = y * dlnSmithG1(g.LN, alpha);
        float den = x + y - x * y;
        dlnDG = dlnD + dx / x + dy / y - (dx + dy - dx * y - x * dy) / den;
        DG = ggxPdf(g.NH, alpha) * smithG2(x, y);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        dC[c] = dM[c] = dA[c] = 0.0f;
        float F0 = lerpr(0.04f, color[c], metal);
        float F = lerpr(g.p5, 1.0f, F0);
        float f = g.lobe == 1 ? (1.0f - F) * (1.0f - metal) * color[c] : (g.lobe == 2 ? F * DG : (1.0f - F) * DG);
        if (g.lobe == 0 || !(f != 0.0f)) continue;
        float dFdC = (1.0f - g.p5) * metal, dFdM = (1.0f - g.p5) * (color[c] - 0.04f);
        if (g.lobe == 1) {
            dC[c] = -dFdC / (1.0f - F) + 1.0f / color[c];
            dM[c] = -dFdM / (1.0f - F) - 1.0f / (1.0f - metal);
        } else if (g.lobe == 2) {
            dC[c] = dFdC / F; dM[c] = dFdM / F; dA[c] = dlnDG;
        } else {
            dC[c] = -dFdC / (1.0f - F); dM[c] = -dFdM / (1.0f - F); dA[c] = dlnDG;
        }
    }
}

// ---------------------------------------------------------- complex IOR
// fresnelReflection (OpenMaterial.cpp:20-38): unpolarised reflectance of
// vacuum -> N = n + ik, principal complex square root.
struct cpx { float re, im; };
DEV cpx cmul(cpx a, cpx b) { cpx r; r.re = a.re * b.re - a.im * b.im; r.im = a.re * b.im + a.im * b.re; return r; }
DEV cpx csqrt_(cpx z) {   // principal branch
    float m = sqrtf(z.re * z.re + z.im * z.im);
    float re = sqrtf(0.5f * (m + fabsf(z.re)));
    cpx r;
    if (re == 0.0f) { r.re = 0.0f; r.im = 0.0f; return r; }
    if (z.re >= 0.0f) { r.re = re; r.im = z.im / (2.0f * re); }
    else { r.re = fabsf(z.im) / (2.0f * re); r.im = z.im < 0.0f ? -re : re; }
    return r;
}
DEV cpx cdiv(cpx a, cpx b) {   // Smith's algorithm
    cpx r;
    if (fabsf(b.re) >= fabsf(b.im)) {
        float q = b.im / b.re, den = b.re + b.im * q;
        r.re = (a.re + a.im * q) / den; r.im = (a.im - a.re * q) / den;
    } else {
        float q = b.re / b.im, den = b.im + b.re * q;
        r.re = (a.re * q + a.im) / den; r.im = (a.im * q - a.re) / den;
    }
    return r;
}
DEV void fresnel(float n, float k, float cosT, float& Rp, float& Rs) {
    cpx N; N.re = n; N.im = k;
    cpx N2 = cmul(N, N);
    cosT = fabsf(cosT);
    const float c2 = cosT * cosT, s2 = 1.0f - c2;
    cpx z; z.re = N2.re - s2; z.im = N2.im;
    cpx sq = csqrt_(z);
    cpx N2c; N2c.re = N2.re * cosT; N2c.im = N2.im * cosT;
    cpx np, dp, ns, ds;
    np.re = N2c.re - sq.re; np.im = N2c.im - sq.im;
    dp.re = N2c.re + sq.re; dp.im = N2c.im + sq.im;
    ns.re = cosT - sq.re; ns.im = -sq.im;
    ds.re = cosT + sq.re; ds.im = sq.im;
    cpx rp = cdiv(np, dp), rs = cdiv(ns, ds);
    Rp = rp.re * rp.re + rp.im * rp.im;
    Rs = rs.re * rs.re + rs.im * rs.im;
}
// d ln(Rp + Rs) / d(n, k) at fixed cos: with N2 = (n+ik)^2, u = sqrt(N2 - s^2),
// d rp/dN2 = c (2u^2 - N2) / (u (N2 c + u)^2), d rs/dN2 = -c / (u (c + u)^2),
// dN2/dn = 2N, dN2/dk = 2iN, d|r|^2 = 2 Re(conj(r) dr).
DEV void fresnelDlog(float n, float k, float cosT, float& dn, float& dk) {
    cpx N; N.re = n; N.im = k;
    cpx N2 = cmul(N, N);
    cosT = fabsf(cosT);
    const float c2 = cosT * cosT, s2 = 1.0f - c2;
    cpx z; z.re = N2.re - s2; z.im = N2.im;
    cpx u = csqrt_(z);
    cpx a; a.re = N2.re * cosT; a.im = N2.im * cosT;
    cpx num, den;
    num.re = a.re - u.re; num.im = a.im - u.im; den.re = a.re + u.re; den.im = a.im + u.im;
    cpx rp = cdiv(num, den);
    num.re = cosT - u.re; num.im = -u.im; den.re = cosT + u.re; den.im = u.im;
    cpx rs = cdiv(num, den);
    cpx apu; apu.re = a.re + u.re; apu.im = a.im + u.im;
    cpx cpu; cpu.re = cosT + u.re; cpu.im = u.im;
    cpx uu = cmul(u, u);
    cpx t1; t1.re = cosT * (2.0f * uu.re - N2.re); t1.im = cosT * (2.0f * uu.im - N2.im);
    cpx drp = cdiv(t1, cmul(u, cmul(apu, apu)));
    cpx mc; mc.re = -cosT; mc.im = 0.0f;
    cpx drs = cdiv(mc, cmul(u, cmul(cpu, cpu)));
    cpx dN2n; dN2n.re = 2.0f * N.re; dN2n.im = 2.0f * N.im;
    cpx dN2k; dN2k.re = -2.0f * N.im; dN2k.im = 2.0f * N.re;
    cpx crp; crp.re = rp.re; crp.im = -rp.im;
    cpx crs; crs.re = rs.re; crs.im = -rs.im;
    const float F = (rp.re * rp.re + rp.im * rp.im) + (rs.re * rs.re + rs.im * rs.im);
    if (!(F != 0.0f)) { dn = dk = 0.0f; return; }
    float dRn = 2.0f * (cmul(crp, cmul(drp, dN2n)).re + cmul(crs, cmul(drs, dN2n)).re);
    float dRk = 2.0f * (cmul(crp, cmul(drp, dN2k)).re + cmul(crs, cmul(drs, dN2k)).re);
    dn = dRn / F;
    dk = dRk / F;
}

// ------------------------------------------------------------ textures
// BitmapTexture::getPixCoords + texture (BitmapTexture.cpp:181-234); fp32
// textures (an extension) use the same addressing without the 1/255 scale.
// The four texels and bilinear weights of (u, v) (BitmapTexture::getPixCoords,
// texel-centre wrap addressing); false for a missing image.
DEV bool texAddress(const OmvTexture& t, float u, float v, int id[4], float w[4]) {
    if (t.w <= 0 || t.h <= 0) return false;
    if (!isfinite(u)) u = 0.0f;
    if (!isfinite(v)) v = 0.0f;
    const float W = (float)t.w, H = (float)t.h;
    float upage = floorf(u - 0.5f / W), vpage = floorf(v - 0.5f / H);
    u -= upage; v -= vpage;
    u *= W; v *= H;
    int iu = (int)u, iv = (int)v;
    float kui = u - (float)iu, kvi = v - (float)iv;
    float ku = 1.0f - kui, kv = 1.0f - kvi;
    int l = iu % t.w, r = (iu + 1) % t.w, b = iv % t.h, tt = (iv + 1) % t.h;
    const int n = t.w * t.h - 1;
    id[0] = min(max(b * t.w + l, 0), n); id[1] = min(max(b * t.w + r, 0), n);
    id[2] = min(max(tt * t.w + l, 0), n); id[3] = min(max(tt * t.w + r, 0), n);
    w[0] = ku * kv; w[1] = kui * kv; w[2] = ku * kvi; w[3] = kui * kvi;
    return true;
}
DEV f4 texLookup(const OmvSceneView& S, int id, float u, float v) {
    const OmvTexture t = S.texs[id];
    f4 out;
    int i[4];
    float w[4];
    if (!texAddress(t, u, v, i, w)) { out.x = out.y = out.z = 0.5f; out.w = 1.0f; return out; }
    if (t.fmt == OMV_TEX_RGBA8) {
        const uint32_t* px = (const uint32_t*)(S.texels + t.offset);
        uint32_t p0 = px[i[0]], p1 = px[i[1]], p2 = px[i[2]], p3 = px[i[3]];
        const float s = 1.0f / 255.0f;
        float o[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            int sh = 8 * c;
            o[c] = (float)((p0 >> sh) & 0xFF) * w[0] * s + (float)((p1 >> sh) & 0xFF) * w[1] * s +
                   (float)((p2 >> sh) & 0xFF) * w[2] * s + (float)((p3 >> sh) & 0xFF) * w[3] * s;
        }
        out.x = o[0]; out.y = o[1]; out.z = o[2]; out.w = o[3];
    } else {
        const float4* px = (const float4*)(S.texels + t.offset);
        float4 p0 = px[i[0]], p1 = px[i[1]], p2 = px[i[2]], p3 = px[i[3]];
        out.x = p0.x * w[0] + p1.x * w[1] + p2.x * w[2] + p3.x * w[3];
        out.y = p0.y * w[0] + p1.y * w[1] + p2.y * w[2] + p3.y * w[3];
        out.z = p0.z * w[0] + p1.z * w[1] + p2.z * w[2] + p3.z * w[3];
        out.w = p0.w * w[0] + p1.w * w[1] + p2.w * w[2] + p3.w * w[3];
    }
    return out;
}

// -------------------------------------------------------- surface point
struct SP {
    int inst, mesh, face;
    int mat;              // the mesh's material slot (read with the mesh record once per hit)
    f3 pos, n, t, b, bary;
    float tw;
    bool texNorm;
    float density, gradient;
};

struct Hit { float t, u, v; int inst, prim; bool tie; };

// BDPTRenderer::computeSurfacePoint + RenderMesh::computeSurfacePoint
// (BDPTRenderer.cpp:57-73, RenderMesh.cpp:159-234)
DEV void computeSurfacePoint(const OmvSceneView& S, const Hit& h, SP& sp, bool fcDensity) {
    const OmvInstance& in = S.insts[h.inst];
    const OmvMesh m = S.meshes[in.mesh_id];
    sp.inst = h.inst; sp.mesh = in.mesh_id; sp.face = h.prim; sp.mat = m.mat_id;
    sp.bary = mk3(1.0f - h.u - h.v, h.u, h.v);
    sp.texNorm = false;
    const uint32_t* f = S.faces + m.face_offset + 3u * (uint32_t)h.prim;
    const float* v0 = S.vtx + m.vtx_offset + f[0] * m.stride;
    const float* v1 = S.vtx + m.vtx_offset + f[1] * m.stride;
    const float* v2 = S.vtx + m.vtx_offset + f[2] * m.stride;
    const float b0 = sp.bary.x, b1 = sp.bary.y, b2 = sp.bary.z;
    f3 pp = mk3(b0 * v0[0] + b1 * v1[0] + b2 * v2[0], b0 * v0[1] + b1 * v1[1] + b2 * v2[1], b0 * v0[2] + b1 * v1[2] + b2 * v2[2]);
    f3 pn = mk3(b0 * v0[3] + b1 * v1[3] + b2 * v2[3], b0 * v0[4] + b1 * v1[4] + b2 * v2[4], b0 * v0[5] + b1 * v1[5] + b2 * v2[5]);
    f3 pt = mk3(b0 * v0[6] + b1 * v1[6] + b2 * v2[6], b0 * v0[7] + b1 * v1[7] + b2 * v2[7], b0 * v0[8] + b1 * v1[8] + b2 * v2[8]);
    float pw = b0 * v0[9] + b1 * v1[9] + b2 * v2[9];
    pw = pw >= 0.0f ? 1.0f : -1.0f;
    pn = normalizeIfNeeded3(pn);
    float tn = dot3(pn, pt);
    pt = mk3(pt.x - tn * pn.x, pt.y - tn * pn.y, pt.z - tn * pn.z);
    pt = normalizeIfNeeded3(pt);
    f3 pbn = cross3(pn, pt);
    pbn = mk3(pbn.x * pw, pbn.y * pw, pbn.z * pw);
    sp.pos = mul34(in.xfm, pp);
    sp.n = normalizeIfNeeded3(mul33(in.nrm, pn));
    sp.t = normalizeIfNeeded3(mul33(in.nrm, pt));
    sp.b = normalizeIfNeeded3(mul33(in.nrm, pbn));
    sp.tw = pw;
    sp.density = 0.0f; sp.gradient = 0.0f;
    if (fcDensity) {
        f3 V0 = mul34(in.xfm, ld3(v0)), V1 = mul34(in.xfm, ld3(v1)), V2 = mul34(in.xfm, ld3(v2));
        f3 tfn = cross3(sub3(V1, V0), sub3(V2, V0));
        float dens = 1.0f / (0.0f + length3(tfn) / 2.0f);
        tfn = normalize3(tfn);
        float ag = 0.0f;
        const float* vv[3] = {v0, v1, v2};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            f3 tvn = normalizeIfNeeded3(mul33(in.nrm, ld3(vv[k] + 3)));
            ag = ag + sqrtf((1.0f - clampr(dot3(tvn, tfn), -1.0f, 1.0f)) / 2.0f);
        }
        sp.gradient = powf(clampr(ag / 3.0f, 1e-6f, 1.0f), 0.286f);
        const float mnD = S.min_density, mxD = S.max_density;
        sp.density = clamp01(1.0f - logf(mxD / clampr(dens, mnD, mxD)) / logf(mxD / mnD));
    }
}

// hit position only (shadow rays): same arithmetic as computeSurfacePoint
DEV f3 hitPosition(const OmvSceneView& S, const Hit& h) {
    const OmvInstance& in = S.insts[h.inst];
    const OmvMesh m = S.meshes[in.mesh_id];
    const uint32_t* f = S.faces + m.face_offset + 3u * (uint32_t)h.prim;
    const float* v0 = S.vtx + m.vtx_offset + f[0] * m.stride;
    const float* v1 = S.vtx + m.vtx_offset + f[1] * m.stride;
    const float* v2 = S.vtx + m.vtx_offset + f[2] * m.stride;
    const float b0 = 1.0f - h.u - h.v, b1 = h.u, b2 = h.v;
    f3 pp = mk3(b0 * v0[0] + b1 * v1[0] + b2 * v2[0], b0 * v0[1] + b1 * v1[1] + b2 * v2[1], b0 * v0[2] + b1 * v1[2] + b2 * v2[2]);
    return mul34(in.xfm, pp);
}

// SurfacePoint::applyTextureNormal (SurfacePoint.cpp:19-34)
DEV void applyTextureNormal(SP& sp, const f3& tn) {
    if (sp.texNorm) return;
    f3 nn = mk3(tn.x * sp.t.x + tn.y * sp.b.x + tn.z * sp.n.x, tn.x * sp.t.y + tn.y * sp.b.y + tn.z * sp.n.y,
                tn.x * sp.t.z + tn.y * sp.b.z + tn.z * sp.n.z);
    sp.n = normalizeIfNeeded3(nn);
    sp.t = normalizeIfNeeded3(cross3(sp.b, sp.n));
    f3 b = cross3(sp.n, sp.t);
    sp.b = mk3(sp.tw * b.x, sp.tw * b.y, sp.tw * b.z);
    sp.texNorm = true;
}

// RenderMesh::getTexCoord (RenderMesh.cpp:145-157)
DEV bool getTexCoord(const OmvSceneView& S, int meshId, int face, const f3& bary, int ch, float& u, float& v) {
    u = 0.0f; v = 0.0f;
    const OmvMesh& m = S.meshes[meshId];
    if (face < 0 || face >= (int)m.num_faces || ch < 0 || ch >= OMV_MAX_UV_CHANNELS) return false;
    const int off = m.tex_off[ch];
    if (off < 0) return false;
    const uint32_t* f = S.faces + m.face_offset + 3u * (uint32_t)face;
    const float* a = S.vtx + m.vtx_offset + f[0] * m.stride + off;
    const float* b = S.vtx + m.vtx_offset + f[1] * m.stride + off;
    const float* c = S.vtx + m.vtx_offset + f[2] * m.stride + off;
    u = a[0] * bary.x + b[0] * bary.y + c[0] * bary.z;
    v = a[1] * bary.x + b[1] * bary.y + c[1] * bary.z;
    return true;
}

// ------------------------------------------------------------ materials
DEV bool pbrTex(const OmvSceneView& S, const SP& sp, const OmvMaterial& M, int slot, f4& out) {
    int id = M.tex_map[slot];
    if (!(0 <= id && id < S.num_texs)) return false;
    float u, v;
    if (!getTexCoord(S, sp.mesh, sp.face, sp.bary, M.tex_ch[slot], u, v)) return false;
    f3 uvT = mul33(M.tex_xf[slot], mk3(u, v, 1.0f));
    out = texLookup(S, id, uvT.x, uvT.y);
    return true;
}
// PhysicallyBasedMaterial::Helper (PhysicallyBasedMaterial.cpp:32-100)
DEV void pbrInputs(const OmvSceneView& S, const SP& sp, const OmvMaterial& M, float col[4], float omr[3]) {
    f4 t;
    if (pbrTex(S, sp, M, OMV_PBR_BASE_COLOR, t)) { col[0] = t.x; col[1] = t.y; col[2] = t.z; col[3] = t.w; }
    else { col[0] = col[1] = col[2] = col[3] = 1.0f; }
#pragma unroll
    for (int i = 0; i < 4; ++i) col[i] = col[i] * M.bcf[i];
    omr[0] = 1.0f; omr[1] = M.metal; omr[2] = M.rough;
    if (pbrTex(S, sp, M, OMV_PBR_METALLIC_ROUGHNESS, t)) { omr[0] *= t.x; omr[1] *= t.z; omr[2] *= t.y; }
    if (M.occ_separate && pbrTex(S, sp, M, OMV_PBR_OCCLUSION, t)) omr[0] = t.x;
}
// PBR texture inputs at a surface point (pbrInputs), read once per vertex
// where several BRDF evaluations there need them (the connections to the
// light vertices, then the bounce): the same texel reads, done once
struct MatIn { float col[4], omr[3]; };
DEV void matInputs(const OmvSceneView& S, const SP& sp, MatIn& mi) {
    const OmvMaterial& M = S.mats[sp.mat];
    if (M.type == OMV_MAT_PBR) pbrInputs(S, sp, M, mi.col, mi.omr);
}
// mi: the inputs from matInputs at this point, or null to read them here
DEV void pbrSetup(const OmvSceneView& S, const SP& sp, const OmvMaterial& M, Bsdf& b, float omr[3], const MatIn* mi = nullptr) {
    float col[4];
    if (mi) {
#pragma unroll
        for (int k = 0; k < 4; ++k) col[k] = mi->col[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) omr[k] = mi->omr[k];
    } else {
        pbrInputs(S, sp, M, col, omr);
    }
    b.c0 = col[0]; b.c1 = col[1]; b.c2 = col[2]; b.c3 = col[3];
    if (M.alpha_mode != OMV_ALPHA_BLEND) b.c3 = 1.0f;
    b.transmissive = M.double_sided || b.c3 < 1.0f;
}
DEV void pbrFinish(const OmvMaterial& M, Bsdf& b, const float omr[3]) {
    b.c0 = lerpr(b.c0, omr[0] * b.c0, M.occ_strength);
    b.c1 = lerpr(b.c1, omr[0] * b.c1, M.occ_strength);
    b.c2 = lerpr(b.c2, omr[0] * b.c2, M.occ_strength);
    b.metal = omr[1];
    b.rough = clampr(omr[2], 0.00001f, 0.99999f);
    b.alpha = powf(b.rough, 2.0f);
}
// PhysicallyBasedMaterial::modifyFrame (:147-167); the other materials
// apply the identity texture normal (OpenMaterial.cpp:66-70, DiffuseColorMaterial.cpp:35-39)
DEV void modifyFrame(const OmvSceneView& S, SP& sp) {
    const OmvMaterial& M = S.mats[sp.mat];
    f3 tn = mk3(0.0f, 0.0f, 1.0f);
    if (M.type == OMV_MAT_PBR) {
        f4 t;
        if (pbrTex(S, sp, M, OMV_PBR_NORMAL, t)) {
            if (t.z <= 0.5f) t.z = 0.5f + FLT_EPSILON;
            float x = 2.0f * t.x + -1.0f, y = 2.0f * t.y + -1.0f, z = 2.0f * t.z + -1.0f;
            x = M.normal_scale * x; y = M.normal_scale * y;
            tn = normalize3(mk3(x, y, z));
        }
    }
    applyTextureNormal(sp, tn);
}
// emission part of PhysicallyBasedMaterial::getRayAndBrdf (:282-287, :102-123)
DEV f3 pbrEmission(const OmvSceneView& S, const SP& sp, const OmvMaterial& M) {
    float u, v;
    getTexCoord(S, sp.mesh, sp.face, sp.bary, M.tex_ch[OMV_PBR_EMISSIVITY], u, v);
    f3 uvT = mul33(M.tex_xf[OMV_PBR_EMISSIVITY], mk3(u, v, 1.0f));
    f4 t = texLookup(S, M.tex_map[OMV_PBR_EMISSIVITY], uvT.x, uvT.y);
    return mk3(t.x * M.ef[0], t.y * M.ef[1], t.z * M.ef[2]);
}
// PhysicallyBasedMaterial::isMasked (:301-316)
DEV bool isMaskedHit(const OmvSceneView& S, int meshId, int face, const f3& bary) {
    const OmvMaterial& M = S.mats[S.meshes[meshId].mat_id];
    if (M.type != OMV_MAT_PBR || M.alpha_mode != OMV_ALPHA_MASK) return false;
    float alpha = M.bcf[3];
    if (alpha < M.alpha_cutoff) return true;
    int id = M.tex_map[OMV_PBR_BASE_COLOR];
    if (0 <= id && id < S.num_texs) {
        float u, v, a;
        if (!getTexCoord(S, meshId, face, bary, M.tex_ch[OMV_PBR_BASE_COLOR], u, v)) a = 1.0f;
        else a = texLookup(S, id, u, v).w;
        alpha *= a;
    }
    return alpha < M.alpha_cutoff;
}
// geometric normal through the normal matrix (OpenMaterial.cpp:76-86)
DEV f3 faceNormal(const OmvSceneView& S, const SP& sp) {
    const OmvMesh m = S.meshes[sp.mesh];
    const uint32_t* f = S.faces + m.face_offset + 3u * (uint32_t)sp.face;
    f3 P0 = ld3(S.vtx + m.vtx_offset + f[0] * m.stride), P1 = ld3(S.vtx + m.vtx_offset + f[1] * m.stride),
       P2 = ld3(S.vtx + m.vtx_offset + f[2] * m.stride);
    f3 fn = cross3(sub3(P1, P0), sub3(P2, P0));
    return normalize3(mul33(S.insts[sp.inst].nrm, fn));
}
// OpenMaterial::defineNextDirection (OpenMaterial.cpp:72-93)
DEV f3 omNextDirection(const OmvSceneView& S, const f3& inV, const SP& sp) {
    f3 out = reflect3(inV, sp.n);
    f3 tfn = faceNormal(S, sp);
    float corr = dot3(out, tfn);
    if (corr < 0.1f) {
        float k = 0.1f - corr;
        out = normalize3(mk3(tfn.x * k + out.x, tfn.y * k + out.y, tfn.z * k + out.z));
    }
    return out;
}
DEV f3 omFresnelBrdf(const OmvMaterial& M, float fCos) {
    float o[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float rp, rs;
        fresnel(M.n[c], M.k[c], fCos, rp, rs);
        o[c] = lerpr(rp, rs, 0.5f) / fabsf(fCos);
    }
    return mk3(o[0], o[1], o[2]);
}
// Estimator extensions (omv_renderer_params.extensions)
struct Ext { bool omRough; float omAlpha; bool nee, splat; };

// OMV_EXT_OM_ROUGH: GGX rough conductor with the complex-IOR Fresnel as F.
// Connection: f = F(|I.H|) D G2 / (4 V.N L.N), zero unless both cosines > 0.
DEV f3 omRoughEval(const OmvMaterial& M, float alpha, const f3& I, const f3& O, const f3& N) {
    float VdotN = -dot3(I, N), LdotN = dot3(O, N);
    if (!(VdotN > 0.0f && LdotN > 0.0f)) return mk3(0.0f, 0.0f, 0.0f);
    f3 H = normalize3(sub3(O, I));
    float VdotH = fabsf(dot3(I, H)), NdotH = fabsf(dot3(N, H));
    float D = ggxPdf(clamp01(NdotH), alpha);
    float G = smithG2(smithG1(VdotN, alpha), smithG1(LdotN, alpha));
    float k = D * G / (4.0f * VdotN * LdotN);
    float o[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float rp, rs;
        fresnel(M.n[c], M.k[c], VdotH, rp, rs);
        o[c] = lerpr(rp, rs, 0.5f) * k;
    }
    return mk3(o[0], o[1], o[2]);
}
// Sampling: H from GGXHemisphere (2 uniforms), O = reflect(I, H); brdf times
// the path cosine is F G |I.H| / (|I.N| |N.H|).
DEV f3 omRoughSample(const OmvMaterial& M, float alpha, const f3& I, const SP& sp, bool inv, f3& O, Rng& r) {
    const f3 N = sp.n;
    O = reflect3(I, N);
    float VdotN = -dot3(I, N);
    if (!(VdotN > 0.0f)) return mk3(0.0f, 0.0f, 0.0f);
    f3 B = cross3(N, sp.t);
    f3 rd = ggxHemisphere(r, alpha * alpha);
    f3 H = normalize3(toBasis(sp.t, B, N, rd));
    f3 o = reflect3(I, H);
    if (!(dot3(N, o) > 0.0f)) return mk3(0.0f, 0.0f, 0.0f);
    O = normalize3(o);
    float LdotN = dot3(O, N), VdotH = fabsf(dot3(I, H)), NdotH = fabsf(dot3(N, H));
    float G = smithG2(smithG1(VdotN, alpha), smithG1(LdotN, alpha));
    float w = G * VdotH / (VdotN * NdotH);
    float ct = inv ? LdotN : VdotN;
    float out[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float rp, rs;
        fresnel(M.n[c], M.k[c], VdotH, rp, rs);
        out[c] = lerpr(rp, rs, 0.5f) * w / ct;
    }
    return mk3(out[0], out[1], out[2]);
}

// RenderMaterial::getBrdf of the three materials for a connection
DEV f3 matGetBrdf(const OmvSceneView& S, const Ext& X, const SP& sp, const f3& inV, const f3& outV, const MatIn* mi = nullptr) {
    const OmvMaterial& M = S.mats[sp.mat];
    if (M.type == OMV_MAT_OM) {   // OpenMaterial::getBrdf (OpenMaterial.cpp:95-163)
        if (!M.has_base) return mk3(0.0f, 0.0f, 0.0f);
        if (X.omRough) return omRoughEval(M, X.omAlpha, inV, outV, sp.n);
        float fCos = -dot3(inV, sp.n);
        if (fabsf(dot3(outV, sp.n) - fCos) > 1e-6f) return mk3(0.0f, 0.0f, 0.0f);
        return omFresnelBrdf(M, fCos);
    }
    if (M.type == OMV_MAT_PBR) {  // PhysicallyBasedMaterial::getBrdf (:178-245)
        Bsdf b; bsdfInit(b);
        float omr[3];
        b.I = inV; b.O = outV; b.N = sp.n;
        pbrSetup(S, sp, M, b, omr, mi);
        if (dot3(inV, sp.n) > 0.0f && b.transmissive) b.N = neg3(b.N);
        pbrFinish(M, b, omr);
        return evaluateDirect(b);
    }
    // DiffuseColorMaterial::getBrdf (DiffuseColorMaterial.cpp:89-104)
    const float pdf = 1.0f / kPi;
    const float brdfPdf = 1.0f / (pdf * kPi);
    return mk3(brdfPdf * M.diffuse[0], brdfPdf * M.diffuse[1], brdfPdf * M.diffuse[2]);
}
// RenderMaterial::getRayAndBrdf of the three materials (computePath, BDPTRenderer.cpp:158-159)
// What the adjoint needs from a bounce (kind 1 PBR: the sampled Bsdf state and
// the occlusion input; kind 2 OM mirror: the Fresnel cosine).
struct BounceInfo {
    int kind;
    Bsdf b;
    float occ, roughRaw, fCos;
};
DEV void matGetRayAndBrdf(const OmvSceneView& S, const Ext& X, const SP& sp, const f3& inV, f3& outV, f3& brdf, f3& erad, bool inv,
                          Rng& r, BounceInfo* info = nullptr, const MatIn* mi = nullptr) {
    const OmvMaterial& M = S.mats[sp.mat];
    erad = mk3(0.0f, 0.0f, 0.0f);
    if (info) info->kind = 0;
    if (M.type == OMV_MAT_OM) {
        if (X.omRough && M.has_base) { brdf = omRoughSample(M, X.omAlpha, inV, sp, inv, outV, r); return; }
        outV = omNextDirection(S, inV, sp);
        brdf = M.has_base ? omFresnelBrdf(M, dot3(outV, sp.n)) : mk3(0.0f, 0.0f, 0.0f);
        if (info && M.has_base) { info->kind = 2; info->fCos = dot3(outV, sp.n); }
        return;
    }
    if (M.type == OMV_MAT_PBR) {
        Bsdf b; bsdfInit(b);
        float omr[3];
        b.I = inV; b.N = sp.n; b.T = sp.t;
        pbrSetup(S, sp, M, b, omr, mi);
        pbrFinish(M, b, omr);
        b.eta = 1.0f / (M.double_sided ? 1.0f : M.ior);
        brdf = sampleIndirect(b, r);
        outV = b.O;
        if (M.emissive) erad = pbrEmission(S, sp, M);
        if (info) { info->kind = 1; info->b = b; info->occ = omr[0]; info->roughRaw = omr[2]; }
        return;
    }
    f3 v = uniformHemisphere(r);
    outV = toBasis(sp.t, sp.b, sp.n, v);
    brdf = inv ? matGetBrdf(S, X, sp, neg3(outV), neg3(inV)) : matGetBrdf(S, X, sp, inV, outV);
}

// ------------------------------------------------------------ traversal
struct Counters { uint32_t nodes, tris, overflow; };

// Per-lane traversal stack: entries 0 .. OMV_STACK_LDS-1 in the kernel's
// dynamic LDS (omv_lds_stack, [entry][lane]), the rest in a global spill
// area.  The two parts are addressed through their own address spaces, so
// the LDS entries compile to ds_read / ds_write (LDS counter only) rather
// than to flat accesses, which would also wait on the vector memory counter
// behind every node and triangle fetch in flight.
typedef __attribute__((address_space(1))) uint32_t GlobalU32;
// Every kernel that traverses runs OMV_BLOCK-lane workgroups with
// OMV_STACK_LDS * OMV_BLOCK words of dynamic LDS, so the LDS stride is a
// constant (a shift, not a multiply, per push and pop).
struct Stack {
    uint32_t col;        // this lane's LDS column: omv_lds_stack[col + entry * OMV_BLOCK]
    GlobalU32* glb;      // this lane's spill area, stride = glbStride
    uint32_t glbStride;
    int sp;
    // lane `tid` of `threads` (spill stacks interleaved by lane in gstack)
    DEV void init(uint32_t* gstack, uint32_t tid, uint32_t threads) {
        col = threadIdx.x;
        glb = (GlobalU32*)(gstack + tid); glbStride = threads;
        sp = 0;
    }
    DEV void push(uint32_t e, Counters& c) {
        if (sp < OMV_STACK_LDS) omv_lds_stack[col + (uint32_t)sp * OMV_BLOCK] = e;
        else if (sp < OMV_STACK_TOTAL) glb[(uint32_t)(sp - OMV_STACK_LDS) * glbStride] = e;
        else { ++c.overflow; return; }
        ++sp;
    }
    // push e if `valid`, without a branch on it: the slot above the top is
    // written either way (free space) and the top moves only for a valid entry
    DEV void pushIf(uint32_t e, bool valid, Counters& c) {
        if (sp < OMV_STACK_LDS) omv_lds_stack[col + (uint32_t)sp * OMV_BLOCK] = e;
        else if (sp < OMV_STACK_TOTAL) glb[(uint32_t)(sp - OMV_STACK_LDS) * glbStride] = e;
        const bool room = sp < OMV_STACK_TOTAL;
        c.overflow += (valid && !room) ? 1u : 0u;
        sp += (valid && room) ? 1 : 0;
    }
    // pushIf of three entries; when no active lane of the wave could reach
    // the spill area (a wave-uniform test) the entries go to LDS without the
    // per-entry LDS / spill / overflow branches
    DEV void push3If(const uint32_t e[3], const bool v[3], Counters& c) {
        if (__ballot(sp > OMV_STACK_LDS - 3) == 0ull) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                omv_lds_stack[col + (uint32_t)sp * OMV_BLOCK] = e[k];
                sp += v[k] ? 1 : 0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) pushIf(e[k], v[k], c);
        }
    }
    DEV uint32_t pop() {
        --sp;
        if (__ballot(sp >= OMV_STACK_LDS) == 0ull) return omv_lds_stack[col + (uint32_t)sp * OMV_BLOCK];   // wave-uniform: LDS only
        return sp < OMV_STACK_LDS ? omv_lds_stack[col + (uint32_t)sp * OMV_BLOCK] : glb[(uint32_t)(sp - OMV_STACK_LDS) * glbStride];
    }
};

enum : uint32_t {
    E_KIND_SHIFT = OMV_ENT_SHIFT,
    E_BLAS_NODE = OMV_ENT_BLAS_NODE, E_BLAS_LEAF = OMV_ENT_BLAS_LEAF, E_TLAS_NODE = OMV_ENT_TLAS_NODE, E_INST = OMV_ENT_INST,
    E_END = OMV_ENT_END, E_NONE = OMV_ENT_NONE,
};

// Component selection for the ray's axis permutation by bit masks: k in
// {0, 1, 2} gives m1 = all ones iff k == 1 and m2 = all ones iff k == 2,
// and pick() takes y over x by m1, then z by m2 (two v_bfi_b32).  The masks
// pass through an empty asm so the optimiser does not turn the selects back
// into a compare chain (lowered as branches) or a dynamic vector index
// (lowered through scratch memory).
struct PermMask { uint32_t m1, m2; };
DEV PermMask permMask(int k) {
    PermMask r;
    r.m1 = 0u - (uint32_t)(k & 1);
    r.m2 = 0u - (uint32_t)((k >> 1) & 1);
    asm("" : "+v"(r.m1), "+v"(r.m2));
    return r;
}
DEV float pick(const f3& v, const PermMask& m) {
    const uint32_t a = (__float_as_uint(v.x) & ~m.m1) | (__float_as_uint(v.y) & m.m1);
    return __uint_as_float((a & ~m.m2) | (__float_as_uint(v.z) & m.m2));
}

struct Pre { int kx, ky, kz; float Sx, Sy, Sz; f3 o, d, idir; };

DEV bool rayPrep(const f3& o, const f3& d, Pre& p) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    if (!(ax > 0.0f || ay > 0.0f || az > 0.0f)) return false;   // zero / NaN direction: no hit
    int kz = 0; float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    float dkz = pick(d, permMask(kz));
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    p.kx = kx; p.ky = ky; p.kz = kz;
    p.Sx = pick(d, permMask(kx)) / dkz;
    p.Sy = pick(d, permMask(ky)) / dkz;
    p.Sz = 1.0f / dkz;
    p.o = o; p.d = d;
    p.idir = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    return true;
}

// Woop, Benthin & Wald 2013; t in (tnear, tfar]
DEV bool triIntersect(const Pre& p, const float4 r0, const float4 r1, const float4 r2, float tnear, float tfar, float& t, float& u, float& v) {
    f3 A = mk3(r0.x - p.o.x, r0.y - p.o.y, r0.z - p.o.z);
    f3 B = mk3(r1.x - p.o.x, r1.y - p.o.y, r1.z - p.o.z);
    f3 C = mk3(r2.x - p.o.x, r2.y - p.o.y, r2.z - p.o.z);
    const PermMask mx = permMask(p.kx), my = permMask(p.ky), mz = permMask(p.kz);
    const float Akz = pick(A, mz), Bkz = pick(B, mz), Ckz = pick(C, mz);
    float Ax = pick(A, mx) - p.Sx * Akz, Ay = pick(A, my) - p.Sy * Akz;
    float Bx = pick(B, mx) - p.Sx * Bkz, By = pick(B, my) - p.Sy * Bkz;
    float Cx = pick(C, mx) - p.Sx * Ckz, Cy = pick(C, my) - p.Sy * Ckz;
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    const bool zero = U == 0.0f || V == 0.0f || W == 0.0f;
    // the FP64 recomputation is rare: a wave-uniform branch around it keeps
    // the compiler from computing it for every triangle and selecting
    if (__ballot(zero) != 0ull && zero) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = p.Sz * Akz, Bz = p.Sz * Bkz, Cz = p.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float tt = T * rcp;
    if (!(tt > tnear && tt <= tfar)) return false;
    t = tt; u = V * rcp; v = W * rcp;
    return true;
}

DEV void offer(Hit& h, bool& found, float t, float u, float v, int inst, int prim) {
    if (!found || t < h.t) { h.t = t; h.u = u; h.v = v; h.inst = inst; h.prim = prim; h.tie = false; found = true; return; }
    if (t == h.t) {
        if (inst != h.inst || prim != h.prim) h.tie = true;
        if (inst < h.inst || (inst == h.inst && prim < h.prim)) { h.u = u; h.v = v; h.inst = inst; h.prim = prim; }
    }
}

// 4-wide slab test, conservative (Ize 2013: tmax scaled by 1 + 2*gamma(3)).
// An empty slot (planes at +inf, entry E_NONE) tests as a miss or, when
// tfar * kUp overflows to +inf, as a hit at +inf; either way its E_NONE
// entry is never pushed (pushSorted), so no entry check is needed here.
DEV void boxTest4(const OmvNode4& nd, const f3& o, const f3& idir, float tnear, float tfar, float tent[4], bool hit[4]) {
    const float kUp = 1.0000004f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float t0x = (nd.lox[c] - o.x) * idir.x, t1x = (nd.hix[c] - o.x) * idir.x;
        float t0y = (nd.loy[c] - o.y) * idir.y, t1y = (nd.hiy[c] - o.y) * idir.y;
        float t0z = (nd.loz[c] - o.z) * idir.z, t1z = (nd.hiz[c] - o.z) * idir.z;
        float tmin = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tnear));
        float tmax = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tfar));
        tent[c] = tmin;
        hit[c] = tmin <= tmax * kUp;
    }
}

// Byte offsets of a ray's near plane rows within a node (OmvNode4's float4
// rows lox hix loy hiy loz hiz child entry), packed x | y << 8 | z << 16:
// the near x plane of every child is hix for a ray going -x (sign bit of
// 1/d.x set, -0 included) and lox otherwise; the far row is the other one
// of the pair (offset ^ 16).
DEV uint32_t nearRows(const f3& idir) {
    const uint32_t sx = __float_as_uint(idir.x) >> 31, sy = __float_as_uint(idir.y) >> 31, sz = __float_as_uint(idir.z) >> 31;
    return (sx << 4) | ((32u | (sy << 4)) << 8) | ((64u | (sz << 4)) << 16);
}
DEV float4 ldRow(const char* base, uint32_t off) { return *reinterpret_cast<const float4*>(base + off); }
// boxTest4 of node idx with the planes fetched by the ray's direction signs:
// the near distance on an axis is (near - o) * idir and the far one
// (far - o) * idir, the same values as the min / max of the two (the
// product is monotonic in the plane), so no per-axis min / max is needed.
// Where the min / max form would drop a NaN distance (o on a plane with a
// zero direction component) the max / min below drop it too, so no box
// that form accepts is rejected.  Also returns the children's entries.
// Node byte offsets must fit in 32 bits (OMV_MAX_NODES).
DEV void boxTest4Rows(const OmvNode4* nodes, uint32_t idx, uint32_t rows, const f3& o, const f3& idir, float tnear, float tfar, float tent[4],
                      bool hit[4], uint32_t ent[4]) {
    const char* base = reinterpret_cast<const char*>(nodes);
    const uint32_t nb = idx << 7;
    const uint32_t ax = nb | (rows & 0xFFu), ay = nb | ((rows >> 8) & 0xFFu), az = nb | (rows >> 16);
    const float4 nx = ldRow(base, ax), fx = ldRow(base, ax ^ 16u);
    const float4 ny = ldRow(base, ay), fy = ldRow(base, ay ^ 16u);
    const float4 nz = ldRow(base, az), fz = ldRow(base, az ^ 16u);
    const uint4 en = *reinterpret_cast<const uint4*>(base + (nb | 112u));
    const float kUp = 1.0000004f;
    const float NX[4] = {nx.x, nx.y, nx.z, nx.w}, FX[4] = {fx.x, fx.y, fx.z, fx.w};
    const float NY[4] = {ny.x, ny.y, ny.z, ny.w}, FY[4] = {fy.x, fy.y, fy.z, fy.w};
    const float NZ[4] = {nz.x, nz.y, nz.z, nz.w}, FZ[4] = {fz.x, fz.y, fz.z, fz.w};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float tmin = fmaxf(fmaxf((NX[c] - o.x) * idir.x, (NY[c] - o.y) * idir.y), fmaxf((NZ[c] - o.z) * idir.z, tnear));
        const float tmax = fminf(fminf((FX[c] - o.x) * idir.x, (FY[c] - o.y) * idir.y), fminf((FZ[c] - o.z) * idir.z, tfar));
        tent[c] = tmin;
        hit[c] = tmin <= tmax * kUp;
    }
    ent[0] = en.x; ent[1] = en.y; ent[2] = en.z; ent[3] = en.w;
}

// push the hit children farthest first; the nearest is not pushed but
// returned in nxt (hasNext), so the next step needs no stack round trip.
// Branch-free: a missed child becomes (key +inf, entry E_NONE), the
// 5-comparator network sorts descending by key with selects, so the misses
// come first and the nearest hit (if any) last; entries are pushed unless
// E_NONE.  Hits at equal keys may come out in either order -- the closest
// hit does not depend on the visiting order.
DEV void pushSorted(Stack& st, Counters& c, uint32_t e[4], float t[4], const bool h[4], uint32_t& nxt, bool& hasNext) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        t[i] = h[i] ? t[i] : INFINITY;
        e[i] = h[i] ? e[i] : (uint32_t)E_NONE;
    }
#define OMV_CSWAP(a, b)                                             \
    {                                                               \
        const bool sw = t[b] > t[a];                                \
        const float ta = t[a], tb = t[b];                           \
        const uint32_t ea = e[a], eb = e[b];                        \
        t[a] = sw ? tb : ta; t[b] = sw ? ta : tb;                   \
        e[a] = sw ? eb : ea; e[b] = sw ? ea : eb;                   \
    }
    OMV_CSWAP(0, 1) OMV_CSWAP(2, 3) OMV_CSWAP(0, 2) OMV_CSWAP(1, 3) OMV_CSWAP(1, 2)
#undef OMV_CSWAP
    const bool v[3] = {e[0] != (uint32_t)E_NONE, e[1] != (uint32_t)E_NONE, e[2] != (uint32_t)E_NONE};
    st.push3If(e, v, c);
    nxt = e[3];
    hasNext = e[3] != (uint32_t)E_NONE;
}

// Closest hit of a world-space ray over the two-level BVH, as a resumable
// traversal: start() sets the TLAS root as the next entry, every step()
// consumes one entry (TLAS/BLAS node, leaf, instance entry or instance end):
// the nearest hit child of the last node when there is one (kept in a
// register), else the top of the stack; more() says whether entries remain.  The result
// does not depend on the visiting order: boxes are tested conservatively
// against the current closest t, triangles accept t in (tnear, t_closest],
// and offer() keeps the (t, instance, primitive) minimum.
struct Trav {
    f3 wo, wd, widir;   // world ray (TLAS boxes, instance transforms)
    Pre op;             // object ray of the current instance (BLAS boxes, triangles)
    uint32_t wrows, orows;   // nearRows of the world and object rays
    Hit h;
    float tnear, tfar;
    float tstop;     // occlusion early-out: the traversal ends at the first hit with t <= tstop (< 0: closest hit only)
    int curInst;
    bool found;
    uint32_t nxt;    // next entry when hasNext
    bool hasNext;
    DEV bool more(const Stack& st) const { return hasNext || st.sp > 0; }
    DEV bool start(const OmvSceneView& S, const f3& o, const f3& d, float tn, float tf, Stack& st, Counters& c) {
        found = false;
        hasNext = false;
        h.tie = false; h.t = tf; h.u = 0.0f; h.v = 0.0f; h.inst = -1; h.prim = 0;
        tnear = tn; tfar = tf;
        tstop = -1.0f;
        st.sp = 0;
        const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
        if (!(ax > 0.0f || ay > 0.0f || az > 0.0f)) return false;   // zero / NaN direction: no hit (as rayPrep)
        wo = o; wd = d;
        widir = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        wrows = nearRows(widir);
        orows = wrows;
        curInst = -1;
        nxt = (E_TLAS_NODE << E_KIND_SHIFT) | (uint32_t)S.tlas_root;
        hasNext = true;
        return true;
    }
    DEV uint32_t take(Stack& st) {
        if (hasNext) { hasNext = false; return nxt; }
        return st.pop();
    }
    DEV static bool isLeaf(uint32_t e) { return e != E_END && (e >> E_KIND_SHIFT) == E_BLAS_LEAF; }
    // triangles of a BLAS leaf entry (object ray of the current instance)
    DEV void leaf(const OmvSceneView& S, uint32_t e, Stack& st, Counters& c) {
        const uint32_t idx = e & 0x3FFFFFFFu;
        const uint32_t first = idx >> 3, cnt = (idx & 7u) + 1u;
        for (uint32_t k = 0; k < cnt; ++k) {
            const float4* rec = reinterpret_cast<const float4*>(S.tris) + 3u * (first + k);
            float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
            ++c.tris;
            float t, u, v;
            if (triIntersect(op, r0, r1, r2, tnear, found ? h.t : tfar, t, u, v)) offer(h, found, t, u, v, curInst, __float_as_int(r0.w));
        }
        // a hit at or before tstop decides an occlusion query: the closest hit
        // is at least as near, so the traversal ends here
        if (found && h.t <= tstop) { st.sp = 0; hasNext = false; }
    }
    // any other entry: instance end, instance entry, TLAS or BLAS node
    DEV void visit(const OmvSceneView& S, uint32_t e, Stack& st, Counters& c) {
        if (e == E_END) { curInst = -1; return; }
        const uint32_t kind = e >> E_KIND_SHIFT;
        const uint32_t idx = e & 0x3FFFFFFFu;
        const float tb = found ? h.t : tfar;
        if (kind == E_INST) {
            const int inst = (int)(idx & 0x1FFFFFFFu);
            const OmvInstance& in = S.insts[inst];
            const float* I = in.inv;
            const f3 o = wo, d = wd;
            f3 oo = mk3(I[0] * o.x + I[1] * o.y + I[2] * o.z + I[3], I[4] * o.x + I[5] * o.y + I[6] * o.z + I[7],
                        I[8] * o.x + I[9] * o.y + I[10] * o.z + I[11]);
            f3 od = mk3(I[0] * d.x + I[1] * d.y + I[2] * d.z, I[4] * d.x + I[5] * d.y + I[6] * d.z, I[8] * d.x + I[9] * d.y + I[10] * d.z);
            if (!rayPrep(oo, od, op)) return;
            orows = nearRows(op.idir);
            curInst = inst;
            st.push(E_END, c);
            nxt = (E_BLAS_NODE << E_KIND_SHIFT) | (uint32_t)in.blas_root;
            hasNext = true;
            return;
        }
        // internal node (TLAS uses the world ray, BLAS the object ray)
        const bool tl = kind == E_TLAS_NODE;
        ++c.nodes;
        float tent[4]; bool hit[4]; uint32_t ent[4];
        boxTest4Rows(S.nodes, idx, tl ? wrows : orows, tl ? wo : op.o, tl ? widir : op.idir, tnear, tb, tent, hit, ent);
        pushSorted(st, c, ent, tent, hit, nxt, hasNext);
    }
    // one entry, whatever its kind (the persistent traversal's unit of work)
    DEV void step(const OmvSceneView& S, Stack& st, Counters& c) {
        const uint32_t e = take(st);
        if (isLeaf(e)) leaf(S, e, st, c);
        else visit(S, e, st, c);
    }
};

DEV bool traceClosest(const OmvSceneView& S, const f3& o, const f3& d, float tnear, float tfar, Hit& h, Stack& st, Counters& c,
                      float tstop = -1.0f) {
    Trav T;
    const bool ok = T.start(S, o, d, tnear, tfar, st, c);
    T.tstop = tstop;
    if (ok) {
        while (T.more(st)) T.step(S, st, c);
    }
    h = T.h;
    return ok && T.found;
}

// --------------------------------------------------------- path helpers
struct RP {   // renderer parameters used per sample
    int camera_bounces, light_bounces, q, atten;
    float light_scale, light_min_dist, cut;
    uint32_t seed;
    Ext x;
};

// sceneIntersect (BDPTRenderer.cpp:75-102); tnear persists across calls of one path.
// tstop >= 0 makes it an occlusion query: the first hit found with t <= tstop
// is returned (instead of the closest), which decides "hit before tstop"
// exactly as the closest hit would.  Scenes with alpha-MASK materials keep the
// closest hit (a cut-out hit is re-traced past, up to 10 times).
// One step of sceneIntersect's mask loop after a trace that found h: 0 = the
// hit is final, 1 = an alpha-MASK cut-out (tnear moved past it, tfar back to
// t1: trace again), 2 = a cut-out with nothing left beyond it (no hit).
DEV int maskStep(const OmvSceneView& S, const f3& org, const f3& dir, const Hit& h, float& tnear, float& tfar, float t1) {
    if (!S.any_masked) return 0;   // no alpha-MASK material: every hit is final
    const int mesh = S.insts[h.inst].mesh_id;
    const OmvMaterial& M = S.mats[S.meshes[mesh].mat_id];
    if (M.type != OMV_MAT_PBR || M.alpha_mode != OMV_ALPHA_MASK) return 0;
    if (!isMaskedHit(S, mesh, h.prim, mk3(1.0f - h.u - h.v, h.u, h.v))) return 0;
    f3 o2 = mk3(org.x + h.t * dir.x, org.y + h.t * dir.y, org.z + h.t * dir.z);
    float v = maxAbs(o2);
    v = v > h.t ? v : h.t;
    float delta = 3.0f * FLT_EPSILON * v;
    tnear = h.t + delta;
    tfar = t1;
    return tfar <= tnear ? 2 : 1;
}

DEV bool sceneIntersect(const OmvSceneView& S, const f3& org, const f3& dir, float& tnear, float tfar, Hit& h, Stack& st, Counters& c, uint32_t& rays,
                        float tstop = -1.0f) {
    const float t1 = tfar;
    if (S.any_masked || S.closest_only) tstop = -1.0f;
    for (int itr = 10; itr > 0; --itr) {
        ++rays;
        if (!traceClosest(S, org, dir, tnear, tfar, h, st, c, tstop)) return false;
        const int m = maskStep(S, org, dir, h, tnear, tfar, t1);
        if (m == 0) return true;
        if (m == 2) return false;
    }
    return false;
}

// computeLightAttenuation (BDPTRenderer.cpp:104-126)
DEV float attenuation(const RP& P, float d, float attD) {
    if (attD >= FLT_MAX) {
        d = d < P.light_min_dist ? P.light_min_dist : d;
        return P.atten == 1 ? 1.0f / d : (P.atten == 2 ? 1.0f / (d * d) : 1.0f);
    }
    float rD = mxf(1.0f - d / attD, 0.0f);
    return P.atten == 1 ? rD : (P.atten == 2 ? sqrtf(rD) : (d < attD ? 1.0f : 0.0f));
}

// isConnected (BDPTRenderer.cpp:208-235)
// Occlusion bound of the connection ray from A to B (isConnected): a hit at
// t <= connectStop is at least 64 ulps of the coordinates' scale farther from
// B than isConnected's acceptance radius 10 eps max(1, |B|), counting the
// rounding of len, of the direction and of the hit point, so the closest hit
// -- no farther than that one -- makes B invisible.  Hits beyond it are
// decided by the closest hit, as before.
DEV float connectStop(const f3& A, const f3& B) {
    const float len = length3(sub3(B, A));
    const float safeB = 10.0f * FLT_EPSILON * maxAbs(B);
    const float margin = 64.0f * FLT_EPSILON * (maxAbs(A) + maxAbs(B) + 2.0f * len);
    return len - safeB - margin;
}
DEV bool isConnected(const OmvSceneView& S, const f3& A, const f3& B, Stack& st, Counters& c, uint32_t& rays, uint32_t& shadow) {
    f3 dir = sub3(B, A);
    float len = length3(dir);
    float maxsz = maxAbs(A);
    float safeD = 2.0f * 3.0f * FLT_EPSILON * maxsz;
    if (len <= safeD) return true;
    dir = mk3(dir.x / len, dir.y / len, dir.z / len);
    float tnear = safeD;
    f3 org = mk3(A.x + 0.0f * dir.x, A.y + 0.0f * dir.y, A.z + 0.0f * dir.z);
    float tfar = len + (0.1f * len > safeD ? 0.1f * len : safeD);
    Hit h;
    ++shadow;
    if (!sceneIntersect(S, org, dir, tnear, tfar, h, st, c, rays, connectStop(A, B))) return true;
    f3 dd = sub3(hitPosition(S, h), B);
    float sd = dot3(dd, dd);
    maxsz = maxAbs(B);
    safeD = 10.0f * FLT_EPSILON * maxsz;
    return sd <= safeD * safeD;
}

// HDRBackground::getRadiance (HdrBackground.cpp:55-97) / Background::getRadiance
DEV f3 background(const OmvSceneView& S, const f3& dir) {
    if (S.bg_type == OMV_BG_CONSTANT) return mk3(S.bg_const[0], S.bg_const[1], S.bg_const[2]);
    f3 rdir = mulT33(dir, S.bg_rot);
    f3 rxy = normalize3(mk3(rdir.x, rdir.y, 0.0f));
    float xang = acosf(clampr(rxy.x, -1.0f, 1.0f));
    xang = rxy.y < 0.0f ? 2.0f * kPi - xang : xang;
    float yang = asinf(clampr(rdir.z, -1.0f, 1.0f));
    yang += 0.5f * kPi;
    float u = xang / (2.0f * kPi);
    float v = 1.0f - yang / kPi;
    const int W = S.bg_w, H = S.bg_h;
    float upage = floorf(u - 0.5f / (float)W), vpage = floorf(v - 0.5f / (float)H);
    u -= upage; v -= vpage; u *= (float)W; v *= (float)H;
    int iu = (int)u, iv = (int)v;
    float kui = u - (float)iu, kvi = v - (float)iv, ku = 1.0f - kui, kv = 1.0f - kvi;
    int l = iu % W, r = (iu + 1) % W, b = iv % H, t = (iv + 1) % H;
    const int n = W * H - 1;
    int i0 = min(max(b * W + l, 0), n), i1 = min(max(b * W + r, 0), n), i2 = min(max(t * W + l, 0), n), i3 = min(max(t * W + r, 0), n);
    float w0 = ku * kv, w1 = kui * kv, w2 = ku * kvi, w3 = kui * kvi;
    const float4* px = reinterpret_cast<const float4*>(S.bg);
    float4 p0 = px[i0], p1 = px[i1], p2 = px[i2], p3 = px[i3];
    f3 val = mk3(p0.x * w0 + p1.x * w1 + p2.x * w2 + p3.x * w3, p0.y * w0 + p1.y * w1 + p2.y * w2 + p3.y * w3,
                 p0.z * w0 + p1.z * w1 + p2.z * w2 + p3.z * w3);
    return mk3(val.x * S.bg_scale, val.y * S.bg_scale, val.z * S.bg_scale);
}

// ------------------------------------------------ OMV_EXT_HDRI_NEE
// Next-event estimation toward the HDR environment at opaque PBR camera
// vertices, combined with the escape of the BSDF-sampled direction by the
// balance heuristic.  The environment sampler picks a row and a column from
// the alias tables built at commit and a uniform point in the texel; its
// solid-angle density is texPdf * W * H / (2 pi^2 sin(theta)).

// Inverse of background()'s mapping for the point (ux, uy) of texel (row, col)
DEV f3 envDirection(const OmvSceneView& S, int row, int col, float ux, float uy, float& sinT) {
    const float u = ((float)col + ux) / (float)S.bg_w, v = ((float)row + uy) / (float)S.bg_h;
    const float yang = (1.0f - v) * kPi;
    const float rz = -cosf(yang);
    const float xang = u * 2.0f * kPi;
    const float r = sqrtf(mxf(1.0f - rz * rz, 0.0f));
    sinT = r;
    return mul33(S.bg_rot, mk3(r * cosf(xang), r * sinf(xang), rz));
}
// Solid-angle density of the environment sampler at dir
DEV float envPdf(const OmvSceneView& S, const f3& dir) {
    f3 rdir = mulT33(dir, S.bg_rot);
    f3 rxy = normalize3(mk3(rdir.x, rdir.y, 0.0f));
    float xang = acosf(clampr(rxy.x, -1.0f, 1.0f));
    xang = rxy.y < 0.0f ? 2.0f * kPi - xang : xang;
    const float z = clampr(rdir.z, -1.0f, 1.0f);
    const float yang = asinf(z) + 0.5f * kPi;
    const float u = xang / (2.0f * kPi), v = 1.0f - yang / kPi;
    const int W = S.bg_w, H = S.bg_h;
    const int col = min(max((int)(u * (float)W), 0), W - 1), row = min(max((int)(v * (float)H), 0), H - 1);
    const float sinT = sqrtf(mxf(1.0f - z * z, 0.0f));
    if (!(sinT > 0.0f)) return 0.0f;
    return S.env_tex_pdf[(size_t)row * W + col] * (float)W * (float)H / (2.0f * kPi * kPi * sinT);
}
// Density of sampleIndirect's lobe mixture (BRDFModel.h:265-337) at an
// opaque vertex: diffuse P LN / pi, reflection (1 - P) D NH / (4 LH)
DEV float bsdfPdf(const Bsdf& b, const f3& O) {
    const float VN = -dot3(b.I, b.N);
    if (VN < 0.0f) return 0.0f;
    const float P = 0.8f * fabsf(VN) + 0.1f;
    const float LN = dot3(O, b.N);
    if (!(LN > 0.0f)) return 0.0f;
    float p = P * LN / kPi;
    const f3 H = normalize3(sub3(O, b.I));
    const float NH = dot3(b.N, H), LH = dot3(O, H);
    if (NH > 0.0f && LH > 0.0f) p += (1.0f - P) * ggxPdf(NH, b.alpha) * NH / (4.0f * LH);
    return p;
}
// The estimator's integrand per solid angle (sampled weight x density x
// cosine, bsdfEvaluate's clamps included) of direction O
DEV f3 bsdfEffCos(const Bsdf& b, const f3& O) {
    const f3 H = normalize3(sub3(O, b.I));
    const float VNc = clampr(fabsf(dot3(b.I, b.N)), 0.0001f, 1.0f);
    const float LNc = clampr(fabsf(dot3(O, b.N)), 0.0001f, 1.0f);
    const float LN = fabsf(dot3(O, b.N));
    const float NH = fabsf(dot3(b.N, H)), VH = fabsf(dot3(b.I, H));
    const float p5 = powf(clamp01(1.0f - VH), 5.0f);
    const float D = ggxPdf(NH, b.alpha);
    const float G = smithG2(smithG1(VNc, b.alpha), smithG1(LNc, b.alpha));
    const float col[3] = {b.c0, b.c1, b.c2};
    float out[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float F0 = lerpr(0.04f, col[c], b.metal);
        const float F = lerpr(p5, 1.0f, F0);
        const float fd = (1.0f - F) * ((1.0f - b.metal) * col[c]) / kPi;
        out[c] = fd * LN * LN / LNc + F * G * D * LN / (4.0f * LNc * LNc);
    }
    return mk3(out[0], out[1], out[2]);
}
// Next-event estimate at camera vertex i (after its emission): the MIS
// weight escW of an escape along the sampled direction outV, and, when the
// estimate is non-zero, its contribution c and the visibility ray (org, w)
// toward the environment (visible iff sceneIntersect finds no hit).
DEV bool neePrepare(const OmvSceneView& S, const RP& P, const SP& sp, const BounceInfo& bi, const f3& ofPrev, const f3& outV, int i,
                    uint64_t sample, float& escW, f3& c, f3& org, f3& w) {
    escW = 1.0f;
    if (!P.x.nee || !S.env_tex_pdf || bi.kind != 1) return false;
    const Bsdf& b = bi.b;
    if (b.transmissive || -dot3(b.I, b.N) < 0.0f || i >= P.camera_bounces || i + 1 > P.q) return false;
    const float pbS = bsdfPdf(b, outV), peS = envPdf(S, outV);
    if (pbS + peS > 0.0f) escW = pbS / (pbS + peS);
    Rng ns; ns.init(P.seed, STREAM_NEE + 16u * (uint32_t)i, sample);
    const float u0 = ns.next(), u1 = ns.next(), u2 = ns.next(), u3 = ns.next(), u4 = ns.next(), u5 = ns.next();
    const int W = S.bg_w, H = S.bg_h;
    int k = min((int)(u0 * (float)H), H - 1);
    const int row = u1 < S.env_row_prob[k] ? k : S.env_row_alias[k];
    k = min((int)(u2 * (float)W), W - 1);
    const size_t rk = (size_t)row * W + k;
    const int col = u3 < S.env_col_prob[rk] ? k : S.env_col_alias[rk];
    float sinT;
    w = envDirection(S, row, col, u4, u5, sinT);
    if (!(sinT > 0.0f)) return false;
    const float pe = S.env_tex_pdf[(size_t)row * W + col] * (float)W * (float)H / (2.0f * kPi * kPi * sinT);
    const float pb = bsdfPdf(b, w);
    if (!(pe > 0.0f) || !(dot3(w, b.N) > 0.0f)) return false;
    const f3 f = bsdfEffCos(b, w);
    const f3 bg = background(S, w);
    const float inv = 1.0f / (pe + pb);
    c = mk3(ofPrev.x * f.x * bg.x * inv, ofPrev.y * f.y * bg.y * inv, ofPrev.z * f.z * bg.z * inv);
    if (c.x == 0.0f && c.y == 0.0f && c.z == 0.0f) return false;
    const float delta = 3.0f * FLT_EPSILON * maxAbs(sp.pos);
    org = mk3(sp.pos.x + delta * w.x, sp.pos.y + delta * w.y, sp.pos.z + delta * w.z);
    return true;
}

// ---------------------------------------------- OMV_EXT_LIGHT_SPLAT
// Light-tracer (t = 1) splats.  Particle weight of a light subpath at its
// first hit: beta_1 = I * lightScale / p(w) * att(r1) * r1^2 with p = 2/pi,
// the density of uniformSphere's + octant; beta_{i+1} = beta_i * w_i *
// |w_out . n_i| (the camera subpath's throughput update).
DEV f3 splatBeta1(const RP& P, const f3& intensity, float att, float r1) {
    const float k1 = P.light_scale * (kPi * 0.5f) * att * r1 * r1;
    return mk3(intensity.x * k1, intensity.y * k1, intensity.z * k1);
}
// Pinhole connection of light vertex sp (arrived along inV) with weight
// beta: the pixel it projects to, its contribution
//   beta * getBrdf(inV, w_c) * |w_c . n| * f^2 / (d^2 cos^3(theta))
// and the visibility ray (from the vertex toward the pinhole, tnear = the
// isConnected offset, tfar = d; visible iff nothing is hit).
DEV bool splatPrepare(const OmvSceneView& S, const Ext& X, const omv_camera_desc& c, const SP& sp, const f3& inV, const f3& beta,
                      f3& ctb, uint32_t& pix, f3& wc, float& d, float& tn) {
    const f3 C = ld3(c.position);
    const f3 y = sp.pos;
    const f3 v = sub3(C, y);
    d = length3(v);
    if (!(d > 0.0f)) return false;
    wc = mk3(v.x / d, v.y / d, v.z / d);
    const f3 l = mulT33(sub3(y, C), c.rotation);
    if (!(l.z > 0.0f)) return false;
    const float px = c.focus * l.x / l.z + c.half_w, py = c.focus * l.y / l.z + c.half_h;
    if (!(px >= 0.0f && px < (float)c.width && py >= 0.0f && py < (float)c.height)) return false;
    const int ix = min((int)px, c.width - 1), iy = min((int)py, c.height - 1);
    const float cosT = l.z / length3(l);
    const f3 f = matGetBrdf(S, X, sp, inV, wc);
    const float k = c.focus * c.focus * fabsf(dot3(wc, sp.n)) / (d * d * cosT * cosT * cosT);
    ctb = mk3(beta.x * f.x * k, beta.y * f.y * k, beta.z * f.z * k);
    if (ctb.x == 0.0f && ctb.y == 0.0f && ctb.z == 0.0f) return false;
    if (!(isfinite(ctb.x) && isfinite(ctb.y) && isfinite(ctb.z))) return false;
    pix = (uint32_t)iy * (uint32_t)c.width + (uint32_t)ix;
    tn = 2.0f * 3.0f * FLT_EPSILON * maxAbs(y);
    return true;
}
// Film accumulation of splats with wave-aggregated atomics: the active lanes
// of the wave that add to the same pixel are summed first and one lane adds
// the sum; lanes alone on their pixel add in one vector atomic.  Call with
// every active lane (want = false for lanes with nothing to add).
DEV void splatAdd(float* film, bool want, uint32_t pix, const f3& v) {
    const int lane = threadIdx.x & 63;
    uint64_t pend = __ballot(want);
    bool solo = false;
    while (pend) {
        const int leader = __ffsll((unsigned long long)pend) - 1;
        const uint32_t pl = __shfl(pix, leader, 64);
        const uint64_t m = __ballot(want && pix == pl) & pend;
        if (m == (1ull << leader)) {
            solo = solo || lane == leader;
        } else {
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            for (uint64_t mm = m; mm; mm &= mm - 1ull) {
                const int j = __ffsll((unsigned long long)mm) - 1;
                sx += __shfl(v.x, j, 64); sy += __shfl(v.y, j, 64); sz += __shfl(v.z, j, 64);
            }
            if (lane == leader) {
                atomicAdd(film + (size_t)pl * 3 + 0, sx);
                atomicAdd(film + (size_t)pl * 3 + 1, sy);
                atomicAdd(film + (size_t)pl * 3 + 2, sz);
            }
        }
        pend &= ~m;
    }
    if (solo) {
        atomicAdd(film + (size_t)pix * 3 + 0, v.x);
        atomicAdd(film + (size_t)pix * 3 + 1, v.y);
        atomicAdd(film + (size_t)pix * 3 + 2, v.z);
    }
}

// DistanceRandom::getValue (DistanceRandom.h:203-220)
DEV int sampleLight(const OmvSceneView& S, float u) {
    const int n = S.num_lights;
    if (n <= 0) return -1;
    float d = u * S.light_max;
    if (n == 1) return 0;
    if (d <= S.light_cdf[0]) return 0;
    if (d >= S.light_cdf[n - 2]) return n - 1;
    int L = 0, R = n - 1;
    while (L < R) { int C = (L + R) / 2; if (d <= S.light_cdf[C]) R = C; else L = C + 1; }
    return (L + R) / 2;
}

// PinholeCamera::getRay (PinholeCamera.cpp:80-95)
DEV void cameraRay(const omv_camera_desc& c, int x, int y, float u0, float u1, f3& O, f3& D) {
    float px = (float)x + 0.5f - c.half_w;
    float py = (float)y + 0.5f - c.half_h;
    float dx = lerpr(-0.5f, 0.5f, u0);
    float dy = lerpr(-0.5f, 0.5f, u1);
    f3 cr = normalize3(mk3(px + dx, py + dy, c.focus));
    D = mul33(c.rotation, cr);
    O = ld3(c.position);
}

DEV uint64_t waveSum(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

DEV void flushCounters(unsigned long long* g, uint32_t rays, uint32_t cam, uint32_t light, uint32_t shadow, const Counters& c) {
    uint64_t v[7] = {rays, cam, light, shadow, c.nodes, c.tris, c.overflow};
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        uint64_t s = waveSum(v[i]);
        if ((threadIdx.x & 63) == 0 && s) atomicAdd(g + i, (unsigned long long)s);
    }
}

inline RP toRP(const omv_renderer_params& p) {
    RP r;
    r.camera_bounces = p.camera_bounces; r.light_bounces = p.light_bounces; r.q = p.max_path_length;
    r.atten = p.light_distance_attenuation; r.light_scale = p.light_scale_value; r.light_min_dist = p.light_min_distance;
    r.cut = p.ray_cut_pix_value; r.seed = p.seed;
    r.x.omRough = (p.extensions & OMV_EXT_OM_ROUGH) != 0;
    r.x.omAlpha = p.om_roughness * p.om_roughness;
    r.x.nee = (p.extensions & OMV_EXT_HDRI_NEE) != 0;
    r.x.splat = (p.extensions & OMV_EXT_LIGHT_SPLAT) != 0;
    return r;
}

}  // namespace

// ================================================= per-sample path function
namespace {

// Per-lane state of a path kernel: traversal stack, counters and this lane's
// column of the light-vertex scratch ([vertex][component][lane]).
struct Lane {
    Stack st;
    Counters cnt;
    uint32_t rays, camRays, lightRays, shadowRays;
    float* splat;   // OMV_EXT_LIGHT_SPLAT film (W*H*3 sums), or null
    float* lv;
    uint64_t lvStride;
    DEV float& lvAt(int vtx, int c) { return lv[((uint64_t)vtx * 6 + c) * lvStride]; }
};

DEV void laneInit(Lane& L, uint32_t* gstack, float* lv, uint32_t gtid, uint32_t gthreads) {
    L.st.init(gstack, gtid, gthreads);
    L.cnt.nodes = 0; L.cnt.tris = 0; L.cnt.overflow = 0;
    L.rays = L.camRays = L.lightRays = L.shadowRays = 0;
    L.splat = nullptr;
    L.lv = lv ? lv + gtid : nullptr; L.lvStride = gthreads;
}

// The light subpath as the camera rounds use it: the sampled light (idx < 0:
// none), its position, intensity and range, and the vertex count.
struct LightSub {
    int idx, len;
    f3 pos, intensity;
    float range;
};
// Light vertices l >= 1 (position, outgoing radiance) in a lane's scratch
struct LaneLightVerts {
    Lane* L;
    DEV f3 pos(int l) const { return mk3(L->lvAt(l, 0), L->lvAt(l, 1), L->lvAt(l, 2)); }
    DEV f3 rad(int l) const { return mk3(L->lvAt(l, 3), L->lvAt(l, 4), L->lvAt(l, 5)); }
};

// Camera rounds i0 .. camera_bounces of one sample: computePath(inv = true)
// and the connections of every camera vertex (BDPTRenderer.cpp:149-176,
// 260-297), then the escape term (:251-258), from the state the rounds
// before i0 left: the ray (origin already offset from its surface point,
// direction, tnear), throughput, escape weight (OMV_EXT_HDRI_NEE), the
// radiance and connection sums so far, the vertex count and the camera
// stream.  Returns the sample, conn + radiance.  The per-sample function runs
// it from round 1; the wavefront schedule hands it the paths still alive at
// a late round with the state its kernels keep, so both give one sample bit
// for bit.
template <bool kLights, bool kNee, class LV>
DEV f3 cameraRounds(const OmvSceneView& S, const RP& P, uint64_t s, int i0, f3 org, f3 pDir, float tnear, f3 of, float escW, f3 radiance,
                    f3 conn, int n, Rng& cs, const LightSub& ls, const LV& lv, Lane& L) {
    const int q = P.q;
    bool exitScene = false;
    for (int i = i0; i < P.camera_bounces + 1; ++i) {
        ++L.camRays;
        Hit h;
        if (!sceneIntersect(S, org, pDir, tnear, FLT_MAX, h, L.st, L.cnt, L.rays)) { exitScene = true; break; }
        SP sp;
        computeSurfacePoint(S, h, sp, false);
        modifyFrame(S, sp);
        MatIn mi;   // PBR texture inputs, read once for the connections and the bounce (kLights)
        if (kLights) matInputs(S, sp, mi);
        // connections of camera vertex j = i with light vertices l, l + j <= q
        if (kLights && ls.idx >= 0) {
            const int lEnd = min(ls.len, q - i + 1);
            const f3 outV = neg3(pDir);
            for (int l = 0; l < lEnd; ++l) {
                f3 lPos = l == 0 ? ls.pos : lv.pos(l);
                f3 lRad;
                f3 LC = normalize3(sub3(sp.pos, lPos));
                if (l == 0) {
                    float lsc = attenuation(P, length3(sub3(sp.pos, lPos)), ls.range);
                    float k = P.light_scale * lsc;
                    lRad = mk3(ls.intensity.x * k, ls.intensity.y * k, ls.intensity.z * k);
                } else {
                    lRad = lv.rad(l);
                }
                float c = fabsf(dot3(LC, sp.n));
                f3 lc = mk3(lRad.x * c * of.x, lRad.y * c * of.y, lRad.z * c * of.z);
                if (intensity(lc) < P.cut) continue;
                f3 brdf = matGetBrdf(S, P.x, sp, LC, outV, &mi);
                brdf = mk3(brdf.x * c, brdf.y * c, brdf.z * c);
                lc = mk3(brdf.x * of.x * lRad.x, brdf.y * of.y * lRad.y, brdf.z * of.z * lRad.z);
                if (intensity(lc) < P.cut) continue;
                if (!isConnected(S, lPos, sp.pos, L.st, L.cnt, L.rays, L.shadowRays)) continue;
                conn = mk3(lc.x + conn.x, lc.y + conn.y, lc.z + conn.z);
            }
        }
        f3 outRay, brdf, erad;
        BounceInfo bi;
        matGetRayAndBrdf(S, P.x, sp, pDir, outRay, brdf, erad, true, cs, kNee ? &bi : nullptr, kLights ? &mi : nullptr);
        float ct = fabsf(dot3(outRay, sp.n));
        const f3 ofPrev = of;
        of = mk3(of.x * brdf.x * ct, of.y * brdf.y * ct, of.z * brdf.z * ct);
        radiance = mk3(radiance.x + of.x * erad.x, radiance.y + of.y * erad.y, radiance.z + of.z * erad.z);
        if (kNee && P.x.nee) {   // OMV_EXT_HDRI_NEE, after the vertex's emission
            f3 c, eo, w;
            if (neePrepare(S, P, sp, bi, ofPrev, outRay, i, s, escW, c, eo, w)) {
                float tn = 0.0f;
                Hit hs;
                ++L.shadowRays;
                if (!sceneIntersect(S, eo, w, tn, FLT_MAX, hs, L.st, L.cnt, L.rays, FLT_MAX))
                    radiance = mk3(radiance.x + c.x, radiance.y + c.y, radiance.z + c.z);
            }
        }
        ++n;
        pDir = outRay;
        // exact shortcuts, as in the oracle: zero throughput, or no emission past q
        if (of.x == 0.0f && of.y == 0.0f && of.z == 0.0f) break;
        if (!S.any_emissive && n > q) break;
        const float delta = 3.0f * FLT_EPSILON * maxAbs(sp.pos);
        org = mk3(sp.pos.x + delta * outRay.x, sp.pos.y + delta * outRay.y, sp.pos.z + delta * outRay.z);
    }
    if (exitScene && S.bg_type != OMV_BG_NONE && n <= q) {
        f3 bg = background(S, pDir);
        bg = mk3(bg.x * of.x, bg.y * of.y, bg.z * of.z);
        if (kNee && P.x.nee) bg = mk3(bg.x * escW, bg.y * escW, bg.z * escW);
        radiance = mk3(bg.x * 1.0f + radiance.x, bg.y * 1.0f + radiance.y, bg.z * 1.0f + radiance.z);
    }
    return mk3(conn.x * 1.0f + radiance.x, conn.y * 1.0f + radiance.y, conn.z * 1.0f + radiance.z);
}

// One sample of BDPTRenderer::computePaths (BDPTRenderer.cpp:237-301):
// sample s = pixel * spp + k.  kLights = false compiles no light subpath
// and no connection code (scenes without point lights, BDPTRenderer.cpp:241-242);
// kNee = false compiles no OMV_EXT_HDRI_NEE code, kSplat = false no
// OMV_EXT_LIGHT_SPLAT code.
template <bool kLights, bool kNee = false, bool kSplat = false>
DEV f3 pathSample(const OmvSceneView& S, const omv_camera_desc& cam, const RP& P, int spp, uint64_t s, Lane& L) {
    const int q = P.q;
    const int lvCap = max(min(P.light_bounces + 1, max(q, 1)), 1);
    const uint64_t pix = s / (uint64_t)spp;
    const int x = (int)(pix % (uint64_t)cam.width), y = (int)(pix / (uint64_t)cam.width);
    Rng cs; cs.init(P.seed, STREAM_CAMERA, s);
    float u0 = cs.next(), u1 = cs.next();
    f3 O, D;
    cameraRay(cam, x, y, u0, u1, O, D);

    // ---- light subpath (tryComputeLightPath + computePath(light), BDPTRenderer.cpp:128-206)
    int lpLen = 0, lightIdx = -1;
    f3 lightPos = mk3(0.0f, 0.0f, 0.0f), lightInt = mk3(0.0f, 0.0f, 0.0f);
    float lightRange = FLT_MAX;
    if (kLights && S.num_lights > 0) {
        Rng ls; ls.init(P.seed, STREAM_LIGHT, s);
        lightIdx = sampleLight(S, ls.next());
        const OmvLight Lt = S.lights[lightIdx];
        lightPos = ld3(Lt.pos); lightInt = ld3(Lt.intensity); lightRange = Lt.range;
        f3 dir = uniformSphere(ls);
        const f3 lRad = mk3(Lt.intensity[0] * P.light_scale, Lt.intensity[1] * P.light_scale, Lt.intensity[2] * P.light_scale);
        // vertex 0: the light itself
        L.lvAt(0, 0) = lightPos.x; L.lvAt(0, 1) = lightPos.y; L.lvAt(0, 2) = lightPos.z;
        L.lvAt(0, 3) = lRad.x; L.lvAt(0, 4) = lRad.y; L.lvAt(0, 5) = lRad.z;
        f3 pPos = lightPos, pDir = dir, of = mk3(1.0f, 1.0f, 1.0f), beta = of;
        float tnear = 0.0f, rD = 1.0f;
        int n = 1;
        // exact shortcut: light vertices with index >= max_path_length are never connected
        for (int i = 1; i < lvCap; ++i) {
            float delta = 3.0f * FLT_EPSILON * maxAbs(pPos);
            f3 org = mk3(pPos.x + delta * pDir.x, pPos.y + delta * pDir.y, pPos.z + delta * pDir.z);
            ++L.lightRays;
            Hit h;
            if (!sceneIntersect(S, org, pDir, tnear, FLT_MAX, h, L.st, L.cnt, L.rays)) break;
            SP sp;
            computeSurfacePoint(S, h, sp, false);
            modifyFrame(S, sp);
            f3 outRay, brdf, erad;
            matGetRayAndBrdf(S, P.x, sp, pDir, outRay, brdf, erad, false, ls);
            float ct = fabsf(dot3(pDir, sp.n));
            of = mk3(of.x * brdf.x * ct, of.y * brdf.y * ct, of.z * brdf.z * ct);
            if (i == 1) rD = attenuation(P, length3(sub3(sp.pos, lightPos)), lightRange);
            if (kSplat && P.x.splat) {   // OMV_EXT_LIGHT_SPLAT: every hit of the walk, before the hard cut
                if (i == 1) beta = splatBeta1(P, lightInt, rD, length3(sub3(sp.pos, lightPos)));
                f3 ctb, wc;
                uint32_t pix = 0;
                float d, tn;
                bool vis = false;
                if (splatPrepare(S, P.x, cam, sp, pDir, beta, ctb, pix, wc, d, tn)) {
                    Hit hs;
                    ++L.shadowRays;
                    vis = !sceneIntersect(S, sp.pos, wc, tn, d, hs, L.st, L.cnt, L.rays, FLT_MAX);
                }
                splatAdd(L.splat, vis, pix, ctb);
                const float co = fabsf(dot3(outRay, sp.n));
                beta = mk3(beta.x * brdf.x * co, beta.y * brdf.y * co, beta.z * brdf.z * co);
            }
            f3 outRad = mk3(of.x * lRad.x * rD, of.y * lRad.y * rD, of.z * lRad.z * rD);
            if (intensity(outRad) < P.cut) break;
            L.lvAt(i, 0) = sp.pos.x; L.lvAt(i, 1) = sp.pos.y; L.lvAt(i, 2) = sp.pos.z;
            L.lvAt(i, 3) = outRad.x; L.lvAt(i, 4) = outRad.y; L.lvAt(i, 5) = outRad.z;
            ++n;
            pPos = sp.pos; pDir = outRay;
        }
        lpLen = n;
    }

    // ---- camera subpath + connections
    LightSub lsub;
    lsub.idx = lightIdx; lsub.len = lpLen; lsub.pos = lightPos; lsub.intensity = lightInt; lsub.range = lightRange;
    LaneLightVerts lvs; lvs.L = &L;
    const float delta = 3.0f * FLT_EPSILON * maxAbs(O);
    const f3 org = mk3(O.x + delta * D.x, O.y + delta * D.y, O.z + delta * D.z);
    return cameraRounds<kLights, kNee>(S, P, s, 1, org, D, 0.0f, mk3(1.0f, 1.0f, 1.0f), 1.0f, mk3(0.0f, 0.0f, 0.0f),
                                       mk3(0.0f, 0.0f, 0.0f), 1, cs, lsub, lvs, L);
}

// One pixel of the normalised film: its spp samples summed in sample order
// (PinholeCamera::hit, PinholeCamera.cpp:97-108), then PinholeCamera::stop
// (rgb / count when count >= 0.5, alpha 1).  The sums and the scale are the
// film kernel's, so the pixel equals the per-sample schedule's bit for bit.
template <bool kLights, bool kNee = false>
DEV float4 renderPixel(const OmvSceneView& S, const omv_camera_desc& cam, const RP& P, int spp, uint64_t p, Lane& L) {
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int k = 0; k < spp; ++k) {
        f3 r = pathSample<kLights, kNee>(S, cam, P, spp, p * (uint64_t)spp + (uint64_t)k, L);
        a0 = a0 + r.x; a1 = a1 + r.y; a2 = a2 + r.z;
    }
    float cnt = (float)spp;
    float sc = cnt < 0.5f ? 1.0f : 1.0f / cnt;
    float4 o; o.x = sc * a0; o.y = sc * a1; o.z = sc * a2; o.w = 1.0f;
    return o;
}

}  // namespace

#endif
