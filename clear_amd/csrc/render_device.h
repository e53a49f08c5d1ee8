// Device-side shading for the gfx950 path tracers: textures, distributions,
// lights, BSDF mixtures, camera and the per-path integrator, reading the
// reference's byte-blob scene ABI (materialsData / texturesData /
// otherResources with byte offsets).
//
// The reference builds each BSDF/EDF into a 512-byte private byte pool
// (pathtracer.cl:57) that lands in scratch memory; here a BSDF is a small
// register-resident record (up to four BxDF components with their evaluated
// textures), which is what the blob walk of BSDFAlloc (reflection.cl:185-342)
// produces, minus the pointer chasing at every evaluation.
//
// Semantics follow the kernels (pathtracer.cl, camera.cl, rng.cl, light.cl,
// distribution.cl, reflection.cl, texture.cl) with the deviations listed in
// the renderer's design notes (32-bit face ids, per-pixel RNG, alpha-rejected
// hits do not shrink t, stale shadow face ids, Ashikhmin-Shirley theta flip
// by vout.z, parameterised bounce cap, environment marginal offsets, FP32
// image texture and Beckmann additions, normalised normal-map axis).
#pragma once

#include <hip/hip_runtime.h>

#include "traverse_device.h"

namespace clr {
namespace dev {

constexpr float kPiF = 3.14159274101257324219f;  // M_PI_F
constexpr float kRayEps = 0.00001f;              // EPSILON (global.cl:28)

// ---- render-side scene ----------------------------------------------------
struct RScene {
    GeomScene g;
    const clr_float3* normals;
    const clr_float3* tangents;
    const float* uvs;
    const clr_light_info* lights;
    const uint8_t* mat;
    const uint8_t* tex;
    const uint8_t* oth;
    const uint8_t* camera;
    const uint8_t* envLP;      // environment light property (LightPropertyInfo)
    const uint8_t* lightDist;  // Discrete1D over lights
    const uint32_t* faceLight; // face -> first light index with that face (or kNone)
    const int* envLightPtr;    // first light at infinity (INT_MAX if none), device word
    int envLight;              // resolved by finishScene: index or -1
    uint32_t* status;          // device status word (stack overflow flag)
};

template <typename T>
__device__ __forceinline__ T ld(const uint8_t* p) {
    return *reinterpret_cast<const T*>(p);
}
__device__ __forceinline__ V3 ld3(const uint8_t* p) {
    return {ld<float>(p), ld<float>(p + 4), ld<float>(p + 8)};
}
__device__ __forceinline__ V3 v3z() { return {0.0f, 0.0f, 0.0f}; }

// Resolves the blob header at kernel start (uniform loads).
__device__ __forceinline__ void finishScene(RScene& S) {
    const uint32_t camOff = ld<uint32_t>(S.oth), envOff = ld<uint32_t>(S.oth + 4),
                   distOff = ld<uint32_t>(S.oth + 8);
    S.camera = S.oth + camOff;
    S.envLP = S.mat + ld<uint32_t>(S.oth + envOff);
    S.lightDist = S.oth + distOff;
    const int e = *S.envLightPtr;
    S.envLight = e == 0x7FFFFFFF ? -1 : e;
}
__device__ __forceinline__ V3 v3s(float s) { return {s, s, s}; }
__device__ __forceinline__ V3& operator+=(V3& a, V3 b) {
    a = a + b;
    return a;
}
__device__ __forceinline__ V3 normalize3(V3 v) {
    const float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return v / l;
}
__device__ __forceinline__ float length3(V3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
__device__ __forceinline__ float lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return omin(omax(v, lo), hi); }

// ---- RNG (rng.cl:21-83): xorshift128 state held in registers --------------
struct Rng {
    uint32_t s0, s1, s2, s3;
    __device__ __forceinline__ uint32_t next() {
        const uint32_t t = s0 ^ (s0 << 11);
        s0 = s1;
        s1 = s2;
        s2 = s3;
        s3 = (s3 ^ (s3 >> 19)) ^ (t ^ (t >> 8));
        return s3;
    }
    __device__ __forceinline__ float uniform() {
        return __uint_as_float((next() >> 9) | 0x3f800000u) - 1.0f;
    }
};
__device__ __forceinline__ uint32_t randUInt(float u, uint32_t maxv) {
    const float f = (float)maxv * u;
    const uint32_t v = f > 0.0f ? (uint32_t)f : 0u;
    return min(v, maxv - 1u);
}
__device__ __forceinline__ void concentricDisk(float u1, float u2, float& dx, float& dy) {
    const float sx = 2 * u1 - 1, sy = 2 * u2 - 1;
    if (sx == 0.0f && sy == 0.0f) {
        dx = dy = 0.0f;
        return;
    }
    float r, th;
    if (sx >= -sy) {
        if (sx > sy) { r = sx; th = sy / sx; }
        else { r = sy; th = 2.0f - sx / sy; }
    } else {
        if (sx > sy) { r = -sy; th = 6.0f + sx / sy; }
        else { r = -sx; th = 4.0f + sy / sx; }
    }
    th *= kPiF / 4.0f;
    dx = r * cosf(th);
    dy = r * sinf(th);
}
__device__ __forceinline__ V3 cosineHemisphere(float u1, float u2) {
    float x, y;
    concentricDisk(u1, u2, x, y);
    return {x, y, sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y))};
}

// ---- textures (texture.cl:58-204) ------------------------------------------
__device__ __forceinline__ uint32_t texel(float c, uint32_t size) {
    const float f = fmodf((float)size * c, (float)size);
    const uint32_t v = f > 0.0f ? (uint32_t)f : 0u;
    return min(v, size - 1u);
}
// IEEE binary16 -> binary32 (exact; subnormals normalised).
__device__ __forceinline__ float half2float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, ex = (h >> 10) & 0x1Fu;
    uint32_t man = h & 0x3FFu, bits;
    if (ex == 0) {
        if (man == 0) {
            bits = sign;
        } else {
            int e = -1;
            do {
                ++e;
                man <<= 1;
            } while (!(man & 0x400u));
            bits = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3FFu) << 13);
        }
    } else if (ex == 31) {
        bits = sign | 0x7F800000u | (man << 13);
    } else {
        bits = sign | ((ex + 112) << 23) | (man << 13);
    }
    return __uint_as_float(bits);
}
__device__ __forceinline__ uint32_t checkerIdx(float u, float v) {
    const uint32_t a = u * 2 > 0 ? (uint32_t)(u * 2) : 0u;
    const uint32_t b = v * 2 > 0 ? (uint32_t)(v * 2) : 0u;
    return (a + b) % 2u;
}
__device__ V3 evalColor(const uint8_t* t, float u, float v) {
    const uint8_t type = t[0];
    if (type == 0) return ld3(t + 16);
    if (type == 4) {
        if (t[1] == 0) return ld3(t + 16 + 16 * checkerIdx(u, v));
        if (t[1] == 1) {
            const float hw = ld<float>(t + 4) * 0.5f;
            float uc = 0.0f, vc = 0.0f;
            const float au = fmodf(fabsf(u), 1.0f), av = fmodf(fabsf(v), 1.0f);
            if (au < hw * 0.5f || au > 1.0f - hw * 0.5f) uc = 1.0f;
            else if (au > 0.5f - hw * 0.5f && au < 0.5f + hw * 0.5f) uc = -1.0f;
            if (av < hw * 0.5f || av > 1.0f - hw * 0.5f) vc = 1.0f;
            else if (av > 0.5f - hw * 0.5f && av < 0.5f + hw * 0.5f) vc = -1.0f;
            if (av > 0.5f) uc *= -1;
            if (au > 0.5f) vc *= -1;
            if (t[8]) { uc *= -1; vc *= -1; }
            const V3 n = normalize3(v3(uc, vc, 1.0f));
            return {0.5f * n.x + 0.5f, 0.5f * n.y + 0.5f, 0.5f * n.z + 0.5f};
        }
        return v3z();
    }
    if (type == 1 || type == 2 || type == 3 || type == 5 || type == 9) {
        const uint32_t w = ld<uint32_t>(t + 4), h = ld<uint32_t>(t + 8);
        const size_t idx = (size_t)texel(v, h) * w + texel(u, w);
        const uint8_t* d = t + ld<int32_t>(t + 12);
        if (type == 1 || type == 2) {
            const uchar4 c = *reinterpret_cast<const uchar4*>(d + 4 * idx);
            return {(float)c.x / 255.0f, (float)c.y / 255.0f, (float)c.z / 255.0f};
        }
        if (type == 3) {
            const ushort4 c = *reinterpret_cast<const ushort4*>(d + 8 * idx);
            return {half2float(c.x), half2float(c.y), half2float(c.z)};
        }
        if (type == 5) {
            const float g = (float)d[idx] / 255.0f;
            return {g, g, g};
        }
        return ld3(d + 12 * idx);
    }
    return v3z();
}
__device__ float evalFloat(const uint8_t* t, float u, float v) {
    const uint8_t type = t[0];
    if (type == 6) return ld<float>(t + 4);
    if (type == 7) {
        const uint32_t w = ld<uint32_t>(t + 4), h = ld<uint32_t>(t + 8);
        return ld<float>(t + ld<int32_t>(t + 12) + 4 * ((size_t)texel(v, h) * w + texel(u, w)));
    }
    if (type == 8) return t[1] == 0 ? ld<float>(t + 4 + 4 * checkerIdx(u, v)) : 0.0f;
    return 0.0f;
}
__device__ float evalAlpha(const uint8_t* t, float u, float v) {
    const uint8_t type = t[0];
    if (type == 2 || type == 5) {
        const uint32_t w = ld<uint32_t>(t + 4), h = ld<uint32_t>(t + 8);
        const size_t idx = (size_t)texel(v, h) * w + texel(u, w);
        const uint8_t* d = t + ld<int32_t>(t + 12);
        return (float)(type == 2 ? d[4 * idx + 3] : d[idx]) / 255.0f;
    }
    if (type == 8) return evalFloat(t, u, v);
    return 1.0f;
}

// ---- distributions (distribution.cl:59-129) -------------------------------
__device__ __forceinline__ uint32_t sampleDiscrete(const uint8_t* d, float u, float& prob) {
    const float* cdf = reinterpret_cast<const float*>(d + ld<int32_t>(d + 12));
    uint32_t start = 0, end = ld<uint32_t>(d + 4) + 1, i = end / 2;
    while (true) {
        if (cdf[i] > u) end = i; else start = i;
        if (end - start == 1) break;
        i = (end - start) / 2 + start;
    }
    prob = reinterpret_cast<const float*>(d + ld<int32_t>(d + 8))[start];
    return start;
}
__device__ __forceinline__ float probDiscrete(const uint8_t* d, uint32_t k) {
    return reinterpret_cast<const float*>(d + ld<int32_t>(d + 8))[k];
}
// 1D piecewise constant; PDF/CDF offsets relative to `base`.
__device__ __forceinline__ float sampleCC1D(const uint8_t* d, const uint8_t* base, float u, float& pdf) {
    const float* cdf = reinterpret_cast<const float*>(base + ld<int32_t>(d + 24));
    uint32_t start = 0, end = ld<uint32_t>(d + 4) + 1, i = end / 2;
    while (true) {
        if (cdf[i] > u) end = i; else start = i;
        if (end - start == 1) break;
        i = (end - start) / 2 + start;
    }
    const float t = (u - cdf[start]) / (cdf[end] - cdf[start]);
    pdf = reinterpret_cast<const float*>(base + ld<int32_t>(d + 20))[start];
    return ld<float>(d + 8) + ((float)start + t) * ld<float>(d + 16);
}
// Clamped to numValues - 1: a point at the domain's upper end would index one
// past the end (distribution.cl:110-111).
__device__ __forceinline__ uint32_t stratumOf(const uint8_t* d, float p) {
    const float f = (p - ld<float>(d + 8)) / ld<float>(d + 16);
    return min(f > 0.0f ? (uint32_t)f : 0u, ld<uint32_t>(d + 4) - 1u);
}
__device__ __forceinline__ void sampleCC2D(const uint8_t* d, float u0, float u1, float& su, float& sv,
                                           float& pdf) {
    const uint8_t* parent = d + 8;
    float pdfTop;
    sv = sampleCC1D(parent, d, u1, pdfTop);
    const uint8_t* child = d + ld<int32_t>(d + 4) + 28 * (size_t)stratumOf(parent, sv);
    su = sampleCC1D(child, child, u0, pdf);
    pdf *= pdfTop;
}
__device__ __forceinline__ float pdfCC2D(const uint8_t* d, float u, float v) {
    const uint8_t* parent = d + 8;
    const float r = reinterpret_cast<const float*>(d + ld<int32_t>(parent + 20))[stratumOf(parent, v)];
    const uint8_t* child = d + ld<int32_t>(d + 4) + 28 * (size_t)stratumOf(parent, v);
    return r * reinterpret_cast<const float*>(child + ld<int32_t>(child + 20))[stratumOf(child, u)];
}

// ---- frames ----------------------------------------------------------------
__device__ __forceinline__ V3 makeTangent(V3 n) {
    if (fabsf(n.x) > fabsf(n.y)) {
        const float inv = 1.0f / sqrtf(n.x * n.x + n.z * n.z);
        return {-n.z * inv, 0.0f, n.x * inv};
    }
    const float inv = 1.0f / sqrtf(n.y * n.y + n.z * n.z);
    return {0.0f, n.z * inv, -n.y * inv};
}
__device__ __forceinline__ V3 toLocal(V3 s, V3 t, V3 n, V3 v) { return {dot(s, v), dot(t, v), dot(n, v)}; }
__device__ __forceinline__ V3 toWorld(V3 s, V3 t, V3 n, V3 v) {
    return {s.x * v.x + t.x * v.y + n.x * v.z, s.y * v.x + t.y * v.y + n.y * v.z,
            s.z * v.x + t.z * v.y + n.z * v.z};
}

// Surface / light point (global.cl:58-81).
struct SurfPt {
    V3 p, gN, sN, sT, uDir;
    float u, v;
    uint32_t face;
    bool hasTangent;
};

// calcHitpointParameters (bvh_traversal.cl:110-158) from the accepted
// triangle's barycentrics.
__device__ void hitParams(const RScene& S, V3 org, V3 dir, float t, uint32_t face, float b1i, float b2i,
                          SurfPt& sp) {
    const clr_face& f = S.g.faces[face];
    const V3 p0 = loadV3(S.g.verts, f.p0), p1 = loadV3(S.g.verts, f.p1), p2 = loadV3(S.g.verts, f.p2);
    sp.face = face;
    sp.p = org + dir * t;
    sp.gN = normalize3(cross(p1 - p0, p2 - p0));
    const float b0 = 1.0f - b1i - b2i;
    const float b1 = b1i;
    const float b2 = 1.0f - b0 - b1;
    const bool hasUV = f.uv0 != kNone;
    sp.u = sp.v = 0.0f;
    float2 uv0{}, uv1{}, uv2{};
    if (hasUV) {
        uv0 = reinterpret_cast<const float2*>(S.uvs)[f.uv0];
        uv1 = reinterpret_cast<const float2*>(S.uvs)[f.uv1];
        uv2 = reinterpret_cast<const float2*>(S.uvs)[f.uv2];
        sp.u = b0 * uv0.x + b1 * uv1.x + b2i * uv2.x;
        sp.v = b0 * uv0.y + b1 * uv1.y + b2i * uv2.y;
    }
    const bool hasVN = f.vn0 != kNone;
    sp.sN = hasVN ? normalize3(b0 * loadV3(S.normals, f.vn0) + b1 * loadV3(S.normals, f.vn1) +
                               b2 * loadV3(S.normals, f.vn2))
                  : sp.gN;
    sp.hasTangent = f.vt0 != kNone;
    if (sp.hasTangent)
        sp.sT = normalize3(b0 * loadV3(S.tangents, f.vt0) + b1 * loadV3(S.tangents, f.vt1) +
                           b2 * loadV3(S.tangents, f.vt2));
    if (hasUV) {
        const float du0 = uv0.x - uv2.x, dv0 = uv0.y - uv2.y, du1 = uv1.x - uv2.x, dv1 = uv1.y - uv2.y;
        const V3 dP0 = p0 - p2, dP1 = p1 - p2;
        const float invDet = 1.0f / (du0 * dv1 - dv0 * du1);
        const V3 uD = invDet * v3(dv1 * dP0.x - dv0 * dP1.x, dv1 * dP0.y - dv0 * dP1.y, dv1 * dP0.z - dv0 * dP1.z);
        sp.uDir = hasVN ? normalize3(cross(cross(sp.sN, uD), sp.sN)) : normalize3(uD);
    }
}

// Alpha rejection through the face's alpha texture (bvh_traversal.cl:94-105).
struct TexAlpha {
    const RScene* S;
    __device__ __forceinline__ bool reject(uint32_t face, float b1, float b2) const {
        const clr_face& f = S->g.faces[face];
        if (f.uv0 == kNone || f.alphaTexPtr == kNone) return false;
        const float b0 = 1.0f - b1 - b2;
        const float2 a = reinterpret_cast<const float2*>(S->uvs)[f.uv0];
        const float2 b = reinterpret_cast<const float2*>(S->uvs)[f.uv1];
        const float2 c = reinterpret_cast<const float2*>(S->uvs)[f.uv2];
        const float u = b0 * a.x + b1 * b.x + b2 * c.x;
        const float v = b0 * a.y + b1 * b.y + b2 * c.y;
        return evalAlpha(S->tex + f.alphaTexPtr, u, v) == 0.0f;
    }
};

// Traversal a render kernel is compiled for: the reference's walk over the
// LBVH, the ordered traversal (over either structure's traversal layout), or
// the reference's walk over the CPU midpoint BVH.
enum : int { kWalkLbvh = 0, kWalkOrdered = 1, kWalkMidpoint = 2 };

// rayIntersection with hit parameters, by WALK. Returns false (face = kNone)
// on a miss.
template <int WALK, class Stack>
__device__ __forceinline__ bool intersect(const RScene& S, V3 org, V3 dir, const Stack& stack, SurfPt& sp) {
    float t = INFINITY, b1 = 0.0f, b2 = 0.0f;
    TraceStats st{0u, 0u};
    bool over = false;
    uint32_t face;
    if constexpr (WALK == kWalkOrdered)  // over the 4-wide records (every render builds them)
        face = wideHit<false, false, TexAlpha, Stack, kStackDepth>(S.g, org, dir, t, stack, st, &b1, &b2,
                                                                   TexAlpha{&S}, &over);
    else if constexpr (WALK == kWalkMidpoint)
        face = bvhWalk<false, false>(S.g, org, dir, t, stack, st, &b1, &b2, TexAlpha{&S}, &over);
    else
        face = lbvhClosest<false, TexAlpha>(S.g, org, dir, t, stack, st, &b1, &b2, TexAlpha{&S}, &over);
    if (over) atomicOr(S.status, kStatusStackOverflow);
    if (face == kNone || isinf(t)) {
        sp.face = kNone;
        return false;
    }
    hitParams(S, org, dir, t, face, b1, b2, sp);
    return true;
}

// ---- MIS weights (pathtracer.cl:106-128,154-163) ----------------------------
// Power heuristic of a light sample, or 1 when light samples are used alone
// (CLR_RENDER_LIGHT_SAMPLING_ONLY).
__device__ __forceinline__ float lightSampleMis(uint32_t flags, float lightPDF, float fsPDF) {
    if (flags & CLR_RENDER_LIGHT_SAMPLING_ONLY) return 1.0f;
    return (lightPDF * lightPDF) / (lightPDF * lightPDF + fsPDF * fsPDF);
}
// Weight of emission reached by a diffuse or glossy BSDF sample: the power
// heuristic, 0 with light samples alone, 1 with BSDF samples alone.
__device__ __forceinline__ float emissionHitMis(uint32_t flags, float dirPDF, float lightPDF) {
    if (flags & CLR_RENDER_LIGHT_SAMPLING_ONLY) return 0.0f;
    if (flags & CLR_RENDER_BSDF_SAMPLING_ONLY) return 1.0f;
    return (dirPDF * dirPDF) / (lightPDF * lightPDF + dirPDF * dirPDF);
}

// ---- BSDF (reflection.cl) --------------------------------------------------
enum : uint32_t { kRefl = 1, kTrans = 2, kDiff = 4, kGlossy = 8, kSpec = 16, kNonSingular = 12, kAll = 31 };

// One decoded BSDF element. Its type flags follow from the element id
// (bxType) and its Fresnel record is an offset into the scene's other
// resources, so an element takes 10 registers.
struct BxDF {
    uint32_t id;
    V3 R;   // R / T / Rs (Ashikhmin specular, Ashikhmin diffuse)
    V3 R2;  // Rd (Ashikhmin diffuse)
    float a, b;
    uint32_t fres;  // Fresnel record: oth + fres (ids 1, 2, 6)
};
struct BSDF {
    int n;
    BxDF c[4];
    V3 N, S, T, NG;
    const uint8_t* oth;  // the scene's other resources (Fresnel records)
};

// c ? a : b field by field (a struct-wide select becomes a copy through a
// selected address, which keeps the element array out of registers).
__device__ __forceinline__ BxDF bxSelect(bool c, const BxDF& a, const BxDF& b) {
    BxDF r;
    r.id = c ? a.id : b.id;
    r.R = v3(c ? a.R.x : b.R.x, c ? a.R.y : b.R.y, c ? a.R.z : b.R.z);
    r.R2 = v3(c ? a.R2.x : b.R2.x, c ? a.R2.y : b.R2.y, c ? a.R2.z : b.R2.z);
    r.a = c ? a.a : b.a;
    r.b = c ? a.b : b.b;
    r.fres = c ? a.fres : b.fres;
    return r;
}

// Type flags of element ids 0..6 (the fxType set with each id,
// reflection.cl:241-329): diffuse (0) and Ashikhmin diffuse (5) are diffuse
// reflection, new Ward (3), Ashikhmin specular (4) and the Beckmann coat (6)
// glossy reflection, 1 specular reflection, 2 specular transmission.
__device__ __forceinline__ uint32_t bxType(uint32_t id) {
    constexpr uint64_t kTable = (uint64_t)(kRefl | kDiff) | (uint64_t)(kRefl | kSpec) << 5 |
                                (uint64_t)(kTrans | kSpec) << 10 | (uint64_t)(kRefl | kGlossy) << 15 |
                                (uint64_t)(kRefl | kGlossy) << 20 | (uint64_t)(kRefl | kDiff) << 25 |
                                (uint64_t)(kRefl | kGlossy) << 30;
    return id < 7u ? (uint32_t)(kTable >> (5u * id)) & 31u : 0u;
}

// BSDF element storage. REG keeps the four elements in registers: element
// i is read and written by field-wise selects over constant indices (an access
// at a variable index would put the array in scratch memory), which suits a
// kernel that shades and ends (the wavefront shade pass). The default indexes
// the array directly, so the compiler keeps it in scratch memory: a
// megakernel that also traverses keeps its registers for the walk.
template <bool REG>
__device__ __forceinline__ BxDF bxGet(const BSDF& B, int i);
template <bool REG>
__device__ __forceinline__ void bxPut(BSDF& B, int i, const BxDF& f) {
    if constexpr (REG) {
#pragma unroll
        for (int k = 0; k < 4; ++k) B.c[k] = bxSelect(k == i, f, B.c[k]);
    } else {
        B.c[i] = f;
    }
}

template <bool REG = false>
__device__ void bsdfAlloc(const RScene& S, uint32_t off, const SurfPt& sp, BSDF& B) {
    const uint8_t* m = S.mat + off;
    B.n = m[0];
    B.N = sp.sN;
    B.S = sp.hasTangent ? sp.sT : makeTangent(sp.sN);
    if (m[1]) {  // normal map (reflection.cl:201-223)
        const V3 tx = evalColor(S.tex + ld<uint32_t>(m + 4), sp.u, sp.v);
        const V3 nb = normalize3(v3(2.0f * tx.x - 1.0f, 2.0f * tx.y - 1.0f, 2.0f * tx.z - 1.0f));
        const V3 vD = cross(sp.uDir, sp.sN);
        B.N = v3(dot(v3(sp.uDir.x, vD.x, sp.sN.x), nb), dot(v3(sp.uDir.y, vD.y, sp.sN.y), nb),
                 dot(v3(sp.uDir.z, vD.z, sp.sN.z), nb));
        V3 ax = cross(sp.sN, B.N);
        const float sinTH = fminf(length3(ax), 1.0f);
        if (sinTH > 0.0001f) {
            ax = normalize3(ax);
            const float cosTH = cosf(asinf(sinTH));
            const float om = 1 - cosTH;
            const V3 s0 = B.S;
            B.S = v3(dot(v3(ax.x * ax.x * om + cosTH, ax.x * ax.y * om - ax.z * sinTH, ax.z * ax.x * om + ax.y * sinTH), s0),
                     dot(v3(ax.x * ax.y * om + ax.z * sinTH, ax.y * ax.y * om + cosTH, ax.y * ax.z * om - ax.x * sinTH), s0),
                     dot(v3(ax.z * ax.x * om - ax.y * sinTH, ax.y * ax.z * om + ax.x * sinTH, ax.z * ax.z * om + cosTH), s0));
        }
    }
    B.T = cross(B.N, B.S);
    B.NG = sp.gN;
    B.oth = S.oth;
    uint32_t p = off + 8;
    for (int i = 0; i < 4; ++i) {
        if (i >= B.n) break;
        p = (p + 3u) & ~3u;
        const uint8_t* e = S.mat + p;
        BxDF f;
        f.id = e[0];
        f.R = v3z();
        f.R2 = v3z();
        f.a = f.b = 0.0f;
        f.fres = 0u;
        switch (e[0]) {
            case 0: {
                f.R = evalColor(S.tex + ld<uint32_t>(e + 4), sp.u, sp.v);
                float s2 = evalFloat(S.tex + ld<uint32_t>(e + 8), sp.u, sp.v);
                s2 *= s2;
                f.a = 1.0f - (s2 / (2.0f * (s2 + 0.33f)));
                f.b = 0.45f * s2 / (s2 + 0.09f);
                p += 12;
                break;
            }
            case 1:
                f.R = evalColor(S.tex + ld<uint32_t>(e + 4), sp.u, sp.v);
                f.fres = ld<uint32_t>(e + 8);
                p += 12;
                break;
            case 2:
                f.R = evalColor(S.tex + ld<uint32_t>(e + 4), sp.u, sp.v);
                f.a = ld<float>(e + 8);
                f.b = ld<float>(e + 12);
                f.fres = ld<uint32_t>(e + 16);
                p += 20;
                break;
            case 3:
            case 4:
                f.R = evalColor(S.tex + ld<uint32_t>(e + 4), sp.u, sp.v);
                f.a = evalFloat(S.tex + ld<uint32_t>(e + 8), sp.u, sp.v);
                f.b = evalFloat(S.tex + ld<uint32_t>(e + 12), sp.u, sp.v);
                p += 16;
                break;
            case 5:
                f.R = evalColor(S.tex + ld<uint32_t>(e + 4), sp.u, sp.v);   // Rs
                f.R2 = evalColor(S.tex + ld<uint32_t>(e + 8), sp.u, sp.v);  // Rd
                p += 12;
                break;
            case 6:
                f.R = evalColor(S.tex + ld<uint32_t>(e + 4), sp.u, sp.v);
                f.a = evalFloat(S.tex + ld<uint32_t>(e + 8), sp.u, sp.v);
                f.fres = ld<uint32_t>(e + 12);
                p += 16;
                break;
            default:
                break;
        }
        bxPut<REG>(B, i, f);
    }
}

__device__ V3 fresnel(const uint8_t* fr, float cosi) {
    const uint8_t type = fr[0];
    if (type == 0) return v3s(1.0f);
    if (type == 1) {
        const V3 eta = ld3(fr + 16), k = ld3(fr + 32);
        cosi = fabsf(cosi);
        const V3 tmp = (eta * eta + k * k) * cosi * cosi;
        const V3 tec = (2.0f * eta) * cosi;
        const V3 one = v3s(1.0f);
        const V3 a = tmp - tec + one, b = tmp + tec + one;
        const V3 tf = eta * eta + k * k;
        const V3 c2 = v3s(cosi * cosi);
        const V3 c = tf - tec + c2, d = tf + tec + c2;
        const V3 Rparl2 = v3(a.x / b.x, a.y / b.y, a.z / b.z);
        const V3 Rperp2 = v3(c.x / d.x, c.y / d.y, c.z / d.z);
        return (Rparl2 + Rperp2) / 2.0f;
    }
    if (type == 2) {
        cosi = clampf(cosi, -1.0f, 1.0f);
        float ei = ld<float>(fr + 4), et = ld<float>(fr + 8);
        if (!(cosi > 0.0f)) {
            const float tmp = ei;
            ei = et;
            et = tmp;
        }
        const float sint = ei / et * sqrtf(fmaxf(0.0f, 1.0f - cosi * cosi));
        if (sint >= 1.0f) return v3s(1.0f);
        const float cost = sqrtf(fmaxf(0.0f, 1.0f - sint * sint));
        cosi = fabsf(cosi);
        const float Rparl = ((et * cosi) - (ei * cost)) / ((et * cosi) + (ei * cost));
        const float Rperp = ((ei * cosi) - (et * cost)) / ((ei * cosi) + (et * cost));
        return v3s((Rparl * Rparl + Rperp * Rperp) / 2.0f);
    }
    return v3z();
}

__device__ __forceinline__ float sinT2(V3 v) { return fmaxf(1.0f - v.z * v.z, 0.0f); }
__device__ __forceinline__ float sinTh(V3 v) { return sqrtf(sinT2(v)); }
__device__ __forceinline__ float cosPhi(V3 v) {
    const float s = sinTh(v);
    return s == 0.0f ? 1.0f : clampf(v.x / s, -1.0f, 1.0f);
}
__device__ __forceinline__ float sinPhi(V3 v) {
    const float s = sinTh(v);
    return s == 0.0f ? 0.0f : clampf(v.y / s, -1.0f, 1.0f);
}

// Beckmann microfacet pieces (addition): D, Smith-Beckmann G1 (Walter 2007
// rational fit), half vector on the +z side.
__device__ __forceinline__ float beckD(V3 h, float a) {
    const float c2 = h.z * h.z;
    if (c2 <= 0.0f) return 0.0f;
    const float t2 = (1.0f - c2) / c2;
    return expf(-t2 / (a * a)) / (kPiF * a * a * c2 * c2);
}
__device__ __forceinline__ float beckG1(V3 v, float a) {
    const float c2 = v.z * v.z;
    if (c2 >= 1.0f) return 1.0f;
    const float tanT = sqrtf(fmaxf(0.0f, 1.0f - c2)) / fabsf(v.z);
    const float x = 1.0f / (a * tanT);
    if (x >= 1.6f) return 1.0f;
    return (3.535f * x + 2.181f * x * x) / (1.0f + 2.276f * x + 2.577f * x * x);
}
__device__ __forceinline__ V3 halfPlusZ(V3 wo, V3 wi) {
    const V3 h = normalize3(wo + wi);
    return h.z < 0.0f ? -h : h;
}

__device__ V3 fx(const BxDF f, const uint8_t* oth, V3 wo, V3 wi) {
    if (wi.z * wo.z <= 0.0f) return v3z();
    switch (f.id) {
        case 0: {
            if (f.a == 1.0f) return f.R / kPiF;
            const float sti = sinTh(wi), sto = sinTh(wo);
            float maxCos = 0.0f;
            if (sti > 1e-4f && sto > 1e-4f) maxCos = fmaxf(0.0f, cosPhi(wi) * cosPhi(wo) + sinPhi(wi) * sinPhi(wo));
            float sinA, tanB;
            if (fabsf(wi.z) > fabsf(wo.z)) { sinA = sto; tanB = sti / fabsf(wi.z); }
            else { sinA = sti; tanB = sto / fabsf(wo.z); }
            return f.R / kPiF * (f.a + f.b * maxCos * sinA * tanB);
        }
        case 3: {
            const V3 h = normalize3(wo + wi);
            const float hx = h.x / f.a, hy = h.y / f.b;
            const float dHN = fabsf(h.z), dHI = dot(h, wi);
            return f.R * expf(-(hx * hx + hy * hy) / (dHN * dHN)) /
                   (4 * kPiF * f.a * f.b * dHI * dHI * dHN * dHN * dHN * dHN);
        }
        case 4: {
            const V3 h = normalize3(wo + wi);
            const float dHV = dot(h, wo);
            const float e = (f.a * h.x * h.x + f.b * h.y * h.y) / (1 - h.z * h.z);
            const V3 fr = f.R + (v3s(1.0f) - f.R) * powf(1.0f - dHV, 5.0f);
            return sqrtf((f.a + 1) * (f.b + 1)) / (8 * kPiF) * powf(fabsf(h.z), e) /
                   (dHV * fmaxf(fabsf(wo.z), fabsf(wi.z))) * fr;
        }
        case 5: {
            const V3 r = 28 * f.R2 / (23 * kPiF) * (v3s(1.0f) - f.R);
            return r * (1.0f - powf(1.0f - fabsf(wo.z) / 2, 5.0f)) * (1.0f - powf(1.0f - fabsf(wi.z) / 2, 5.0f));
        }
        case 6: {
            const V3 h = halfPlusZ(wo, wi);
            const float D = beckD(h, f.a);
            const float G = beckG1(wo, f.a) * beckG1(wi, f.a);
            const V3 F = fresnel(oth + f.fres, dot(wo, h));
            const float k = D * G / (4.0f * fabsf(wo.z) * fabsf(wi.z));
            return (f.R * F) * k;
        }
        default:
            return v3z();
    }
}

__device__ float fxPdf(const BxDF& f, V3 wo, V3 wi) {
    switch (f.id) {
        case 0:
        case 5:
            return fabsf(wi.z) / kPiF;
        case 3: {
            if (wi.z * wo.z <= 0) return 0.0f;
            const V3 h = normalize3(wo + wi);
            const float hx = h.x / f.a, hy = h.y / f.b;
            const float dHN = fabsf(h.z), dHI = dot(h, wi);
            const float num = expf(-(hx * hx + hy * hy) / (dHN * dHN));
            return num / (4 * kPiF * f.a * f.b * dHI * dHN * dHN * dHN);
        }
        case 4: {
            if (wo.z * wi.z <= 0) return 0.0f;
            const V3 h = normalize3(wo + wi);
            const float e = (f.a * h.x * h.x + f.b * h.y * h.y) / (1 - h.z * h.z);
            return sqrtf((f.a + 1) * (f.b + 1)) / (2 * kPiF) * powf(fabsf(h.z), e) / (4 * dot(wo, h));
        }
        case 6: {
            if (wo.z * wi.z <= 0) return 0.0f;
            const V3 h = halfPlusZ(wo, wi);
            return beckD(h, f.a) * fabsf(h.z) / (4.0f * fabsf(dot(wo, h)));
        }
        default:
            return 0.0f;
    }
}

__device__ V3 sampleFx(const BxDF& f, const uint8_t* oth, V3 wo, float u0, float u1, V3& wi, float& pdf) {
    switch (f.id) {
        case 0:
        case 5:
            wi = cosineHemisphere(u0, u1);
            if (wo.z < 0.0f) wi.z *= -1;
            pdf = fabsf(wi.z) / kPiF;
            return fx(f, oth, wo, wi);
        case 1: {
            wi = v3(-wo.x, -wo.y, wo.z);
            pdf = 1.0f;
            const V3 r = f.R * fresnel(oth + f.fres, wo.z);
            return r / fabsf(wi.z);
        }
        case 2: {
            const bool entering = wo.z > 0.0f;
            float ei = f.a, et = f.b;
            if (!entering) {
                const float tmp = ei;
                ei = et;
                et = tmp;
            }
            const float eta = ei / et;
            const float sint2 = eta * eta * sinT2(wo);
            if (sint2 >= 1.0f) return v3z();
            float cost = sqrtf(fmaxf(0.0f, 1.0f - sint2));
            if (entering) cost = -cost;
            wi = v3(eta * -wo.x, eta * -wo.y, cost);
            pdf = 1.0f;
            const V3 r = (v3s(1.0f) - fresnel(oth + f.fres, wo.z)) * f.R;
            return r / fabsf(wi.z);
        }
        case 3: {
            const float quad = 2 * kPiF * u1;
            const float phi = atan2f(f.b * sinf(quad), f.a * cosf(quad));
            const float cpa = cosf(phi) / f.a, spa = sinf(phi) / f.b;
            float th = atanf(sqrtf(-logf(1 - u0) / (cpa * cpa + spa * spa)));
            if (wo.z < 0) th = kPiF - th;
            const V3 h = v3(sinf(th) * cosf(phi), sinf(th) * sinf(phi), cosf(th));
            wi = 2 * dot(wo, h) * h - wo;
            if (wi.z * wo.z <= 0) {
                pdf = 0.0f;
                return v3z();
            }
            const float hx = h.x / f.a, hy = h.y / f.b;
            const float dHN = fabsf(h.z), dHI = dot(h, wi);
            const float num = expf(-(hx * hx + hy * hy) / (dHN * dHN));
            const float den = 4 * kPiF * f.a * f.b * dHI * dHN * dHN * dHN;
            pdf = num / den;
            return f.R * (num / (den * dHI * dHN));
        }
        case 4: {
            const float quad = 2 * kPiF * u1;
            const float phi = atan2f(sqrtf(f.a + 1) * sinf(quad), sqrtf(f.b + 1) * cosf(quad));
            const float cp = cosf(phi), sp = sinf(phi);
            float th = acosf(powf(1 - u0, 1.0f / (f.a * cp * cp + f.b * sp * sp + 1)));
            if (wo.z < 0) th = kPiF - th;
            const V3 h = v3(sinf(th) * cosf(phi), sinf(th) * sinf(phi), cosf(th));
            wi = 2 * dot(wo, h) * h - wo;
            if (wo.z * wi.z <= 0) {
                pdf = 0.0f;
                return v3z();
            }
            pdf = fxPdf(f, wo, wi);
            return fx(f, oth, wo, wi);
        }
        case 6: {
            const float a = f.a;
            const float t2 = -a * a * logf(1.0f - u0);
            const float phi = 2.0f * kPiF * u1;
            const float cth = 1.0f / sqrtf(1.0f + t2);
            const float sth = sqrtf(fmaxf(0.0f, 1.0f - cth * cth));
            V3 h = v3(sth * cosf(phi), sth * sinf(phi), cth);
            if (wo.z < 0.0f) h = -h;
            wi = 2 * dot(wo, h) * h - wo;
            if (wi.z * wo.z <= 0) {
                pdf = 0.0f;
                return v3z();
            }
            pdf = fxPdf(f, wo, wi);
            return fx(f, oth, wo, wi);
        }
        default:
            pdf = 0.0f;
            return v3z();
    }
}

__device__ __forceinline__ bool matchType(uint32_t type, uint32_t mask) { return (type & mask) == type; }

// Element i of a BSDF (see bxPut). Loops over the elements stay rolled, so
// each inlines the element code once.
template <bool REG>
__device__ __forceinline__ BxDF bxGet(const BSDF& B, int i) {
    if constexpr (REG) {
        BxDF f = B.c[0];
#pragma unroll
        for (int k = 1; k < 4; ++k) f = bxSelect(k == i, B.c[k], f);
        return f;
    } else {
        return B.c[i];
    }
}
__device__ __forceinline__ bool hasNonSpecular(const BSDF& B) {
    bool r = false;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < B.n) r |= (bxType(B.c[i].id) & kNonSingular) != 0;
    return r;
}

template <bool REG = false>
__device__ V3 sampleFs(const BSDF& B, V3 vout, float uc, float u0, float u1, V3& vin, float& pdf,
                       uint32_t& sampledType) {
    const V3 wo = toLocal(B.S, B.T, B.N, vout);
    const uint32_t which = randUInt(uc, (uint32_t)B.n);
    const BxDF f = bxGet<REG>(B, (int)which);
    V3 wi;
    V3 ret = sampleFx(f, B.oth, wo, u0, u1, wi, pdf);
    if (pdf == 0.0f) {
        sampledType = 0;
        return v3z();
    }
    const uint32_t fType = bxType(f.id);
    sampledType = fType;
    vin = toWorld(B.S, B.T, B.N, wi);
    if (!(fType & kSpec) && B.n > 1)
        for (int i = 0; i < B.n; ++i) {
            const BxDF e = bxGet<REG>(B, i);
            if ((uint32_t)i != which && matchType(bxType(e.id), kAll)) pdf += fxPdf(e, wo, wi);
        }
    pdf /= (float)B.n;
    if (!(fType & kSpec)) {
        ret = v3z();
        const uint32_t flags = dot(vin, B.NG) * dot(vout, B.NG) > 0 ? (kAll & ~kTrans) : (kAll & ~kRefl);
        for (int i = 0; i < B.n; ++i) {
            const BxDF e = bxGet<REG>(B, i);
            if (matchType(bxType(e.id), flags)) ret += fx(e, B.oth, wo, wi);
        }
    }
    return ret;
}
template <bool REG = false>
__device__ V3 evalFs(const BSDF& B, V3 vout, V3 vin) {
    const V3 wo = toLocal(B.S, B.T, B.N, vout), wi = toLocal(B.S, B.T, B.N, vin);
    const uint32_t flags = dot(vin, B.NG) * dot(vout, B.NG) > 0 ? (kAll & ~kTrans) : (kAll & ~kRefl);
    V3 r = v3z();
    for (int i = 0; i < B.n; ++i) {
        const BxDF e = bxGet<REG>(B, i);
        if (matchType(bxType(e.id), flags)) r += fx(e, B.oth, wo, wi);
    }
    return r;
}
template <bool REG = false>
__device__ float evalFsPdf(const BSDF& B, V3 vout, V3 vin) {
    const V3 wo = toLocal(B.S, B.T, B.N, vout), wi = toLocal(B.S, B.T, B.N, vin);
    float p = 0.0f;
    for (int i = 0; i < B.n; ++i) {
        const BxDF e = bxGet<REG>(B, i);
        if (matchType(bxType(e.id), kAll)) p += fxPdf(e, wo, wi);
    }
    return p / (float)B.n;
}

// ---- emitters (light.cl) ---------------------------------------------------
struct EDF {
    bool env;
    int n;
    V3 L;  // sum over elements (each element's Le is direction independent
           // apart from the one-sided cosine test of area emitters)
    V3 N, S, T, NG;
};
__device__ void edfArea(const RScene& Sc, uint32_t off, const SurfPt& lp, float u, float v, EDF& E) {
    const uint8_t* info = Sc.mat + off;
    E.env = false;
    E.n = info[0];
    E.N = lp.sN;
    E.S = lp.hasTangent ? lp.sT : makeTangent(lp.sN);
    E.T = cross(E.N, E.S);
    E.NG = lp.gN;
    E.L = v3z();
    uint32_t p = off + 1;
    for (int i = 0; i < E.n; ++i) {
        p = (p + 3u) & ~3u;
        const uint8_t* e = Sc.mat + p;
        if (e[0] != 0) break;
        E.L += evalColor(Sc.tex + ld<uint32_t>(e + 4), u, v);
        p += 8;
    }
}
__device__ void edfEnv(const RScene& Sc, float u, float v, EDF& E) {
    E.env = true;
    E.n = Sc.envLP[0];
    E.L = v3z();
    uint32_t p = (uint32_t)(Sc.envLP - Sc.mat) + 1;
    for (int i = 0; i < E.n; ++i) {
        p = (p + 3u) & ~3u;
        const uint8_t* e = Sc.mat + p;
        if (e[0] != 0) break;
        E.L += evalColor(Sc.tex + ld<uint32_t>(e + 4), u, v) * ld<float>(e + 8);
        p += 16;
    }
}
// Le: diffuse emission is M / pi on the front side (light.cl:382-393), the
// environment returns its radiance. Summation order per element matches the
// reference when a property has a single element (every scene here).
__device__ __forceinline__ V3 evalLe(const EDF& E, V3 vout) {
    if (E.env) return E.L;
    const float z = dot(E.N, vout);
    return z > 0.0f ? E.L / kPiF : v3z();
}
__device__ __forceinline__ float absCosEDF(const EDF& E, V3 v) { return E.env ? 1.0f : fabsf(dot(E.N, v)); }

// Light sample (light.cl:138-245). Returns the light point in `lp` (face or
// kNone for the environment) and multiplies areaPDF in.
__device__ void sampleLight(const RScene& Sc, float ul, float u0, float u1, SurfPt& lp, bool& atInf, EDF& E,
                            float& areaPDF) {
    const uint32_t li = sampleDiscrete(Sc.lightDist, ul, areaPDF);
    const clr_light_info L = Sc.lights[li];
    if (L.atInfinity) {
        atInf = true;
        lp.face = kNone;
        float su = 0.0f, sv = 0.0f;
        const uint32_t num = Sc.envLP[0];
        uint32_t p = (uint32_t)(Sc.envLP - Sc.mat) + 1;
        for (uint32_t i = 0; i < num; ++i) {
            p = (p + 3u) & ~3u;
            const uint8_t* e = Sc.mat + p;
            if (e[0] != 0) break;
            if (i == 0) {
                float uvPDF;
                sampleCC2D(Sc.oth + ld<uint32_t>(e + 12), u0, u1, su, sv, uvPDF);
                const float theta = sv * kPiF, phi = su * 2 * kPiF;
                const float st = sinf(theta);
                lp.p = v3(-sinf(phi) * st, cosf(theta), cosf(phi) * st);
                areaPDF *= uvPDF / (2 * kPiF * kPiF * st);
                break;
            }
            p += 16;
        }
        lp.u = su;
        lp.v = sv;
        edfEnv(Sc, su, sv, E);
        return;
    }
    atInf = false;
    lp.face = L.reference;
    const clr_face& f = Sc.g.faces[lp.face];
    const V3 p0 = loadV3(Sc.g.verts, f.p0), p1 = loadV3(Sc.g.verts, f.p1), p2 = loadV3(Sc.g.verts, f.p2);
    const float su = sqrtf(u0);
    const float b0 = 1.0f - su, b1 = u1 * su, b2 = 1.0f - b0 - b1;
    const V3 ng = cross(p1 - p0, p2 - p0);
    const float area = 0.5f * length3(ng);
    areaPDF *= 1.0f / area;
    lp.p = b0 * p0 + b1 * p1 + b2 * p2;
    lp.gN = normalize3(ng);
    const bool hasVN = f.vn0 != kNone && f.vn1 != kNone && f.vn2 != kNone;
    lp.sN = hasVN ? normalize3(b0 * loadV3(Sc.normals, f.vn0) + b1 * loadV3(Sc.normals, f.vn1) +
                               b2 * loadV3(Sc.normals, f.vn2))
                  : lp.gN;
    lp.hasTangent = f.vt0 != kNone && f.vt1 != kNone && f.vt2 != kNone;
    if (lp.hasTangent)
        lp.sT = normalize3(b0 * loadV3(Sc.tangents, f.vt0) + b1 * loadV3(Sc.tangents, f.vt1) +
                           b2 * loadV3(Sc.tangents, f.vt2));
    lp.u = lp.v = 0.0f;
    if (f.uv0 != kNone && f.uv1 != kNone && f.uv2 != kNone) {
        const float2 a = reinterpret_cast<const float2*>(Sc.uvs)[f.uv0];
        const float2 b = reinterpret_cast<const float2*>(Sc.uvs)[f.uv1];
        const float2 c = reinterpret_cast<const float2*>(Sc.uvs)[f.uv2];
        lp.u = b0 * a.x + b1 * b.x + b2 * c.x;
        lp.v = b0 * a.y + b1 * b.y + b2 * c.y;
        // uDir is not used by emitters.
    }
    edfArea(Sc, f.lightPtr, lp, lp.u, lp.v, E);
}

// getAreaPDF (light.cl:247-292).
__device__ float areaPdfOfFace(const RScene& Sc, uint32_t face) {
    const uint32_t li = Sc.faceLight[face];
    if (li == kNone) return 0.0f;
    const clr_face& f = Sc.g.faces[face];
    const V3 p0 = loadV3(Sc.g.verts, f.p0), p1 = loadV3(Sc.g.verts, f.p1), p2 = loadV3(Sc.g.verts, f.p2);
    const V3 ng = cross(p1 - p0, p2 - p0);
    return 1.0f / (0.5f * length3(ng)) * probDiscrete(Sc.lightDist, li);
}
__device__ float areaPdfEnv(const RScene& Sc, float u, float v) {
    float areaPDF = 0.0f;
    const uint32_t num = Sc.envLP[0];
    uint32_t p = (uint32_t)(Sc.envLP - Sc.mat) + 1;
    for (uint32_t i = 0; i < num; ++i) {
        p = (p + 3u) & ~3u;
        const uint8_t* e = Sc.mat + p;
        if (e[0] != 0) break;
        const float uvPDF = pdfCC2D(Sc.oth + ld<uint32_t>(e + 12), u, v);
        areaPDF += uvPDF / (2 * kPiF * kPiF * sinf(v * kPiF)) * 1.0f;
        p += 16;
    }
    if (Sc.envLight < 0) return 0.0f;
    return areaPDF * probDiscrete(Sc.lightDist, (uint32_t)Sc.envLight);
}

// ---- camera (camera.cl:60-123, matrix.cl) ----------------------------------
__device__ __forceinline__ float mt(const uint8_t* m, int i) { return ld<float>(m + 4 * i); }
__device__ __forceinline__ V3 mulVec(const uint8_t* m, V3 v) {
    return {mt(m, 0) * v.x + mt(m, 4) * v.y + mt(m, 8) * v.z, mt(m, 1) * v.x + mt(m, 5) * v.y + mt(m, 9) * v.z,
            mt(m, 2) * v.x + mt(m, 6) * v.y + mt(m, 10) * v.z};
}
__device__ __forceinline__ V3 mulPoint(const uint8_t* m, V3 p) {
    V3 r = {mt(m, 0) * p.x + mt(m, 4) * p.y + mt(m, 8) * p.z + mt(m, 12) * 1.0f,
            mt(m, 1) * p.x + mt(m, 5) * p.y + mt(m, 9) * p.z + mt(m, 13) * 1.0f,
            mt(m, 2) * p.x + mt(m, 6) * p.y + mt(m, 10) * p.z + mt(m, 14) * 1.0f};
    const float w = mt(m, 3) * p.x + mt(m, 7) * p.y + mt(m, 11) * p.z + mt(m, 15) * 1.0f;
    if (w != 1.0f) r = r / w;
    return r;
}
__device__ __forceinline__ V3 mulRaster(const uint8_t* m, float x, float y) {
    float px = mt(m, 0) * x + mt(m, 4) * y + mt(m, 8) * 0.0f + mt(m, 12) * 1.0f;
    float py = mt(m, 1) * x + mt(m, 5) * y + mt(m, 9) * 0.0f + mt(m, 13) * 1.0f;
    float pz = mt(m, 2) * x + mt(m, 6) * y + mt(m, 10) * 0.0f + mt(m, 14) * 1.0f;
    const float pw = mt(m, 3) * x + mt(m, 7) * y + mt(m, 11) * 0.0f + mt(m, 15) * 1.0f;
    if (pw != 1.0f) {
        px /= pw;
        py /= pw;
        pz /= pw;
    }
    return {px, py, pz};
}

// Camera ray for pixel (gx, gy): consumes 2 lens draws then 2 jitter draws.
// Returns the initial path weight |cos| / dirPDF (sample_We returns 1).
__device__ __forceinline__ void cameraRay(const RScene& Sc, uint32_t gx, uint32_t gy, Rng& rng, V3& org,
                                          V3& dir, float& weight, float& dirPDF) {
    const uint8_t* cam = Sc.camera;
    const uint8_t* l2w = cam + 64;
    const float lensR = ld<float>(cam + 136), objPD = ld<float>(cam + 140), vpa = ld<float>(cam + 132);
    const float c0 = rng.uniform(), c1 = rng.uniform();
    float ux, uy;
    concentricDisk(c0, c1, ux, uy);
    const V3 idfDir = mulVec(l2w, v3(0.0f, 0.0f, 1.0f));
    const V3 lensLocal = v3(ux * lensR, uy * lensR, 0.0f);
    org = mulPoint(l2w, lensLocal);
    const float px = (float)gx + rng.uniform();
    const float py = (float)gy + rng.uniform();
    V3 d = normalize3(mulRaster(cam + 192, px, py));
    if (lensR > 0.0f) {
        const float ft = -objPD / d.z;
        const V3 pf = v3z() + ft * d;
        d = normalize3(pf - v3(lensLocal.x, lensLocal.y, 0.0f));
    }
    dir = mulVec(l2w, d);
    dirPDF = 1.0f / (powf(-d.z, 3.0f) * vpa);
    weight = fabsf(dot(idfDir, dir)) / dirPDF;
}

// Primary ray from given samples: lens sample (c0, c1) and raster position
// (px, py) (sampleLensPos and sample_We, camera.cl:60-113).
__device__ __forceinline__ void cameraRayAt(const RScene& Sc, float c0, float c1, float px, float py, V3& org,
                                            V3& dir) {
    const uint8_t* cam = Sc.camera;
    const uint8_t* l2w = cam + 64;
    const float lensR = ld<float>(cam + 136), objPD = ld<float>(cam + 140);
    float ux, uy;
    concentricDisk(c0, c1, ux, uy);
    const V3 lensLocal = v3(ux * lensR, uy * lensR, 0.0f);
    org = mulPoint(l2w, lensLocal);
    V3 d = normalize3(mulRaster(cam + 192, px, py));
    if (lensR > 0.0f) {
        const float ft = -objPD / d.z;
        const V3 pf = v3z() + ft * d;
        d = normalize3(pf - v3(lensLocal.x, lensLocal.y, 0.0f));
    }
    dir = mulVec(l2w, d);
}

// getBaseColor_fs (reflection.cl:730-772): the sum over the BSDF's elements
// of R (diffuse, specular reflection, Ward), Rs (Ashikhmin specular) and Rd
// (Ashikhmin diffuse); transmission and other elements add nothing.
__device__ __forceinline__ V3 baseColor(const BSDF& B) {
    V3 c = v3z();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= B.n) break;
        const BxDF& f = B.c[i];
        if (f.id == 0 || f.id == 1 || f.id == 3 || f.id == 4) c = c + f.R;
        else if (f.id == 5) c = c + f.R2;
    }
    return c;
}

}  // namespace dev
}  // namespace clr
