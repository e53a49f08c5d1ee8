// CPU oracle for the path tracer — TEST INFRASTRUCTURE ONLY.
//
// Serial restatement of the reference's render kernel and its includes
// (OpenCL_PathTracer/OpenCL/pathtracer.cl, camera.cl, rng.cl, light.cl,
// distribution.cl, reflection.cl, texture.cl, bvh_traversal.cl, global.cl,
// matrix.cl; the CPU twin is Equivalent/sim_*.hpp), reading the same byte
// blobs as the reference (materialsData, texturesData, otherResources).
// Loaded by tests/ and bench.py's cpu_baseline leg only.
//
// Arithmetic: -ffp-contract=off; vector/scalar division divides each
// component (OpenCL semantics; the Equivalent multiplies by the reciprocal,
// sim_base.hpp:633-647); dot/cross left to right; host libm for sin, cos,
// acos, asin, atan2, atan, exp, log, pow, fmod.
//
// Documented deviations from the reference, each shared with the GPU path:
//  D1 face indices are 32-bit (bvh_traversal.cl:43,55 truncate to 16 bits);
//  D2 RNG state is per global pixel (pathtracer.cl:54 indexes per tile slot);
//  D3 an alpha-rejected triangle does not shrink t (bvh_traversal.cl:87);
//  D4 a shadow ray that misses has no face; an environment sample is visible
//     only when the shadow ray misses (pathtracer.cl:121-122 reads a stale or
//     uninitialised faceID);
//  D5 Ashikhmin-Shirley sampling mirrors theta_h by vout.z (reflection.cl:474
//     reads vin->z before it is written);
//  D6 the bounce cap is a parameter (pathtracer.cl:166 fixes ray.depth > 99);
//  D7 the environment marginal's PDF/CDF offsets are taken relative to the 2D
//     distribution header, where the host writes them
//     (CreateFunctions.cpp:325-326 vs distribution.cl:116,125);
//  D8 additions: FP32 RGB image texture (type 9) and a Beckmann microfacet
//     BxDF (element/BxDF id 6), pinned by closed forms in the tests;
//  D9 normal mapping normalises the rotation axis as the kernel does
//     (reflection.cl:210; the Equivalent omits it, sim_reflection.hpp:208-211);
//  D10 stratum indices of piecewise-constant distributions are clamped to
//     numValues - 1 (distribution.cl:110-111 indexes one past the end for a
//     point at the domain's upper end, as its own comment there notes);
//  D11 a light sample whose density is zero or infinite (an environment
//     sample exactly at a pole, u = 0, where sin(theta) = 0 makes
//     light.cl:169's pdf infinite and the MIS weight inf/inf) adds nothing.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/clear_amd.h"
#include "ordered_traversal.h"

namespace {

constexpr float kPi = 3.14159274101257324219f;  // M_PI_F
constexpr float kEps = 0.00001f;                // EPSILON (global.cl:28)
constexpr uint32_t kNoFace = 0xFFFFFFFFu;

// Planted mistakes, for tests only: orc_set_plant(k) makes the restatement
// commit mistake k, so that a test can show its pin detects that mistake.
// 0 (the default) plants nothing.
enum Plant : int {
    kPlantLightNoInvArea = 1,     // sampleLightPos: pdf not divided by the triangle area
    kPlantAreaPdfNoInvArea = 2,   // sampleLightPos and getAreaPDF both drop 1 / area
    kPlantNeeNoCosL = 3,          // light-sample weight without |cos| at the light
    kPlantMisNoCosL = 4,          // BSDF pdf taken to area measure without |cos| at the light
    kPlantNeeMisSwapped = 5,      // light-sample MIS weight with the BSDF pdf on top
    kPlantImplicitNoDist2 = 6,    // emission-hit light pdf taken to solid angle without dist^2
    kPlantMixNoAverage = 7,       // mixture pdf not divided by the element count
    kPlantMixOwnPdfOnly = 8,      // sample_fs returns the sampled element's pdf alone
    kPlantNoLightChoiceProb = 9,  // the probability of the chosen light dropped
    kPlantLeNoInvPi = 10,         // diffuse emission M instead of M / pi
    kPlantLeTwoSided = 11,        // diffuse emission on both sides
    kPlantBeckmannNo4 = 12,       // Beckmann f without the 4 of its denominator
    kPlantMixPickFirst = 13,      // sample_fs always samples element 0
};
int g_plant = 0;
inline bool planted(int k) { return g_plant == k; }

// Traversal work of renders with the ordered walk (its 4-wide form), for the
// render roofline (orc_work_counters): [0] rays traced, [1] record visits,
// [2] triangle tests, all as closest-hit walks; of those, environment shadow rays:
// [3] count, [4] visits, [5] tests, and [6] visits, [7] tests of the any-hit
// walk the wavefront renderer uses for them instead (counted on the side;
// the restatement's results do not change).
constexpr int kWorkWords = 8;
thread_local uint64_t t_work[kWorkWords];
std::atomic<uint64_t> g_work[kWorkWords];
std::atomic<bool> g_countWork{false};
void flushWork() {
    for (int k = 0; k < kWorkWords; ++k) {
        g_work[k] += t_work[k];
        t_work[k] = 0;
    }
}

struct V3 {
    float x, y, z;
};
inline V3 mk(float x, float y, float z) { return {x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
inline V3& operator*=(V3& a, V3 b) { a = a * b; return a; }
inline V3& operator*=(V3& a, float s) { a = a * s; return a; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline V3 normalize(V3 v) { return v / length(v); }
inline bool zeroVec(V3 v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; }
inline float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
inline float omax(float a, float b) { return (b > a || a != a) ? b : a; }
inline float omin(float a, float b) { return (b < a || a != a) ? b : a; }
inline float clampf(float v, float lo, float hi) { return omin(omax(v, lo), hi); }

template <typename T>
inline T rd(const uint8_t* p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}
inline V3 rd3(const uint8_t* p) { return {rd<float>(p), rd<float>(p + 4), rd<float>(p + 8)}; }

struct SceneC {
    const clr_float3* vertices;
    const clr_float3* normals;
    const clr_float3* tangents;
    const float* uvs;
    const clr_face* faces;
    const clr_light_info* lights;
    const uint8_t* materials;
    const uint8_t* textures;
    const uint8_t* others;
    const clr_lbvh_inode* inodes;
    const clr_lbvh_leaf* leaves;
    uint32_t numFaces;
    const uint8_t* camera;
    const uint8_t* environment;
    const uint8_t* lightDist;
    bool ordered;  // ordered traversal (ordered_traversal.h) instead of the reference's walk
    // CPU midpoint BVH: the reference's walk uses it instead of the LBVH
    // (the ordered traversal's hits do not depend on the structure).
    const clr_bvh_node* bvh;
};

inline V3 vtx(const clr_float3* a, uint32_t i) { return {a[i].x, a[i].y, a[i].z}; }

// ---- rng.cl:21-83 ----
inline uint32_t getUInt(uint32_t* s) {
    uint32_t t = s[0] ^ (s[0] << 11);
    s[0] = s[1];
    s[1] = s[2];
    s[2] = s[3];
    return s[3] = (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8));
}
inline float getFloat(uint32_t* s) {
    uint32_t b = (getUInt(s) >> 9) | 0x3f800000u;
    float f;
    std::memcpy(&f, &b, 4);
    return f - 1.0f;
}
inline uint32_t randUInt(float u, uint32_t maxv) {
    float f = (float)maxv * u;
    uint32_t v = f > 0.0f ? (uint32_t)f : 0u;
    return std::min(v, maxv - 1);
}
void concentricSampleDisk(float u1, float u2, float* dx, float* dy) {
    float r, theta;
    float sx = 2 * u1 - 1;
    float sy = 2 * u2 - 1;
    if (sx == 0.0f && sy == 0.0f) { *dx = 0.0f; *dy = 0.0f; return; }
    if (sx >= -sy) {
        if (sx > sy) { r = sx; theta = sy / sx; }
        else { r = sy; theta = 2.0f - sx / sy; }
    } else {
        if (sx > sy) { r = -sy; theta = 6.0f + sx / sy; }
        else { r = -sx; theta = 4.0f + sy / sx; }
    }
    theta *= kPi / 4.0f;
    *dx = r * std::cos(theta);
    *dy = r * std::sin(theta);
}
V3 cosineSampleHemisphere(float u1, float u2) {
    float x, y;
    concentricSampleDisk(u1, u2, &x, &y);
    return {x, y, std::sqrt(std::max(0.0f, 1.0f - x * x - y * y))};
}

// ---- textures (texture.cl:58-204) ----
inline float halfToFloat(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16, exp = (h >> 10) & 0x1Fu, man = h & 0x3FFu, bits;
    if (exp == 0) {
        if (man == 0) bits = sign;
        else {
            int e = -1;
            do { ++e; man <<= 1; } while (!(man & 0x400u));
            bits = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3FFu) << 13);
        }
    } else if (exp == 31) bits = sign | 0x7F800000u | (man << 13);
    else bits = sign | ((exp + 112) << 23) | (man << 13);
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}
inline uint32_t texel(float coord, uint32_t size) {
    float f = std::fmod((float)size * coord, (float)size);
    uint32_t v = f > 0.0f ? (uint32_t)f : 0u;
    return std::min(v, size - 1);
}
V3 proceduralColor(const uint8_t* t, float u, float v) {
    switch (t[1]) {
        case 0: {  // checkerboard
            uint32_t i = ((u * 2 > 0 ? (uint32_t)(u * 2) : 0u) + (v * 2 > 0 ? (uint32_t)(v * 2) : 0u)) % 2;
            return rd3(t + 16 + 16 * i);
        }
        case 1: {  // checkerboard bump
            float halfWidth = rd<float>(t + 4) * 0.5f;
            float uComp = 0.0f, vComp = 0.0f;
            float au = std::fmod(std::fabs(u), 1.0f), av = std::fmod(std::fabs(v), 1.0f);
            if (au < halfWidth * 0.5f || au > 1.0f - halfWidth * 0.5f) uComp = 1.0f;
            else if (au > 0.5f - halfWidth * 0.5f && au < 0.5f + halfWidth * 0.5f) uComp = -1.0f;
            if (av < halfWidth * 0.5f || av > 1.0f - halfWidth * 0.5f) vComp = 1.0f;
            else if (av > 0.5f - halfWidth * 0.5f && av < 0.5f + halfWidth * 0.5f) vComp = -1.0f;
            if (av > 0.5f) uComp *= -1;
            if (au > 0.5f) vComp *= -1;
            if (t[8]) { uComp *= -1; vComp *= -1; }
            V3 n = normalize(mk(uComp, vComp, 1.0f));
            return mk(0.5f * n.x + 0.5f, 0.5f * n.y + 0.5f, 0.5f * n.z + 0.5f);
        }
        default: return mk(0, 0, 0);
    }
}
V3 evalColor(const uint8_t* t, float u, float v) {
    switch (t[0]) {
        case 0: return rd3(t + 16);
        case 1:
        case 2: {
            uint32_t w = rd<uint32_t>(t + 4), h = rd<uint32_t>(t + 8);
            const uint8_t* d = t + rd<int32_t>(t + 12) + 4 * ((size_t)texel(v, h) * w + texel(u, w));
            return mk((float)d[0] / 255.0f, (float)d[1] / 255.0f, (float)d[2] / 255.0f);
        }
        case 3: {
            uint32_t w = rd<uint32_t>(t + 4), h = rd<uint32_t>(t + 8);
            const uint8_t* d = t + rd<int32_t>(t + 12) + 8 * ((size_t)texel(v, h) * w + texel(u, w));
            return mk(halfToFloat(rd<uint16_t>(d)), halfToFloat(rd<uint16_t>(d + 2)), halfToFloat(rd<uint16_t>(d + 4)));
        }
        case 4: return proceduralColor(t, u, v);
        case 5: {
            uint32_t w = rd<uint32_t>(t + 4), h = rd<uint32_t>(t + 8);
            float g = (float)t[rd<int32_t>(t + 12) + (size_t)texel(v, h) * w + texel(u, w)] / 255.0f;
            return mk(g, g, g);
        }
        case 9: {
            uint32_t w = rd<uint32_t>(t + 4), h = rd<uint32_t>(t + 8);
            return rd3(t + rd<int32_t>(t + 12) + 12 * ((size_t)texel(v, h) * w + texel(u, w)));
        }
        default: return mk(0, 0, 0);
    }
}
float evalFloat(const uint8_t* t, float u, float v) {
    switch (t[0]) {
        case 6: return rd<float>(t + 4);
        case 7: {
            uint32_t w = rd<uint32_t>(t + 4), h = rd<uint32_t>(t + 8);
            return rd<float>(t + rd<int32_t>(t + 12) + 4 * ((size_t)texel(v, h) * w + texel(u, w)));
        }
        case 8: {
            if (t[1] != 0) return 0.0f;
            uint32_t i = ((u * 2 > 0 ? (uint32_t)(u * 2) : 0u) + (v * 2 > 0 ? (uint32_t)(v * 2) : 0u)) % 2;
            return rd<float>(t + 4 + 4 * i);
        }
        default: return 0.0f;
    }
}
float evalAlpha(const uint8_t* t, float u, float v) {
    switch (t[0]) {
        case 2: {
            uint32_t w = rd<uint32_t>(t + 4), h = rd<uint32_t>(t + 8);
            return (float)t[rd<int32_t>(t + 12) + 4 * ((size_t)texel(v, h) * w + texel(u, w)) + 3] / 255.0f;
        }
        case 5: {
            uint32_t w = rd<uint32_t>(t + 4), h = rd<uint32_t>(t + 8);
            return (float)t[rd<int32_t>(t + 12) + (size_t)texel(v, h) * w + texel(u, w)] / 255.0f;
        }
        case 8: return evalFloat(t, u, v);
        default: return 1.0f;
    }
}

// ---- distributions (distribution.cl:59-129) ----
uint32_t sampleDiscrete1D(const uint8_t* d, float u, float* prob) {
    const uint8_t* cdf = d + rd<int32_t>(d + 12);
    uint32_t start = 0, end = rd<uint32_t>(d + 4) + 1, i = end / 2;
    while (true) {
        if (rd<float>(cdf + 4 * i) > u) end = i; else start = i;
        if (end - start == 1) { *prob = rd<float>(d + rd<int32_t>(d + 8) + 4 * start); return start; }
        i = (end - start) / 2 + start;
    }
}
inline float probDiscrete1D(const uint8_t* d, uint32_t k) { return rd<float>(d + rd<int32_t>(d + 8) + 4 * k); }
// 1D piecewise-constant; `base` is where the PDF/CDF offsets are relative to.
float sampleCC1D(const uint8_t* d, const uint8_t* base, float u, float* pdf) {
    const uint8_t* cdf = base + rd<int32_t>(d + 24);
    uint32_t start = 0, end = rd<uint32_t>(d + 4) + 1, i = end / 2;
    while (true) {
        if (rd<float>(cdf + 4 * i) > u) end = i; else start = i;
        if (end - start == 1) {
            float c0 = rd<float>(cdf + 4 * start), c1 = rd<float>(cdf + 4 * end);
            float t = (u - c0) / (c1 - c0);
            *pdf = rd<float>(base + rd<int32_t>(d + 20) + 4 * start);
            return rd<float>(d + 8) + ((float)start + t) * rd<float>(d + 16);
        }
        i = (end - start) / 2 + start;
    }
}
inline uint32_t stratum(const uint8_t* d, float p) {  // D10
    float f = (p - rd<float>(d + 8)) / rd<float>(d + 16);
    return std::min(f > 0.0f ? (uint32_t)f : 0u, rd<uint32_t>(d + 4) - 1u);
}
inline float pdfCC1D(const uint8_t* d, const uint8_t* base, float p) {
    return rd<float>(base + rd<int32_t>(d + 20) + 4 * stratum(d, p));
}
void sampleCC2D(const uint8_t* d, float u0, float u1, float* uv, float* pdf) {
    const uint8_t* parent = d + 8;
    float pdfTop;
    uv[1] = sampleCC1D(parent, d, u1, &pdfTop);  // D7: parent offsets relative to d
    const uint8_t* child = d + rd<int32_t>(d + 4) + 28 * (size_t)stratum(parent, uv[1]);
    uv[0] = sampleCC1D(child, child, u0, pdf);
    *pdf *= pdfTop;
}
float pdfCC2D(const uint8_t* d, float u, float v) {
    const uint8_t* parent = d + 8;
    float r = pdfCC1D(parent, d, v);
    const uint8_t* child = d + rd<int32_t>(d + 4) + 28 * (size_t)stratum(parent, v);
    return r * pdfCC1D(child, child, u);
}

// ---- geometry records (global.cl:38-88) ----
struct SurfacePoint {
    V3 p{}, gNormal{}, sNormal{}, sTangent{}, uDir{};
    float uv[2] = {0, 0};
    uint32_t faceID = kNoFace;
    bool hasTangent = false;
};
struct LightPoint {
    V3 p{}, gNormal{}, sNormal{}, sTangent{}, uDir{};
    float uv[2] = {0, 0};
    uint32_t faceID = kNoFace;
    bool hasTangent = false, atInfinity = false;
};
void makeTangent(V3 n, V3* t) {
    if (std::fabs(n.x) > std::fabs(n.y)) {
        float inv = 1.0f / std::sqrt(n.x * n.x + n.z * n.z);
        *t = mk(-n.z * inv, 0.0f, n.x * inv);
    } else {
        float inv = 1.0f / std::sqrt(n.y * n.y + n.z * n.z);
        *t = mk(0.0f, n.z * inv, -n.y * inv);
    }
}
inline V3 worldToLocal(V3 s, V3 t, V3 n, V3 v) { return {dot(s, v), dot(t, v), dot(n, v)}; }
inline V3 localToWorld(V3 s, V3 t, V3 n, V3 v) {
    return {s.x * v.x + t.x * v.y + n.x * v.z, s.y * v.x + t.y * v.y + n.y * v.z, s.z * v.x + t.z * v.y + n.z * v.z};
}

// ---- traversal (bvh_traversal.cl) ----
bool boxTest(const float* bmin, const float* bmax, V3 org, V3 dir, float tBound) {
    const float o[3] = {org.x, org.y, org.z}, dv[3] = {dir.x, dir.y, dir.z};
    int lmu[3];
    for (int k = 0; k < 3; ++k) lmu[k] = o[k] < bmin[k] ? -1 : (o[k] > bmax[k] ? 1 : 0);
    if (lmu[0] | lmu[1] | lmu[2]) {
        float t[3];
        for (int k = 0; k < 3; ++k) {
            float plane = o[k] < bmin[k] ? bmin[k] : (o[k] > bmax[k] ? bmax[k] : 0.0f);
            t[k] = (lmu[k] != 0 && dv[k] != 0.0f) ? ((plane - o[k]) / dv[k]) : -1.0f;
        }
        float maxt = omax(t[0], omax(t[1], t[2]));
        if (maxt < 0.0f || maxt > tBound) return false;
        for (int k = 0; k < 3; ++k) {
            float c = o[k] + maxt * dv[k];
            if (!((c >= bmin[k] && c <= bmax[k]) || t[k] == maxt)) return false;
        }
    }
    return true;
}
struct Isect {
    uint32_t face;
    float b1, b2;
};
bool triTest(const SceneC& s, V3 org, V3 dir, uint32_t fi, float* t, Isect* is) {
    const clr_face& f = s.faces[fi];
    V3 p0 = vtx(s.vertices, f.p0), p1 = vtx(s.vertices, f.p1), p2 = vtx(s.vertices, f.p2);
    V3 e01 = p1 - p0, e02 = p2 - p0;
    V3 p = cross(dir, e02);
    float det = dot(e01, p);
    if (det > -0.000001f && det < 0.000001f) return false;
    float invDet = 1.0f / det;
    V3 d = org - p0;
    float b1 = dot(d, p) * invDet;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    V3 q = cross(d, e01);
    float b2 = dot(dir, q) * invDet;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
    float tt = dot(e02, q) * invDet;
    if (tt < 0.0f || tt > *t) return false;
    if (f.uv0 != 0xFFFFFFFFu && f.alphaTexPtr != 0xFFFFFFFFu) {
        float b0 = 1.0f - b1 - b2;
        const float* uv0 = s.uvs + 2 * (size_t)f.uv0;
        const float* uv1 = s.uvs + 2 * (size_t)f.uv1;
        const float* uv2 = s.uvs + 2 * (size_t)f.uv2;
        float u = b0 * uv0[0] + b1 * uv1[0] + b2 * uv2[0];
        float v = b0 * uv0[1] + b1 * uv1[1] + b2 * uv2[1];
        if (evalAlpha(s.textures + f.alphaTexPtr, u, v) == 0.0f) return false;  // D3
    }
    *t = tt;
    is->face = fi;
    is->b1 = b1;
    is->b2 = b2;
    return true;
}
// calcHitpointParameters (bvh_traversal.cl:110-158).
void hitParams(const SceneC& s, V3 org, V3 dir, float t, const Isect& is, SurfacePoint* sp) {
    const clr_face& f = s.faces[is.face];
    V3 p0 = vtx(s.vertices, f.p0), p1 = vtx(s.vertices, f.p1), p2 = vtx(s.vertices, f.p2);
    sp->faceID = is.face;
    sp->p = org + dir * t;
    sp->gNormal = normalize(cross(p1 - p0, p2 - p0));
    float b0 = 1.0f - is.b1 - is.b2;  // param = (b0, b1)
    float b1 = is.b1;
    float b2 = 1.0f - b0 - b1;
    bool hasUV = f.uv0 != 0xFFFFFFFFu;
    sp->uv[0] = sp->uv[1] = 0.0f;
    if (hasUV) {
        const float* uv0 = s.uvs + 2 * (size_t)f.uv0;
        const float* uv1 = s.uvs + 2 * (size_t)f.uv1;
        const float* uv2 = s.uvs + 2 * (size_t)f.uv2;
        float ib2 = is.b2;
        sp->uv[0] = b0 * uv0[0] + b1 * uv1[0] + ib2 * uv2[0];
        sp->uv[1] = b0 * uv0[1] + b1 * uv1[1] + ib2 * uv2[1];
    }
    bool hasVN = f.vn0 != 0xFFFFFFFFu;
    if (hasVN)
        sp->sNormal = normalize(b0 * vtx(s.normals, f.vn0) + b1 * vtx(s.normals, f.vn1) + b2 * vtx(s.normals, f.vn2));
    else
        sp->sNormal = sp->gNormal;
    sp->hasTangent = f.vt0 != 0xFFFFFFFFu;
    if (sp->hasTangent)
        sp->sTangent = normalize(b0 * vtx(s.tangents, f.vt0) + b1 * vtx(s.tangents, f.vt1) + b2 * vtx(s.tangents, f.vt2));
    if (hasUV) {
        const float* uv0 = s.uvs + 2 * (size_t)f.uv0;
        const float* uv1 = s.uvs + 2 * (size_t)f.uv1;
        const float* uv2 = s.uvs + 2 * (size_t)f.uv2;
        float du0 = uv0[0] - uv2[0], dv0 = uv0[1] - uv2[1];
        float du1 = uv1[0] - uv2[0], dv1 = uv1[1] - uv2[1];
        V3 dP0 = p0 - p2, dP1 = p1 - p2;
        float invDetUV = 1.0f / (du0 * dv1 - dv0 * du1);
        V3 uDir = invDetUV * mk(dv1 * dP0.x - dv0 * dP1.x, dv1 * dP0.y - dv0 * dP1.y, dv1 * dP0.z - dv0 * dP1.z);
        if (hasVN)
            sp->uDir = normalize(cross(cross(sp->sNormal, uDir), sp->sNormal));
        else
            sp->uDir = normalize(uDir);
    }
}
// rayIntersection (bvh_traversal.cl:182-215), or the ordered traversal: its
// 4-wide form for renders (the library's renderers walk that form; the hit
// is the same), the binary form for the BVH view. boxHits (optional):
// rayIntersectionVis's count (bvh_traversal.cl:216-319: internal and leaf
// boxes passed on the LBVH, node boxes passed on the midpoint BVH), or the
// binary ordered traversal's entered nodes plus tests.
bool rayIntersection(const SceneC& s, V3 org, V3 dir, SurfacePoint* sp, uint32_t* boxHits = nullptr) {
    uint32_t hits = 0;
    struct Commit {
        uint32_t* out;
        uint32_t& n;
        ~Commit() {
            if (out) *out = n;
        }
    } commit{boxHits, hits};
    float t = INFINITY;
    Isect best{kNoFace, 0, 0}, cur{};
    if (s.ordered) {
        const float o[3] = {org.x, org.y, org.z}, d[3] = {dir.x, dir.y, dir.z};
        auto tri = [&](uint32_t face, float bound, float* tt) {
            float x = bound;
            Isect tmp{};
            if (!triTest(s, org, dir, face, &x, &tmp)) return false;
            *tt = x;
            return true;
        };
        orc_ordered::Counts cnt;
        const uint32_t f =
            boxHits ? orc_ordered::traverse(s.inodes, s.leaves, s.numFaces, o, d, t, tri, cnt, false, nullptr)
                    : orc_ordered::traverseWide(s.inodes, s.leaves, s.numFaces, o, d, t, tri, cnt, false, nullptr);
        hits = cnt.visits + cnt.tests;
        ++t_work[0];
        t_work[1] += cnt.visits;
        t_work[2] += cnt.tests;
        const bool hit = f != kNoFace && !std::isinf(t);
        if (hit) {
            // The accepted triangle's barycentrics: the same test again.
            float x = INFINITY;
            triTest(s, org, dir, f, &x, &best);
            hitParams(s, org, dir, t, best, sp);
        } else {
            sp->faceID = kNoFace;  // D4
        }
        return hit;
    }
    auto tryLeaf = [&](uint32_t li) {
        const clr_lbvh_leaf& L = s.leaves[li];
        if (!boxTest(L.min, L.max, org, dir, t)) return;
        ++hits;
        if (triTest(s, org, dir, L.objIdx, &t, &cur)) best = cur;
    };
    if (s.bvh) {
        // rayIntersection without __USE_LBVH (bvh_traversal.cl:256-285).
        uint32_t stack[64];
        uint32_t depth = 0;
        stack[depth++] = 0;
        while (depth > 0) {
            const clr_bvh_node& N = s.bvh[stack[--depth]];
            if (!boxTest(N.min, N.max, org, dir, t)) continue;
            ++hits;
            if (N.children[1] != 0xFFFFFFFFu) {
                if (depth + 2 <= 64) {
                    stack[depth++] = N.children[1];
                    stack[depth++] = N.children[0];
                }
            } else if (triTest(s, org, dir, N.children[0], &t, &cur)) {
                best = cur;
            }
        }
    } else if (s.numFaces == 1) {
        tryLeaf(0);
    } else {
        uint32_t stack[64];
        uint32_t depth = 0;
        stack[depth++] = 0;
        while (depth > 0) {
            const clr_lbvh_inode& N = s.inodes[stack[--depth]];
            if (!boxTest(N.min, N.max, org, dir, t)) continue;
            ++hits;
            for (int i = 0; i < 2; ++i) {
                if (!N.isChild[i]) {
                    if (depth < 64) stack[depth++] = N.c[i];
                    continue;
                }
                tryLeaf(N.c[i]);
            }
        }
    }
    bool hit = best.face != kNoFace && !std::isinf(t);
    if (hit) hitParams(s, org, dir, t, best, sp);
    else sp->faceID = kNoFace;  // D4
    return hit;
}

// ---- BSDF (reflection.cl) ----
enum : uint32_t { BxDF_Reflection = 1, BxDF_Transmission = 2, BxDF_Diffuse = 4, BxDF_Glossy = 8, BxDF_Specular = 16,
                  BxDF_Non_Singular = 12, BxDF_All = 31 };
struct BxDF {
    uint8_t id;
    uint32_t type;
    V3 R{}, Rs{};
    float a = 0, b = 0;
    const uint8_t* fresnel = nullptr;
};
struct BSDF {
    int n = 0;
    BxDF c[4];
    V3 N{}, S{}, T{}, NG{};
};
inline float cosT(V3 v) { return v.z; }
inline float absCosT(V3 v) { return std::fabs(v.z); }
inline float sinT2(V3 v) { return std::max(1.0f - v.z * v.z, 0.0f); }
inline float sinTh(V3 v) { return std::sqrt(sinT2(v)); }
inline float cosPhi(V3 v) { float s = sinTh(v); if (s == 0.0f) return 1.0f; return clampf(v.x / s, -1.0f, 1.0f); }
inline float sinPhi(V3 v) { float s = sinTh(v); if (s == 0.0f) return 0.0f; return clampf(v.y / s, -1.0f, 1.0f); }
inline bool matchType(const BxDF& f, uint32_t mask) { return (f.type & mask) == f.type; }

// Normal mapping (reflection.cl:201-223): the texel (rgb in [0, 1]) decodes
// to a tangent-space normal nb = normalize(2 rgb - 1) in the frame (uDir,
// vDir = uDir x sNormal, sNormal); N is that normal in world space, and the
// tangent S is turned by the rotation that takes sNormal to N (axis
// sNormal x N, normalised: D9), skipped when the angle's sine is <= 1e-4.
void normalMapFrame(V3 sNormal, V3 uDir, V3 tex, V3* N, V3* S) {
    const V3 nb = normalize(mk(2.0f * tex.x - 1.0f, 2.0f * tex.y - 1.0f, 2.0f * tex.z - 1.0f));
    const V3 vDir = cross(uDir, sNormal);
    *N = mk(dot(mk(uDir.x, vDir.x, sNormal.x), nb), dot(mk(uDir.y, vDir.y, sNormal.y), nb),
            dot(mk(uDir.z, vDir.z, sNormal.z), nb));
    V3 ax = cross(sNormal, *N);
    const float sinTH = std::fmin(length(ax), 1.0f);
    if (sinTH > 0.0001f) {
        ax = normalize(ax);  // D9
        const float cosTH = std::cos(std::asin(sinTH));
        const float oneM = 1 - cosTH;
        const V3 S0 = *S;
        *S = mk(dot(mk(ax.x * ax.x * oneM + cosTH, ax.x * ax.y * oneM - ax.z * sinTH, ax.z * ax.x * oneM + ax.y * sinTH), S0),
                dot(mk(ax.x * ax.y * oneM + ax.z * sinTH, ax.y * ax.y * oneM + cosTH, ax.y * ax.z * oneM - ax.x * sinTH), S0),
                dot(mk(ax.z * ax.x * oneM - ax.y * sinTH, ax.y * ax.z * oneM + ax.x * sinTH, ax.z * ax.z * oneM + cosTH), S0));
    }
}

void bsdfAlloc(const SceneC& s, uint32_t off, const SurfacePoint& sp, BSDF* B) {
    const uint8_t* m = s.materials + off;
    B->n = m[0];
    B->N = sp.sNormal;
    if (sp.hasTangent) B->S = sp.sTangent;
    else makeTangent(sp.sNormal, &B->S);
    if (m[1]) {  // normal map (reflection.cl:201-223)
        const V3 tex = evalColor(s.textures + rd<uint32_t>(m + 4), sp.uv[0], sp.uv[1]);
        normalMapFrame(sp.sNormal, sp.uDir, tex, &B->N, &B->S);
    }
    B->T = cross(B->N, B->S);
    B->NG = sp.gNormal;
    size_t p = off + 8;
    for (int i = 0; i < B->n; ++i) {
        p = (p + 3) & ~size_t(3);
        const uint8_t* e = s.materials + p;
        BxDF& f = B->c[i];
        f.id = e[0];
        switch (e[0]) {
            case 0: {
                f.type = BxDF_Reflection | BxDF_Diffuse;
                f.R = evalColor(s.textures + rd<uint32_t>(e + 4), sp.uv[0], sp.uv[1]);
                float s2 = evalFloat(s.textures + rd<uint32_t>(e + 8), sp.uv[0], sp.uv[1]);
                s2 *= s2;
                f.a = 1.0f - (s2 / (2.0f * (s2 + 0.33f)));
                f.b = 0.45f * s2 / (s2 + 0.09f);
                p += 12;
                break;
            }
            case 1:
                f.type = BxDF_Reflection | BxDF_Specular;
                f.R = evalColor(s.textures + rd<uint32_t>(e + 4), sp.uv[0], sp.uv[1]);
                f.fresnel = s.others + rd<uint32_t>(e + 8);
                p += 12;
                break;
            case 2:
                f.type = BxDF_Transmission | BxDF_Specular;
                f.R = evalColor(s.textures + rd<uint32_t>(e + 4), sp.uv[0], sp.uv[1]);
                f.a = rd<float>(e + 8);
                f.b = rd<float>(e + 12);
                f.fresnel = s.others + rd<uint32_t>(e + 16);
                p += 20;
                break;
            case 3:
                f.type = BxDF_Reflection | BxDF_Glossy;
                f.R = evalColor(s.textures + rd<uint32_t>(e + 4), sp.uv[0], sp.uv[1]);
                f.a = evalFloat(s.textures + rd<uint32_t>(e + 8), sp.uv[0], sp.uv[1]);
                f.b = evalFloat(s.textures + rd<uint32_t>(e + 12), sp.uv[0], sp.uv[1]);
                p += 16;
                break;
            case 4:
                f.type = BxDF_Reflection | BxDF_Glossy;
                f.R = evalColor(s.textures + rd<uint32_t>(e + 4), sp.uv[0], sp.uv[1]);
                f.a = evalFloat(s.textures + rd<uint32_t>(e + 8), sp.uv[0], sp.uv[1]);
                f.b = evalFloat(s.textures + rd<uint32_t>(e + 12), sp.uv[0], sp.uv[1]);
                p += 16;
                break;
            case 5:
                f.type = BxDF_Reflection | BxDF_Diffuse;
                f.R = evalColor(s.textures + rd<uint32_t>(e + 4), sp.uv[0], sp.uv[1]);   // Rs
                f.Rs = evalColor(s.textures + rd<uint32_t>(e + 8), sp.uv[0], sp.uv[1]);  // Rd
                p += 12;
                break;
            case 6:
                f.type = BxDF_Reflection | BxDF_Glossy;
                f.R = evalColor(s.textures + rd<uint32_t>(e + 4), sp.uv[0], sp.uv[1]);
                f.a = evalFloat(s.textures + rd<uint32_t>(e + 8), sp.uv[0], sp.uv[1]);
                f.fresnel = s.others + rd<uint32_t>(e + 12);
                p += 16;
                break;
            default: f.type = 0; break;
        }
    }
}

V3 evalFresnel(const uint8_t* fr, float cosi) {
    switch (fr[0]) {
        case 0: return mk(1, 1, 1);
        case 1: {
            V3 eta = rd3(fr + 16), k = rd3(fr + 32);
            cosi = std::fabs(cosi);
            V3 tmp = (eta * eta + k * k) * cosi * cosi;
            V3 one = mk(1, 1, 1);
            V3 twoEtaCos = (2.0f * eta) * cosi;
            V3 a = tmp - twoEtaCos + one, b = tmp + twoEtaCos + one;
            V3 Rparl2 = mk(a.x / b.x, a.y / b.y, a.z / b.z);
            V3 tmpf = eta * eta + k * k;
            V3 c2 = mk(cosi * cosi, cosi * cosi, cosi * cosi);
            V3 c = tmpf - twoEtaCos + c2, d = tmpf + twoEtaCos + c2;
            V3 Rperp2 = mk(c.x / d.x, c.y / d.y, c.z / d.z);
            return (Rparl2 + Rperp2) / 2.0f;
        }
        case 2: {
            cosi = clampf(cosi, -1.0f, 1.0f);
            bool entering = cosi > 0.0f;
            float ei = rd<float>(fr + 4), et = rd<float>(fr + 8);
            if (!entering) std::swap(ei, et);
            float sint = ei / et * std::sqrt(std::fmax(0.0f, 1.0f - cosi * cosi));
            if (sint >= 1.0f) return mk(1, 1, 1);
            float cost = std::sqrt(std::fmax(0.0f, 1.0f - sint * sint));
            cosi = std::fabs(cosi);
            float Rparl = ((et * cosi) - (ei * cost)) / ((et * cosi) + (ei * cost));
            float Rperp = ((ei * cosi) - (et * cost)) / ((ei * cosi) + (et * cost));
            float v = (Rparl * Rparl + Rperp * Rperp) / 2.0f;
            return mk(v, v, v);
        }
        default: return mk(0, 0, 0);
    }
}

// Beckmann microfacet (addition D8): D(h) = exp(-tan^2/a^2) / (pi a^2 cos^4),
// Smith-Beckmann G1 (Walter et al. 2007 rational fit), f = R F(wo.h) D G / (4 |wo.z||wi.z|),
// sampling h proportional to D(h) cos(theta_h), pdf = D |cos_h| / (4 |wo.h|).
float beckmannD(V3 h, float a) {
    float c2 = h.z * h.z;
    if (c2 <= 0.0f) return 0.0f;
    float t2 = (1.0f - c2) / c2;
    return std::exp(-t2 / (a * a)) / (kPi * a * a * c2 * c2);
}
float beckmannG1(V3 v, float a) {
    float c2 = v.z * v.z;
    if (c2 >= 1.0f) return 1.0f;
    float tanT = std::sqrt(std::fmax(0.0f, 1.0f - c2)) / std::fabs(v.z);
    float x = 1.0f / (a * tanT);
    if (x >= 1.6f) return 1.0f;
    return (3.535f * x + 2.181f * x * x) / (1.0f + 2.276f * x + 2.577f * x * x);
}
inline V3 halfDir(V3 wo, V3 wi) {
    V3 h = normalize(wo + wi);
    return h.z < 0.0f ? -h : h;
}

V3 fx(const BxDF& f, V3 wo, V3 wi);
float fxPdf(const BxDF& f, V3 wo, V3 wi);

V3 sampleFx(const BxDF& f, V3 wo, float u0, float u1, V3* wi, float* pdf) {
    switch (f.id) {
        case 0:
        case 5:
            *wi = cosineSampleHemisphere(u0, u1);
            if (wo.z < 0.0f) wi->z *= -1;
            *pdf = absCosT(*wi) / kPi;
            return fx(f, wo, *wi);
        case 1:
            *wi = mk(-wo.x, -wo.y, wo.z);
            *pdf = 1.0f;
            {
                V3 F = evalFresnel(f.fresnel, cosT(wo));
                V3 r = f.R * F;
                return r / absCosT(*wi);
            }
        case 2: {
            bool entering = cosT(wo) > 0.0f;
            float ei = f.a, et = f.b;
            if (!entering) std::swap(ei, et);
            float sini2 = sinT2(wo);
            float eta = ei / et;
            float sint2 = eta * eta * sini2;
            if (sint2 >= 1.0f) return mk(0, 0, 0);
            float cost = std::sqrt(std::fmax(0.0f, 1.0f - sint2));
            if (entering) cost = -cost;
            *wi = mk(eta * -wo.x, eta * -wo.y, cost);
            *pdf = 1.0f;
            V3 F = evalFresnel(f.fresnel, cosT(wo));
            V3 r = (mk(1, 1, 1) - F) * f.R;
            return r / absCosT(*wi);
        }
        case 3: {
            float quad = 2 * kPi * u1;
            float phi = std::atan2(f.b * std::sin(quad), f.a * std::cos(quad));
            float cpa = std::cos(phi) / f.a, spa = std::sin(phi) / f.b;
            float th = std::atan(std::sqrt(-std::log(1 - u0) / (cpa * cpa + spa * spa)));
            if (wo.z < 0) th = kPi - th;
            V3 h = mk(std::sin(th) * std::cos(phi), std::sin(th) * std::sin(phi), std::cos(th));
            *wi = 2 * dot(wo, h) * h - wo;
            if (wi->z * wo.z <= 0) { *pdf = 0.0f; return mk(0, 0, 0); }
            float hx = h.x / f.a, hy = h.y / f.b;
            float dHN = absCosT(h), dHI = dot(h, *wi);
            float num = std::exp(-(hx * hx + hy * hy) / (dHN * dHN));
            float den = 4 * kPi * f.a * f.b * dHI * dHN * dHN * dHN;
            *pdf = num / den;
            return f.R * (num / (den * dHI * dHN));
        }
        case 4: {
            float quad = 2 * kPi * u1;
            float phi = std::atan2(std::sqrt(f.a + 1) * std::sin(quad), std::sqrt(f.b + 1) * std::cos(quad));
            float cp = std::cos(phi), sp = std::sin(phi);
            float th = std::acos(std::pow(1 - u0, 1.0f / (f.a * cp * cp + f.b * sp * sp + 1)));
            if (wo.z < 0) th = kPi - th;  // D5
            V3 h = mk(std::sin(th) * std::cos(phi), std::sin(th) * std::sin(phi), std::cos(th));
            *wi = 2 * dot(wo, h) * h - wo;
            if (wo.z * wi->z <= 0) { *pdf = 0.0f; return mk(0, 0, 0); }
            *pdf = fxPdf(f, wo, *wi);
            return fx(f, wo, *wi);
        }
        case 6: {
            float a = f.a;
            float t2 = -a * a * std::log(1.0f - u0);
            float phi = 2.0f * kPi * u1;
            float cth = 1.0f / std::sqrt(1.0f + t2);
            float sth = std::sqrt(std::fmax(0.0f, 1.0f - cth * cth));
            V3 h = mk(sth * std::cos(phi), sth * std::sin(phi), cth);
            if (wo.z < 0.0f) h = -h;
            *wi = 2 * dot(wo, h) * h - wo;
            if (wi->z * wo.z <= 0) { *pdf = 0.0f; return mk(0, 0, 0); }
            *pdf = fxPdf(f, wo, *wi);
            return fx(f, wo, *wi);
        }
        default: break;
    }
    return mk(0, 0, 0);
}

V3 fx(const BxDF& f, V3 wo, V3 wi) {
    switch (f.id) {
        case 0: {
            if (wi.z * wo.z <= 0.0f) return mk(0, 0, 0);
            if (f.a == 1.0f) return f.R / kPi;
            float sti = sinTh(wi), sto = sinTh(wo);
            float maxCos = 0.0f;
            if (sti > 1e-4f && sto > 1e-4f) {
                float dcos = cosPhi(wi) * cosPhi(wo) + sinPhi(wi) * sinPhi(wo);
                maxCos = std::fmax(0.0f, dcos);
            }
            float sinA, tanB;
            if (absCosT(wi) > absCosT(wo)) { sinA = sto; tanB = sti / absCosT(wi); }
            else { sinA = sti; tanB = sto / absCosT(wo); }
            return f.R / kPi * (f.a + f.b * maxCos * sinA * tanB);
        }
        case 1:
        case 2: return mk(0, 0, 0);
        case 3: {
            if (wi.z * wo.z <= 0.0f) return mk(0, 0, 0);
            V3 h = normalize(wo + wi);
            float hx = h.x / f.a, hy = h.y / f.b;
            float dHN = absCosT(h), dHI = dot(h, wi);
            return f.R * std::exp(-(hx * hx + hy * hy) / (dHN * dHN)) /
                   (4 * kPi * f.a * f.b * dHI * dHI * dHN * dHN * dHN * dHN);
        }
        case 4: {
            if (wi.z * wo.z <= 0.0f) return mk(0, 0, 0);
            V3 h = normalize(wo + wi);
            float dHV = dot(h, wo);
            float e = (f.a * h.x * h.x + f.b * h.y * h.y) / (1 - h.z * h.z);
            V3 fr = f.R + (mk(1, 1, 1) - f.R) * std::pow(1.0f - dHV, 5.0f);
            return std::sqrt((f.a + 1) * (f.b + 1)) / (8 * kPi) * std::pow(std::fabs(h.z), e) /
                   (dHV * std::fmax(absCosT(wo), absCosT(wi))) * fr;
        }
        case 5: {
            if (wi.z * wo.z <= 0.0f) return mk(0, 0, 0);
            // 28 Rd / (23 pi) (1 - Rs) (1 - (1 - |cos_o|/2)^5) (1 - (1 - |cos_i|/2)^5)
            V3 r = 28 * f.Rs / (23 * kPi) * (mk(1, 1, 1) - f.R);
            return r * (1.0f - std::pow(1.0f - absCosT(wo) / 2, 5.0f)) * (1.0f - std::pow(1.0f - absCosT(wi) / 2, 5.0f));
        }
        case 6: {
            if (wi.z * wo.z <= 0.0f) return mk(0, 0, 0);
            V3 h = halfDir(wo, wi);
            float D = beckmannD(h, f.a);
            float G = beckmannG1(wo, f.a) * beckmannG1(wi, f.a);
            V3 F = evalFresnel(f.fresnel, dot(wo, h));
            float k = D * G / ((planted(kPlantBeckmannNo4) ? 1.0f : 4.0f) * absCosT(wo) * absCosT(wi));
            return (f.R * F) * k;
        }
        default: break;
    }
    return mk(0, 0, 0);
}

float fxPdf(const BxDF& f, V3 wo, V3 wi) {
    switch (f.id) {
        case 0:
        case 5: return absCosT(wi) / kPi;
        case 1:
        case 2: return 0.0f;
        case 3: {
            if (wi.z * wo.z <= 0) return 0.0f;
            V3 h = normalize(wo + wi);
            float hx = h.x / f.a, hy = h.y / f.b;
            float dHN = absCosT(h), dHI = dot(h, wi);
            float num = std::exp(-(hx * hx + hy * hy) / (dHN * dHN));
            return num / (4 * kPi * f.a * f.b * dHI * dHN * dHN * dHN);
        }
        case 4: {
            if (wo.z * wi.z <= 0) return 0.0f;
            V3 h = normalize(wo + wi);
            float e = (f.a * h.x * h.x + f.b * h.y * h.y) / (1 - h.z * h.z);
            return std::sqrt((f.a + 1) * (f.b + 1)) / (2 * kPi) * std::pow(std::fabs(h.z), e) / (4 * dot(wo, h));
        }
        case 6: {
            if (wo.z * wi.z <= 0) return 0.0f;
            V3 h = halfDir(wo, wi);
            return beckmannD(h, f.a) * std::fabs(h.z) / (4.0f * std::fabs(dot(wo, h)));
        }
        default: break;
    }
    return 0.0f;
}

bool hasNonSpecular(const BSDF& B) {
    for (int i = 0; i < B.n; ++i)
        if (B.c[i].type & BxDF_Non_Singular) return true;
    return false;
}
V3 sampleFs(const BSDF& B, V3 vout, float uc, float u0, float u1, V3* vin, float* pdf, uint32_t* sampledType) {
    uint32_t flags = BxDF_All;
    uint32_t nm = (uint32_t)B.n;
    V3 wo = worldToLocal(B.S, B.T, B.N, vout);
    uint32_t which = planted(kPlantMixPickFirst) ? 0u : randUInt(uc, (uint32_t)B.n);
    const BxDF& f = B.c[which];
    V3 wi{};
    V3 ret = sampleFx(f, wo, u0, u1, &wi, pdf);
    if (*pdf == 0.0f) { *sampledType = 0; return mk(0, 0, 0); }
    *sampledType = f.type;
    *vin = localToWorld(B.S, B.T, B.N, wi);
    if (!(f.type & BxDF_Specular) && nm > 1 && !planted(kPlantMixOwnPdfOnly))
        for (int i = 0; i < B.n; ++i)
            if ((uint32_t)i != which && matchType(B.c[i], flags)) *pdf += fxPdf(B.c[i], wo, wi);
    if (!planted(kPlantMixNoAverage)) *pdf /= (float)nm;
    if (!(f.type & BxDF_Specular)) {
        ret = mk(0, 0, 0);
        if (dot(*vin, B.NG) * dot(vout, B.NG) > 0) flags &= ~BxDF_Transmission;
        else flags &= ~BxDF_Reflection;
        for (int i = 0; i < B.n; ++i)
            if (matchType(B.c[i], flags)) ret += fx(B.c[i], wo, wi);
    }
    return ret;
}
V3 fs(const BSDF& B, V3 vout, V3 vin) {
    uint32_t flags = BxDF_All;
    V3 wo = worldToLocal(B.S, B.T, B.N, vout), wi = worldToLocal(B.S, B.T, B.N, vin);
    V3 ret = mk(0, 0, 0);
    if (dot(vin, B.NG) * dot(vout, B.NG) > 0) flags &= ~BxDF_Transmission;
    else flags &= ~BxDF_Reflection;
    for (int i = 0; i < B.n; ++i)
        if (matchType(B.c[i], flags)) ret += fx(B.c[i], wo, wi);
    return ret;
}
float fsPdf(const BSDF& B, V3 vout, V3 vin) {
    V3 wo = worldToLocal(B.S, B.T, B.N, vout), wi = worldToLocal(B.S, B.T, B.N, vin);
    float pdf = 0.0f;
    for (int i = 0; i < B.n; ++i)
        if (matchType(B.c[i], BxDF_All)) pdf += fxPdf(B.c[i], wo, wi);
    return planted(kPlantMixNoAverage) ? pdf : pdf / (float)B.n;
}
inline float absCosNsBSDF(const BSDF& B, V3 v) { return std::fabs(dot(B.N, v)); }

// ---- lights (light.cl) ----
struct EDF {
    bool env = false;
    int n = 0;
    V3 L[4];  // M (area) or Le (environment) per element
    V3 N{}, S{}, T{}, NG{};
};
void edfAlloc(const SceneC& s, uint32_t off, const LightPoint& lp, EDF* E) {
    if (lp.atInfinity) {
        E->env = true;
        const uint8_t* lpInfo = s.materials + rd<uint32_t>(s.environment);
        E->n = lpInfo[0];
        size_t p = (size_t)(lpInfo - s.materials) + 1;
        for (int i = 0; i < E->n && i < 4; ++i) {
            p = (p + 3) & ~size_t(3);
            const uint8_t* e = s.materials + p;
            if (e[0] == 0) {
                E->L[i] = evalColor(s.textures + rd<uint32_t>(e + 4), lp.uv[0], lp.uv[1]) * rd<float>(e + 8);
                p += 16;
            } else E->L[i] = mk(0, 0, 0);
        }
        return;
    }
    E->env = false;
    const uint8_t* lpInfo = s.materials + off;
    E->n = lpInfo[0];
    E->N = lp.sNormal;
    if (lp.hasTangent) E->S = lp.sTangent;
    else makeTangent(lp.sNormal, &E->S);
    E->T = cross(E->N, E->S);
    E->NG = lp.gNormal;
    size_t p = off + 1;
    for (int i = 0; i < E->n && i < 4; ++i) {
        p = (p + 3) & ~size_t(3);
        const uint8_t* e = s.materials + p;
        if (e[0] == 0) {
            E->L[i] = evalColor(s.textures + rd<uint32_t>(e + 4), lp.uv[0], lp.uv[1]);
            p += 8;
        } else E->L[i] = mk(0, 0, 0);
    }
}
V3 Le(const EDF& E, V3 vout) {
    V3 r = mk(0, 0, 0);
    if (!E.env) {
        V3 wo = worldToLocal(E.S, E.T, E.N, vout);
        for (int i = 0; i < E.n; ++i)
            r += (wo.z > 0.0f || planted(kPlantLeTwoSided)) ? (planted(kPlantLeNoInvPi) ? E.L[i] : E.L[i] / kPi)
                                                            : mk(0, 0, 0);
    } else {
        for (int i = 0; i < E.n; ++i) r += E.L[i];
    }
    return r;
}
inline float absCosNsEDF(const EDF& E, V3 v) { return !E.env ? std::fabs(dot(E.N, v)) : 1.0f; }

void sampleLightPos(const SceneC& s, float ul, float u0, float u1, LightPoint* lp, EDF* E, float* areaPDF) {
    clr_light_info li = s.lights[sampleDiscrete1D(s.lightDist, ul, areaPDF)];
    if (planted(kPlantNoLightChoiceProb)) *areaPDF = 1.0f;
    uint32_t prop = 0xFFFFu;
    if (li.atInfinity) {
        lp->atInfinity = true;
        const uint8_t* lpInfo = s.materials + rd<uint32_t>(s.environment);
        size_t p = (size_t)(lpInfo - s.materials) + 1;
        for (uint32_t i = 0; i < lpInfo[0]; ++i) {
            p = (p + 3) & ~size_t(3);
            const uint8_t* e = s.materials + p;
            if (e[0] == 0) {
                if (i == 0) {
                    float uv[2], uvPDF;
                    sampleCC2D(s.others + rd<uint32_t>(e + 12), u0, u1, uv, &uvPDF);
                    lp->uv[0] = uv[0];
                    lp->uv[1] = uv[1];
                    float theta = uv[1] * kPi, phi = uv[0] * 2 * kPi;
                    float sinTheta = std::sin(theta);
                    lp->p = mk(-std::sin(phi) * sinTheta, std::cos(theta), std::cos(phi) * sinTheta);
                    *areaPDF *= uvPDF / (2 * kPi * kPi * sinTheta);
                    break;
                }
                p += 16;
            }
        }
    } else {
        lp->atInfinity = false;
        lp->faceID = li.reference;
        const clr_face& f = s.faces[lp->faceID];
        V3 p0 = vtx(s.vertices, f.p0), p1 = vtx(s.vertices, f.p1), p2 = vtx(s.vertices, f.p2);
        float su = std::sqrt(u0);
        float b0 = 1.0f - su, b1 = u1 * su, b2 = 1.0f - b0 - b1;
        V3 ng = cross(p1 - p0, p2 - p0);
        float area = 0.5f * length(ng);
        if (!planted(kPlantLightNoInvArea) && !planted(kPlantAreaPdfNoInvArea)) *areaPDF *= 1.0f / area;
        lp->p = b0 * p0 + b1 * p1 + b2 * p2;
        lp->gNormal = normalize(ng);
        bool hasVN = f.vn0 != 0xFFFFFFFFu && f.vn1 != 0xFFFFFFFFu && f.vn2 != 0xFFFFFFFFu;
        if (hasVN)
            lp->sNormal = normalize(b0 * vtx(s.normals, f.vn0) + b1 * vtx(s.normals, f.vn1) + b2 * vtx(s.normals, f.vn2));
        else
            lp->sNormal = lp->gNormal;
        lp->hasTangent = f.vt0 != 0xFFFFFFFFu && f.vt1 != 0xFFFFFFFFu && f.vt2 != 0xFFFFFFFFu;
        if (lp->hasTangent)
            lp->sTangent = normalize(b0 * vtx(s.tangents, f.vt0) + b1 * vtx(s.tangents, f.vt1) + b2 * vtx(s.tangents, f.vt2));
        bool hasUV = f.uv0 != 0xFFFFFFFFu && f.uv1 != 0xFFFFFFFFu && f.uv2 != 0xFFFFFFFFu;
        lp->uv[0] = lp->uv[1] = 0.0f;
        if (hasUV) {
            const float* uv0 = s.uvs + 2 * (size_t)f.uv0;
            const float* uv1 = s.uvs + 2 * (size_t)f.uv1;
            const float* uv2 = s.uvs + 2 * (size_t)f.uv2;
            lp->uv[0] = b0 * uv0[0] + b1 * uv1[0] + b2 * uv2[0];
            lp->uv[1] = b0 * uv0[1] + b1 * uv1[1] + b2 * uv2[1];
            float du0 = uv0[0] - uv2[0], dv0 = uv0[1] - uv2[1], du1 = uv1[0] - uv2[0], dv1 = uv1[1] - uv2[1];
            V3 dP0 = p0 - p2, dP1 = p1 - p2;
            float invDetUV = 1.0f / (du0 * dv1 - dv0 * du1);
            V3 uDir = invDetUV * mk(dv1 * dP0.x - dv0 * dP1.x, dv1 * dP0.y - dv0 * dP1.y, dv1 * dP0.z - dv0 * dP1.z);
            lp->uDir = hasVN ? normalize(cross(cross(lp->sNormal, uDir), lp->sNormal)) : normalize(uDir);
        }
        prop = f.lightPtr;
    }
    edfAlloc(s, prop, *lp, E);
}
float getAreaPDF(const SceneC& s, const LightPoint& lp) {
    uint32_t nItems = rd<uint32_t>(s.lightDist + 4);
    if (lp.atInfinity) {
        const uint8_t* lpInfo = s.materials + rd<uint32_t>(s.environment);
        float areaPDF = 0.0f;
        size_t p = (size_t)(lpInfo - s.materials) + 1;
        for (uint32_t i = 0; i < lpInfo[0]; ++i) {
            p = (p + 3) & ~size_t(3);
            const uint8_t* e = s.materials + p;
            if (e[0] == 0) {
                float uvPDF = pdfCC2D(s.others + rd<uint32_t>(e + 12), lp.uv[0], lp.uv[1]);
                areaPDF += uvPDF / (2 * kPi * kPi * std::sin(lp.uv[1] * kPi)) * 1.0f;
                p += 16;
            }
        }
        for (uint32_t i = 0; i < nItems; ++i)
            if (s.lights[i].atInfinity) return areaPDF * probDiscrete1D(s.lightDist, i);
    } else {
        const clr_face& f = s.faces[lp.faceID];
        V3 p0 = vtx(s.vertices, f.p0), p1 = vtx(s.vertices, f.p1), p2 = vtx(s.vertices, f.p2);
        V3 ng = cross(p1 - p0, p2 - p0);
        float areaPDF = planted(kPlantAreaPdfNoInvArea) ? 1.0f : 1.0f / (0.5f * length(ng));
        for (uint32_t i = 0; i < nItems; ++i)
            if (s.lights[i].reference == lp.faceID)
                return areaPDF * (planted(kPlantNoLightChoiceProb) ? 1.0f : probDiscrete1D(s.lightDist, i));
    }
    return 0.0f;
}
void lightPointFromSurface(const SurfacePoint& sp, LightPoint* lp) {
    lp->p = sp.p; lp->gNormal = sp.gNormal; lp->sNormal = sp.sNormal; lp->sTangent = sp.sTangent;
    lp->uDir = sp.uDir; lp->uv[0] = sp.uv[0]; lp->uv[1] = sp.uv[1]; lp->faceID = sp.faceID;
    lp->hasTangent = sp.hasTangent; lp->atInfinity = false;
}

// ---- camera (camera.cl, matrix.cl) ----
inline float m16(const uint8_t* m, int i) { return rd<float>(m + 4 * i); }
V3 mulV3(const uint8_t* m, V3 v) {
    return {m16(m, 0) * v.x + m16(m, 4) * v.y + m16(m, 8) * v.z, m16(m, 1) * v.x + m16(m, 5) * v.y + m16(m, 9) * v.z,
            m16(m, 2) * v.x + m16(m, 6) * v.y + m16(m, 10) * v.z};
}
V3 mulP3(const uint8_t* m, V3 p) {
    V3 r = {m16(m, 0) * p.x + m16(m, 4) * p.y + m16(m, 8) * p.z + m16(m, 12) * 1.0f,
            m16(m, 1) * p.x + m16(m, 5) * p.y + m16(m, 9) * p.z + m16(m, 13) * 1.0f,
            m16(m, 2) * p.x + m16(m, 6) * p.y + m16(m, 10) * p.z + m16(m, 14) * 1.0f};
    float w = m16(m, 3) * p.x + m16(m, 7) * p.y + m16(m, 11) * p.z + m16(m, 15) * 1.0f;
    if (w != 1.0f) r = r / w;
    return r;
}
V3 mulP4xyz(const uint8_t* m, float x, float y) {  // (x, y, 0, 1), xyz after the w divide
    float px = m16(m, 0) * x + m16(m, 4) * y + m16(m, 8) * 0.0f + m16(m, 12) * 1.0f;
    float py = m16(m, 1) * x + m16(m, 5) * y + m16(m, 9) * 0.0f + m16(m, 13) * 1.0f;
    float pz = m16(m, 2) * x + m16(m, 6) * y + m16(m, 10) * 0.0f + m16(m, 14) * 1.0f;
    float pw = m16(m, 3) * x + m16(m, 7) * y + m16(m, 11) * 0.0f + m16(m, 15) * 1.0f;
    if (pw != 1.0f) { px /= pw; py /= pw; pz /= pw; }
    return {px, py, pz};
}

// Camera ray for film point (px, py) in raster units and lens sample
// (lu0, lu1): sampleLensPos and sample_We (camera.cl:60-113). The lens point
// is the concentric disk sample scaled by lensRadius in the camera's z = 0
// plane; with a lens, the direction aims at the pinhole ray's point on the
// plane z = -objPDistance (the focal plane). dirPDF = 1 / (cos^3 * area of the
// image plane at unit distance).
void cameraRay(const SceneC& s, float px, float py, float lu0, float lu1, V3* org, V3* rdir, float* dirPDF) {
    const uint8_t* cam = s.camera;
    const uint8_t* l2w = cam + 64;
    const float lensRadius = rd<float>(cam + 136), objPD = rd<float>(cam + 140), vpa = rd<float>(cam + 132);
    float ux, uy;
    concentricSampleDisk(lu0, lu1, &ux, &uy);
    const V3 lensLocal = mk(ux * lensRadius, uy * lensRadius, 0.0f);
    *org = mulP3(l2w, lensLocal);
    V3 dir = normalize(mulP4xyz(cam + 192, px, py));
    if (lensRadius > 0.0f) {
        const float ft = -objPD / dir.z;
        const V3 pf = mk(0, 0, 0) + ft * dir;
        dir = normalize(pf - mk(lensLocal.x, lensLocal.y, 0.0f));
    }
    *rdir = mulV3(l2w, dir);
    *dirPDF = 1.0f / (std::pow(-dir.z, 3.0f) * vpa);
}

struct RenderArgs {
    uint32_t width, height, maxDepth;
    bool noRR;       // CLR_RENDER_NO_RR
    bool lightOnly;  // CLR_RENDER_LIGHT_SAMPLING_ONLY
    bool bsdfOnly;   // CLR_RENDER_BSDF_SAMPLING_ONLY
};
RenderArgs renderArgs(uint32_t width, uint32_t height, uint32_t maxDepth, uint32_t flags) {
    return {width, height, maxDepth, (flags & CLR_RENDER_NO_RR) != 0, (flags & CLR_RENDER_LIGHT_SAMPLING_ONLY) != 0,
            (flags & CLR_RENDER_BSDF_SAMPLING_ONLY) != 0};
}
// Weight of an emission term reached by a diffuse or glossy BSDF sample
// (pathtracer.cl:154-163; per strategy: 0 with light samples alone, 1 with
// BSDF samples alone).
inline float emissionHitMIS(const RenderArgs& ra, float dirPDF, float lightPDF) {
    if (ra.lightOnly) return 0.0f;
    if (ra.bsdfOnly) return 1.0f;
    return (dirPDF * dirPDF) / (lightPDF * lightPDF + dirPDF * dirPDF);
}

// Observer of a path's radiance terms (the adjoint's differentiates them).
struct NoHook {
    void emission(V3, uint32_t) {}
    void lightSample(V3, uint32_t, uint32_t, const BSDF&, V3, V3, const SurfacePoint&) {}
    void bounce(uint32_t, const BSDF&, V3, V3, const SurfacePoint&, V3) {}
};

// Work of an environment shadow ray: its closest-hit counts since (v0, c0)
// and the counts of an any-hit walk of the same ray.
void countAnyHit(const SceneC& s, V3 org, V3 dir, uint64_t v0, uint64_t c0) {
    ++t_work[3];
    t_work[4] += t_work[1] - v0;
    t_work[5] += t_work[2] - c0;
    const float o[3] = {org.x, org.y, org.z}, d[3] = {dir.x, dir.y, dir.z};
    auto tri = [&](uint32_t face, float bound, float* tt) {
        float x = bound;
        Isect tmp{};
        if (!triTest(s, org, dir, face, &x, &tmp)) return false;
        *tt = x;
        return true;
    };
    float t = INFINITY;
    orc_ordered::Counts cnt;
    orc_ordered::traverseWide(s.inodes, s.leaves, s.numFaces, o, d, t, tri, cnt, true, nullptr);
    t_work[6] += cnt.visits;
    t_work[7] += cnt.tests;
}

// One camera path of pathtracing (pathtracer.cl:25-197) for pixel (x, y).
template <class Hook = NoHook>
V3 tracePath(const SceneC& s, const RenderArgs& ra, uint32_t gx, uint32_t gy, uint32_t* rds, Hook& hook) {
    V3 pix = mk(0, 0, 0);
    // sampleLensPos draws the lens sample, sample_We the film jitter.
    const float cu0 = getFloat(rds), cu1 = getFloat(rds);
    const float px = (float)gx + getFloat(rds);
    const float py = (float)gy + getFloat(rds);
    V3 org, rdir;
    float dirPDF;
    cameraRay(s, px, py, cu0, cu1, &org, &rdir, &dirPDF);
    const V3 idfDir = mulV3(s.camera + 64, mk(0.0f, 0.0f, 1.0f));
    V3 alpha = mk(1, 1, 1) * (std::fabs(dot(idfDir, rdir)) / dirPDF);
    float initY = luminance(alpha);

    SurfacePoint sp;
    LightPoint lp;
    EDF E;
    BSDF B;
    bool traceContinue = true;
    uint32_t sampledType = 0;
    uint32_t depth = 0;
    V3 vout{};
    float MIS, lightPDF;
    if (rayIntersection(s, org, rdir, &sp)) {
        const clr_face* face = &s.faces[sp.faceID];
        vout = -rdir;
        if (face->lightPtr != 0xFFFFu) {
            lightPointFromSurface(sp, &lp);
            edfAlloc(s, face->lightPtr, lp, &E);
            const V3 term = alpha * Le(E, vout);
            pix += term;
            hook.emission(term, face->lightPtr);
        }
        while (true) {
            bsdfAlloc(s, face->matPtr, sp, &B);
            float dist2;
            if (hasNonSpecular(B)) {
                float lu = getFloat(rds), l0 = getFloat(rds), l1 = getFloat(rds);
                if (!ra.bsdfOnly) {  // light sample (skipped, its numbers drawn, with BSDF samples alone)
                    lp = LightPoint();
                    sampleLightPos(s, lu, l0, l1, &lp, &E, &lightPDF);
                    V3 vinL = lp.p - sp.p;
                    dist2 = dot(vinL, vinL);
                    vinL = vinL * (1.0f / std::sqrt(dist2));
                    if (lp.atInfinity) { vinL = lp.p; dist2 = 1.0f; }
                        V3 sorg = sp.p + vinL * kEps;
                    SurfacePoint lsp;
                    const uint64_t v0 = t_work[1], c0 = t_work[2];
                    bool hit = rayIntersection(s, sorg, vinL, &lsp);
                    if (lp.atInfinity && s.ordered && g_countWork) countAnyHit(s, sorg, vinL, v0, c0);
                    bool visible = lp.atInfinity ? !hit : (hit && lsp.faceID == lp.faceID);  // D4
                    if (visible && lightPDF > 0.0f && lightPDF < INFINITY) {  // D11
                        V3 vinLrev = -vinL;
                        float fsPDF = fsPdf(B, vout, vinL) * (planted(kPlantMisNoCosL) ? 1.0f : absCosNsEDF(E, vinL)) / dist2;
                        if (ra.lightOnly) MIS = 1.0f;
                        else if (planted(kPlantNeeMisSwapped)) MIS = (fsPDF * fsPDF) / (lightPDF * lightPDF + fsPDF * fsPDF);
                        else MIS = (lightPDF * lightPDF) / (lightPDF * lightPDF + fsPDF * fsPDF);
                        const float cosL = planted(kPlantNeeNoCosL) ? 1.0f : absCosNsEDF(E, vinL);
                        float w = MIS * cosL * absCosNsBSDF(B, vinL) / dist2 / lightPDF;
                        const V3 term = w * ((alpha * Le(E, vinLrev)) * fs(B, vout, vinL));
                        pix += term;
                        hook.lightSample(term, lp.atInfinity ? 0xFFFFFFFFu : (uint32_t)s.faces[lp.faceID].lightPtr,
                                         face->matPtr, B, vout, vinL, sp);
                    }
                }
            }
            traceContinue = true;
            float bc = getFloat(rds), b0 = getFloat(rds), b1 = getFloat(rds);
            V3 nd{};
            V3 f = sampleFs(B, vout, bc, b0, b1, &nd, &dirPDF, &sampledType);
            if (zeroVec(f) || dirPDF == 0.0f) { traceContinue = false; break; }
            hook.bounce(face->matPtr, B, vout, nd, sp, pix);
            alpha *= f * (absCosNsBSDF(B, nd) / dirPDF);
            rdir = nd;
            org = sp.p + rdir * kEps;
            ++depth;
            if (!rayIntersection(s, org, rdir, &sp)) break;
            face = &s.faces[sp.faceID];
            vout = -rdir;
            if (face->lightPtr != 0xFFFFu) {
                lightPointFromSurface(sp, &lp);
                edfAlloc(s, face->lightPtr, lp, &E);
                MIS = 1.0f;
                if (sampledType & BxDF_Non_Singular) {
                    V3 dd = sp.p - org;
                    float d2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z;
                    if (planted(kPlantImplicitNoDist2)) d2 = 1.0f;
                    lightPDF = getAreaPDF(s, lp) * d2 / absCosNsEDF(E, vout);
                    MIS = emissionHitMIS(ra, dirPDF, lightPDF);
                }
                const V3 term = (MIS * alpha) * Le(E, vout);
                pix += term;
                hook.emission(term, face->lightPtr);
            }
            if (depth >= ra.maxDepth) { traceContinue = false; break; }  // D6
            if (ra.noRR) continue;
            float cp = std::fmin(luminance(alpha) * (1.0f / initY), 1.0f);
            if (getFloat(rds) < cp) {
                alpha *= 1.0f / cp;
                dirPDF *= cp;
            } else {
                traceContinue = false;
                break;
            }
        }
    }
    if (traceContinue) {
        vout = -rdir;
        lp = LightPoint();
        lp.atInfinity = true;
        float theta = std::acos(clampf(rdir.y, -1.0f, 1.0f));
        float phi = std::fmod(std::atan2(-rdir.x, rdir.z) + 2.0f * kPi, 2 * kPi);
        lp.uv[0] = phi / (2.0f * kPi);
        lp.uv[1] = theta / kPi;
        edfAlloc(s, 0xFFFFu, lp, &E);
        MIS = 1.0f;
        if ((sampledType & BxDF_Non_Singular) != 0 && depth > 0) {
            lightPDF = getAreaPDF(s, lp) * 1.0f / absCosNsEDF(E, vout);
            MIS = emissionHitMIS(ra, dirPDF, lightPDF);
        }
        pix += (MIS * alpha) * Le(E, vout);
    }
    return pix;
}

// Path-replay adjoint (restates the library's): per differentiable material
// (one Oren-Nayar element) albedo and sigma, and one emitter's emission M.
// Light-sample and emission terms are differentiated where they occur; at
// each bounce from a differentiable vertex the radiance so far is recorded
// and the later terms, which all carry that vertex's sampled BSDF value, are
// L - prefix at the end of the path. Accumulates in double.
struct AdjHookC {
    const SceneC& s;
    const std::vector<int>& table;  // matPtr -> parameter index or -1
    uint32_t lightProp;
    V3 M, adj;
    double* gAlb;
    double* gSig;
    double* gEm;
    struct Rec {
        int m;
        V3 prefix, rho;
        float dlog;
    };
    std::vector<Rec> recs;

    int param(uint32_t matPtr, const BSDF& B) const {
        if (B.n != 1 || B.c[0].id != 0) return -1;
        return table[matPtr & 0xFFFFu];
    }
    float dlogSigma(uint32_t matPtr, const BSDF& B, const SurfacePoint& sp, V3 vout, V3 vin) const {
        const BxDF& f = B.c[0];
        if (f.a == 1.0f) return 0.0f;
        const uint8_t* e = s.materials + ((matPtr + 8u + 3u) & ~3u);
        const float sigma = evalFloat(s.textures + rd<uint32_t>(e + 8), sp.uv[0], sp.uv[1]);
        const V3 wo = worldToLocal(B.S, B.T, B.N, vout), wi = worldToLocal(B.S, B.T, B.N, vin);
        const float sti = sinTh(wi), sto = sinTh(wo);
        float maxCos = 0.0f;
        if (sti > 1e-4f && sto > 1e-4f) maxCos = std::fmax(0.0f, cosPhi(wi) * cosPhi(wo) + sinPhi(wi) * sinPhi(wo));
        float sinA, tanB;
        if (absCosT(wi) > absCosT(wo)) { sinA = sto; tanB = sti / absCosT(wi); }
        else { sinA = sti; tanB = sto / absCosT(wo); }
        const float X = maxCos * sinA * tanB;
        const float s2 = sigma * sigma, qa = s2 + 0.33f, qb = s2 + 0.09f;
        const float dA = -0.33f * sigma / (qa * qa), dB = 0.081f * sigma / (qb * qb);
        return (dA + dB * X) / (f.a + f.b * X);
    }
    void addAlbedo(int m, V3 term, V3 rho) {
        if (rho.x != 0.0f) gAlb[3 * m + 0] += (double)adj.x * term.x / rho.x;
        if (rho.y != 0.0f) gAlb[3 * m + 1] += (double)adj.y * term.y / rho.y;
        if (rho.z != 0.0f) gAlb[3 * m + 2] += (double)adj.z * term.z / rho.z;
    }
    void addEmission(V3 term) {
        if (M.x != 0.0f) gEm[0] += (double)adj.x * term.x / M.x;
        if (M.y != 0.0f) gEm[1] += (double)adj.y * term.y / M.y;
        if (M.z != 0.0f) gEm[2] += (double)adj.z * term.z / M.z;
    }
    void emission(V3 term, uint32_t lightPtr) {
        if (lightPtr == lightProp) addEmission(term);
    }
    void lightSample(V3 term, uint32_t lightPtr, uint32_t matPtr, const BSDF& B, V3 vout, V3 vinL,
                     const SurfacePoint& sp) {
        if (lightPtr == lightProp) addEmission(term);
        const int m = param(matPtr, B);
        if (m >= 0) {
            addAlbedo(m, term, B.c[0].R);
            gSig[m] += (double)dot(adj, term) * dlogSigma(matPtr, B, sp, vout, vinL);
        }
    }
    void bounce(uint32_t matPtr, const BSDF& B, V3 vout, V3 vin, const SurfacePoint& sp, V3 pixSoFar) {
        const int m = param(matPtr, B);
        if (m < 0) return;
        recs.push_back({m, pixSoFar, B.c[0].R, dlogSigma(matPtr, B, sp, vout, vin)});
    }
    void finish(V3 L) {
        for (const Rec& r : recs) {
            const V3 after = L - r.prefix;
            addAlbedo(r.m, after, r.rho);
            gSig[r.m] += (double)dot(adj, after) * r.dlog;
        }
        recs.clear();
    }
};

}  // namespace

extern "C" {

// Per-pixel xorshift128 states: seed = (y * width + x) ^ 0x9E3779B9, expanded
// as s_k = seed = 1812433253 * (seed ^ (seed >> 30)) + k (main.cpp:101-104,
// unsigned, no discards).
int orc_seed_states(uint32_t width, uint32_t height, uint32_t* states) {
    for (uint32_t i = 0; i < width * height; ++i) {
        uint32_t seed = i ^ 0x9E3779B9u;
        for (uint32_t k = 0; k < 4; ++k) states[4 * (size_t)i + k] = seed = 1812433253u * (seed ^ (seed >> 30)) + k;
    }
    return 0;
}

// Renders `spp` samples for pixels [pixBegin, pixEnd) (row-major index),
// accumulating into pixels (clr_float3 per pixel) and advancing states.
// arrays: vertices, normals, tangents, uvs, faces, lights, materials, textures,
// others, inodes, leaves, midpoint BVH nodes (or null). threads <= 0 uses all
// hardware threads.
// Adjoint of orc_render over pixels [0, width * height): gradients (float out)
// of sum_p dLdI[p] . (radiance added for p) / spp, states not modified.
int orc_render_adjoint(const void* const* arrays, uint32_t numFaces, uint32_t width, uint32_t height,
                       uint32_t spp, uint32_t maxDepth, uint32_t flags, uint32_t numMat,
                       const uint32_t* matOffsets, uint32_t lightProp, const uint32_t* states,
                       const clr_float3* dLdI, float* gAlb, float* gSig, float* gEm, int threads) {
    SceneC s{};
    s.vertices = (const clr_float3*)arrays[0];
    s.normals = (const clr_float3*)arrays[1];
    s.tangents = (const clr_float3*)arrays[2];
    s.uvs = (const float*)arrays[3];
    s.faces = (const clr_face*)arrays[4];
    s.lights = (const clr_light_info*)arrays[5];
    s.materials = (const uint8_t*)arrays[6];
    s.textures = (const uint8_t*)arrays[7];
    s.others = (const uint8_t*)arrays[8];
    s.inodes = (const clr_lbvh_inode*)arrays[9];
    s.leaves = (const clr_lbvh_leaf*)arrays[10];
    s.bvh = (const clr_bvh_node*)arrays[11];
    s.numFaces = numFaces;
    s.camera = s.others + rd<uint32_t>(s.others + 0);
    s.environment = s.others + rd<uint32_t>(s.others + 4);
    s.lightDist = s.others + rd<uint32_t>(s.others + 8);
    s.ordered = (flags & CLR_RENDER_REFERENCE_TRAVERSAL) == 0;
    const RenderArgs ra = renderArgs(width, height, maxDepth, flags);
    std::vector<int> table(65536, -1);
    for (uint32_t k = 0; k < numMat; ++k) table[matOffsets[k] & 0xFFFFu] = (int)k;
    V3 M = mk(0, 0, 0);
    if (lightProp != 0xFFFFFFFFu) {
        const uint8_t* e = s.materials + ((lightProp + 1u + 3u) & ~3u);
        M = evalColor(s.textures + rd<uint32_t>(e + 4), 0.0f, 0.0f);
    }
    const uint32_t numPix = width * height;
    const int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::vector<double>> part(nt, std::vector<double>(4 * (size_t)numMat + 3, 0.0));
    std::atomic<uint32_t> next{0};
    auto work = [&](int ti) {
        double* gA = part[ti].data();
        double* gS = gA + 3 * (size_t)numMat;
        double* gE = gS + numMat;
        AdjHookC hook{s, table, lightProp, M, mk(0, 0, 0), gA, gS, gE, {}};
        while (true) {
            const uint32_t b = next.fetch_add(64);
            if (b >= numPix) break;
            for (uint32_t i = b; i < std::min(numPix, b + 64); ++i) {
                uint32_t st[4] = {states[4 * (size_t)i], states[4 * (size_t)i + 1], states[4 * (size_t)i + 2],
                                  states[4 * (size_t)i + 3]};
                const float inv = 1.0f / (float)spp;
                hook.adj = mk(dLdI[i].x * inv, dLdI[i].y * inv, dLdI[i].z * inv);
                for (uint32_t k = 0; k < spp; ++k) hook.finish(tracePath(s, ra, i % width, i / width, st, hook));
            }
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < nt; ++t) ts.emplace_back(work, t);
    for (auto& t : ts) t.join();
    for (uint32_t k = 0; k < 4 * numMat + 3; ++k) {
        double v = 0.0;
        for (int t = 0; t < nt; ++t) v += part[t][k];
        if (k < 3 * numMat) gAlb[k] = (float)v;
        else if (k < 4 * numMat) gSig[k - 3 * numMat] = (float)v;
        else if (gEm) gEm[k - 4 * numMat] = (float)v;
    }
    return 0;
}

// The ray from the lens centre (lens sample 0.5, 0.5) through the centre of
// pixel (gx, gy) (sampleLensPos, sample_We; camera.cl:60-113).
void centreRay(const SceneC& s, uint32_t gx, uint32_t gy, V3* org, V3* rdir) {
    float dirPDF;
    cameraRay(s, gx + 0.5f, gy + 0.5f, 0.5f, 0.5f, org, rdir, &dirPDF);
}

// Those rays for every pixel as float[7][W * H] SoA planes (tmax = inf), so
// tests can trace them with the traversal restatement.
int orc_centre_rays(const void* const* arrays, uint32_t width, uint32_t height, float* rays) {
    SceneC s{};
    s.others = (const uint8_t*)arrays[8];
    s.camera = s.others + rd<uint32_t>(s.others + 0);
    const size_t R = (size_t)width * height;
    for (uint32_t gy = 0; gy < height; ++gy)
        for (uint32_t gx = 0; gx < width; ++gx) {
            V3 o, d;
            centreRay(s, gx, gy, &o, &d);
            const size_t i = (size_t)gy * width + gx;
            rays[i] = o.x;
            rays[R + i] = o.y;
            rays[2 * R + i] = o.z;
            rays[3 * R + i] = d.x;
            rays[4 * R + i] = d.y;
            rays[5 * R + i] = d.z;
            rays[6 * R + i] = INFINITY;
        }
    return 0;
}

// bvh_visualizer (bvh_visualizer.cl:25-84) over every pixel: the ray from
// the lens centre (sample 0.5, 0.5) through the pixel centre; pixel += the
// base colour (getBaseColor_fs, reflection.cl:730-772) of the surface hit,
// then += 0.025 * box hits (rayIntersection's count).
int orc_bvh_visualizer(const void* const* arrays, uint32_t numFaces, uint32_t width, uint32_t height,
                       uint32_t flags, clr_float3* pixels) {
    SceneC s{};
    s.ordered = (flags & CLR_RENDER_REFERENCE_TRAVERSAL) == 0;
    s.vertices = (const clr_float3*)arrays[0];
    s.normals = (const clr_float3*)arrays[1];
    s.tangents = (const clr_float3*)arrays[2];
    s.uvs = (const float*)arrays[3];
    s.faces = (const clr_face*)arrays[4];
    s.lights = (const clr_light_info*)arrays[5];
    s.materials = (const uint8_t*)arrays[6];
    s.textures = (const uint8_t*)arrays[7];
    s.others = (const uint8_t*)arrays[8];
    s.inodes = (const clr_lbvh_inode*)arrays[9];
    s.leaves = (const clr_lbvh_leaf*)arrays[10];
    s.bvh = (const clr_bvh_node*)arrays[11];
    s.numFaces = numFaces;
    s.camera = s.others + rd<uint32_t>(s.others + 0);
    s.environment = s.others + rd<uint32_t>(s.others + 4);
    s.lightDist = s.others + rd<uint32_t>(s.others + 8);
    for (uint32_t gy = 0; gy < height; ++gy)
        for (uint32_t gx = 0; gx < width; ++gx) {
            V3 org, rdir;
            centreRay(s, gx, gy, &org, &rdir);
            SurfacePoint sp;
            uint32_t hits = 0;
            clr_float3& p = pixels[(size_t)gy * width + gx];
            if (rayIntersection(s, org, rdir, &sp, &hits)) {
                BSDF B;
                bsdfAlloc(s, s.faces[sp.faceID].matPtr, sp, &B);
                V3 c = mk(0, 0, 0);
                for (int i = 0; i < B.n; ++i) {
                    const BxDF& f = B.c[i];
                    if (f.id == 0 || f.id == 1 || f.id == 3 || f.id == 4) c = c + f.R;
                    else if (f.id == 5) c = c + f.Rs;  // Ashikhmin diffuse: Rd
                }
                p.x += c.x;
                p.y += c.y;
                p.z += c.z;
            }
            const float k = (float)hits * 0.025f;
            p.x += k;
            p.y += k;
            p.z += k;
        }
    return 0;
}

int orc_render(const void* const* arrays, uint32_t numFaces, uint32_t width, uint32_t height, uint32_t spp,
               uint32_t maxDepth, uint32_t* states, clr_float3* pixels, uint32_t pixBegin, uint32_t pixEnd,
               int threads, uint32_t flags) {
    SceneC s{};
    s.ordered = (flags & CLR_RENDER_REFERENCE_TRAVERSAL) == 0;
    s.vertices = (const clr_float3*)arrays[0];
    s.normals = (const clr_float3*)arrays[1];
    s.tangents = (const clr_float3*)arrays[2];
    s.uvs = (const float*)arrays[3];
    s.faces = (const clr_face*)arrays[4];
    s.lights = (const clr_light_info*)arrays[5];
    s.materials = (const uint8_t*)arrays[6];
    s.textures = (const uint8_t*)arrays[7];
    s.others = (const uint8_t*)arrays[8];
    s.inodes = (const clr_lbvh_inode*)arrays[9];
    s.leaves = (const clr_lbvh_leaf*)arrays[10];
    s.bvh = (const clr_bvh_node*)arrays[11];
    s.numFaces = numFaces;
    s.camera = s.others + rd<uint32_t>(s.others + 0);
    s.environment = s.others + rd<uint32_t>(s.others + 4);
    s.lightDist = s.others + rd<uint32_t>(s.others + 8);
    const RenderArgs ra = renderArgs(width, height, maxDepth, flags);
    if (pixEnd > width * height) pixEnd = width * height;
    int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{pixBegin};
    auto work = [&]() {
        while (true) {
            uint32_t b = next.fetch_add(64);
            if (b >= pixEnd) break;
            uint32_t e = std::min(pixEnd, b + 64);
            for (uint32_t i = b; i < e; ++i) {
                uint32_t gx = i % width, gy = i / width;
                // Radiance of the call's samples summed in order, then added to
                // the pixel once (the reference adds each contribution to the
                // pixel as it goes, pathtracer.cl:93-195; only the rounding of
                // the running sum differs).
                V3 acc = mk(0, 0, 0);
                NoHook hook;
                for (uint32_t k = 0; k < spp; ++k) acc += tracePath(s, ra, gx, gy, states + 4 * (size_t)i, hook);
                pixels[i].x += acc.x;
                pixels[i].y += acc.y;
                pixels[i].z += acc.z;
            }
        }
        flushWork();
    };
    if (nt == 1) work();
    else {
        std::vector<std::thread> ts;
        for (int t = 0; t < nt; ++t) ts.emplace_back(work);
        for (auto& t : ts) t.join();
    }
    return 0;
}

// Closed-form probes of the BSDF restatement for the pin tests:
// evaluates fx / fx_pdf / sample_fx of a single BxDF given in the local
// frame. kind: 0 Oren-Nayar diffuse (p0 = sigma), 3 Ward (p0 = ax, p1 = ay),
// 4 Ashikhmin-Shirley specular (p0 = nu, p1 = nv), 5 Ashikhmin diffuse,
// 6 Beckmann (p0 = alpha, p1 = eta; dielectric Fresnel, eta 0 = no Fresnel).
// out[0..2] = f, out[3] = pdf, out[4..6] = sampled wi, out[7] = sampled pdf,
// out[8..10] = sampled f.
int orc_bxdf_probe(int kind, float p0, float p1, const float* R, const float* wo3, const float* wi3, float u0,
                   float u1, float* out) {
    BxDF f;
    f.id = (uint8_t)kind;
    f.R = mk(R[0], R[1], R[2]);
    f.Rs = mk(R[0], R[1], R[2]);
    uint8_t fres[12] = {0};
    if (kind == 0) {
        float s2 = p0 * p0;
        f.a = 1.0f - (s2 / (2.0f * (s2 + 0.33f)));
        f.b = 0.45f * s2 / (s2 + 0.09f);
        f.type = BxDF_Reflection | BxDF_Diffuse;
    } else {
        f.a = p0;
        f.b = p1;
        f.type = BxDF_Reflection | BxDF_Glossy;
        if (kind == 5) f.type = BxDF_Reflection | BxDF_Diffuse;
        if (kind == 6) {
            fres[0] = p1 > 0.0f ? 2 : 0;
            float one = 1.0f;
            std::memcpy(fres + 4, &one, 4);
            std::memcpy(fres + 8, &p1, 4);
        }
    }
    f.fresnel = fres;
    V3 wo = mk(wo3[0], wo3[1], wo3[2]), wi = mk(wi3[0], wi3[1], wi3[2]);
    V3 v = fx(f, wo, wi);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
    out[3] = fxPdf(f, wo, wi);
    V3 swi{};
    float spdf = 0.0f;
    V3 sv = sampleFx(f, wo, u0, u1, &swi, &spdf);
    out[4] = swi.x; out[5] = swi.y; out[6] = swi.z;
    out[7] = spdf;
    out[8] = sv.x; out[9] = sv.y; out[10] = sv.z;
    return 0;
}

int orc_fresnel_probe(int type, const float* eta, const float* k, float etaExt, float etaInt, float cosi,
                      float* out) {
    uint8_t buf[48] = {0};
    buf[0] = (uint8_t)type;
    if (type == 1) {
        std::memcpy(buf + 16, eta, 12);
        std::memcpy(buf + 32, k, 12);
    } else if (type == 2) {
        std::memcpy(buf + 4, &etaExt, 4);
        std::memcpy(buf + 8, &etaInt, 4);
    }
    V3 r = evalFresnel(buf, cosi);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
    return 0;
}

// Camera probe: origin (3), direction (3) and direction pdf of the camera
// ray for film point (px, py) and lens sample (lu0, lu1) of the scene's camera.
int orc_camera_ray(const void* const* arrays, float px, float py, float lu0, float lu1, float* out) {
    SceneC s{};
    s.others = (const uint8_t*)arrays[8];
    s.camera = s.others + rd<uint32_t>(s.others + 0);
    V3 o, d;
    float pdf;
    cameraRay(s, px, py, lu0, lu1, &o, &d, &pdf);
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
    out[3] = d.x; out[4] = d.y; out[5] = d.z;
    out[6] = pdf;
    return 0;
}

// Normal-map probe: shading normal N, tangent S and T = N x S for a surface
// frame (sNormal, uDir, tangent S) and a texel colour.
int orc_normal_map_probe(const float* sNormal, const float* uDir, const float* S, const float* tex, float* out) {
    V3 N{}, S1 = mk(S[0], S[1], S[2]);
    normalMapFrame(mk(sNormal[0], sNormal[1], sNormal[2]), mk(uDir[0], uDir[1], uDir[2]),
                   mk(tex[0], tex[1], tex[2]), &N, &S1);
    const V3 T = cross(N, S1);
    out[0] = N.x; out[1] = N.y; out[2] = N.z;
    out[3] = S1.x; out[4] = S1.y; out[5] = S1.z;
    out[6] = T.x; out[7] = T.y; out[8] = T.z;
    return 0;
}

// Texture probe: evalColor of one texture record (texture.cl) at (u, v).
int orc_texture_probe(const uint8_t* record, float u, float v, float* out) {
    const V3 c = evalColor(record, u, v);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
    return 0;
}

int orc_disk_probe(float u1, float u2, float* out) {
    concentricSampleDisk(u1, u2, out, out + 1);
    return 0;
}

// Traversal work of the renders since the last reset (the kWorkWords
// counters above), and whether later renders count the any-hit walks of
// environment shadow rays (count != 0).
int orc_work_counters(uint64_t* out, int reset, int count) {
    for (int k = 0; k < kWorkWords; ++k) {
        if (out) out[k] = g_work[k];
        if (reset) g_work[k] = 0;
    }
    g_countWork = count != 0;
    return 0;
}

// Plants mistake k (the Plant list above) in the restatement until the next
// call; 0 removes it. Tests use it to show that a pin detects a mistake.
int orc_set_plant(int k) {
    g_plant = k;
    return 0;
}

// Camera probe over many film points: film[4i..4i+3] = (px, py, lu0, lu1);
// out[7i..7i+6] = origin, direction, direction pdf (orc_camera_ray), and
// out[7n..7n+2] = the camera's +z axis in world space (idfDir of tracePath).
int orc_camera_rays(const void* const* arrays, uint32_t n, const float* film, float* out) {
    SceneC s{};
    s.others = (const uint8_t*)arrays[8];
    s.camera = s.others + rd<uint32_t>(s.others + 0);
    for (uint32_t i = 0; i < n; ++i) {
        V3 o, d;
        float pdf;
        cameraRay(s, film[4 * i], film[4 * i + 1], film[4 * i + 2], film[4 * i + 3], &o, &d, &pdf);
        float* r = out + 7 * (size_t)i;
        r[0] = o.x; r[1] = o.y; r[2] = o.z;
        r[3] = d.x; r[4] = d.y; r[5] = d.z;
        r[6] = pdf;
    }
    const V3 z = mulV3(s.camera + 64, mk(0.0f, 0.0f, 1.0f));
    out[7 * (size_t)n] = z.x; out[7 * (size_t)n + 1] = z.y; out[7 * (size_t)n + 2] = z.z;
    return 0;
}

// BSDF mixture probe: the material at byte offset matOffset of the scene's
// materialsData, allocated (bsdfAlloc) at a surface point with geometric and
// shading normal +z, uv (0, 0) and no tangent. Per i (directions in world
// space): out[12i..] = fs(wo_i, wi_i) (3), fs_pdf(wo_i, wi_i), then
// sample_fs(wo_i, uc, u0, u1 = u[3i..3i+2]): direction (3), pdf, value (3)
// and the sampled element's type bits.
int orc_bsdf_probe(const void* const* arrays, uint32_t matOffset, uint32_t n, const float* wo, const float* wi,
                   const float* u, float* out) {
    SceneC s{};
    s.materials = (const uint8_t*)arrays[6];
    s.textures = (const uint8_t*)arrays[7];
    s.others = (const uint8_t*)arrays[8];
    SurfacePoint sp{};
    sp.p = mk(0, 0, 0);
    sp.gNormal = sp.sNormal = mk(0, 0, 1);
    sp.uDir = mk(1, 0, 0);
    sp.hasTangent = false;
    BSDF B;
    bsdfAlloc(s, matOffset, sp, &B);
    for (uint32_t i = 0; i < n; ++i) {
        const V3 vo = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), vi = mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        float* r = out + 12 * (size_t)i;
        const V3 f = fs(B, vo, vi);
        r[0] = f.x; r[1] = f.y; r[2] = f.z;
        r[3] = fsPdf(B, vo, vi);
        V3 sd{};
        float spdf = 0.0f;
        uint32_t type = 0;
        const V3 sf = sampleFs(B, vo, u[3 * i], u[3 * i + 1], u[3 * i + 2], &sd, &spdf, &type);
        r[4] = sd.x; r[5] = sd.y; r[6] = sd.z;
        r[7] = spdf;
        r[8] = sf.x; r[9] = sf.y; r[10] = sf.z;
        r[11] = (float)type;
    }
    return 0;
}

}  // extern "C"
