// Device-side ray/box, ray/triangle and LBVH closest-hit routines shared by
// the trace entry point and the renderers.
//
// Arithmetic is kept in the reference's operation order with no contraction
// (the library is built with -ffp-contract=off) so that results equal the CPU
// oracle's bit for bit: dot/cross left to right, IEEE division.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/clear_amd.h"

namespace clr {
namespace dev {

constexpr int kStackDepth = 64;
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float omax(float a, float b) { return (b > a || a != a) ? b : a; }
__device__ __forceinline__ float omin(float a, float b) { return (b < a || a != a) ? b : a; }

__device__ __forceinline__ V3 loadV3(const clr_float3* p, uint32_t i) {
    const float4 v = *reinterpret_cast<const float4*>(p + i);
    return {v.x, v.y, v.z};
}

struct GeomScene {
    const clr_lbvh_inode* inodes;
    const clr_lbvh_leaf* leaves;
    uint32_t n;
    const clr_float3* verts;
    const clr_face* faces;
};

struct TraceStats {
    uint32_t visits, tests;
};

// rayAABBIntersection (bvh_traversal.cl:160-179): per axis the candidate plane
// on the origin's side, t = (plane - org) / dir (-1 when inside or dir == 0),
// maxt = max t; reject if maxt < 0 or > tBound; the entry point must lie in the
// box on every axis (or that axis produced maxt).
__device__ __forceinline__ bool boxTest(float4 mn, float4 mx, V3 o, V3 d, float tBound) {
    const int lx = o.x < mn.x ? -1 : (o.x > mx.x ? 1 : 0);
    const int ly = o.y < mn.y ? -1 : (o.y > mx.y ? 1 : 0);
    const int lz = o.z < mn.z ? -1 : (o.z > mx.z ? 1 : 0);
    if ((lx | ly | lz) == 0) return true;
    const float px = o.x < mn.x ? mn.x : (o.x > mx.x ? mx.x : 0.0f);
    const float py = o.y < mn.y ? mn.y : (o.y > mx.y ? mx.y : 0.0f);
    const float pz = o.z < mn.z ? mn.z : (o.z > mx.z ? mx.z : 0.0f);
    const float tx = (lx != 0 && d.x != 0.0f) ? ((px - o.x) / d.x) : -1.0f;
    const float ty = (ly != 0 && d.y != 0.0f) ? ((py - o.y) / d.y) : -1.0f;
    const float tz = (lz != 0 && d.z != 0.0f) ? ((pz - o.z) / d.z) : -1.0f;
    const float maxt = omax(tx, omax(ty, tz));
    if (maxt < 0.0f || maxt > tBound) return false;
    const float cx = o.x + maxt * d.x;
    const float cy = o.y + maxt * d.y;
    const float cz = o.z + maxt * d.z;
    const bool hx = (cx >= mn.x && cx <= mx.x) || tx == maxt;
    const bool hy = (cy >= mn.y && cy <= mx.y) || ty == maxt;
    const bool hz = (cz >= mn.z && cz <= mx.z) || tz == maxt;
    return hx && hy && hz;
}

// Geometric part of rayTriangleIntersection (bvh_traversal.cl:54-92) with a
// 32-bit face index. On acceptance returns true and sets t, b1, b2.
__device__ __forceinline__ bool triTest(const GeomScene& g, V3 o, V3 d, uint32_t faceIdx, float& t,
                                        float& b1Out, float& b2Out) {
    const uint32_t* f = reinterpret_cast<const uint32_t*>(g.faces + faceIdx);
    const uint2 f01 = *reinterpret_cast<const uint2*>(f);
    const uint32_t f2 = f[2];
    const V3 p0 = loadV3(g.verts, f01.x);
    const V3 p1 = loadV3(g.verts, f01.y);
    const V3 p2 = loadV3(g.verts, f2);
    const V3 edge01 = p1 - p0;
    const V3 edge02 = p2 - p0;
    const V3 p = cross(d, edge02);
    const float det = dot(edge01, p);
    if (det > -0.000001f && det < 0.000001f) return false;
    const float invDet = 1.0f / det;
    const V3 dd = o - p0;
    const float b1 = dot(dd, p) * invDet;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    const V3 q = cross(dd, edge01);
    const float b2 = dot(d, q) * invDet;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
    const float tt = dot(edge02, q) * invDet;
    if (tt < 0.0f || tt > t) return false;
    t = tt;
    b1Out = b1;
    b2Out = b2;
    return true;
}

// Alpha rejection hook: NoAlpha accepts every geometric hit.
struct NoAlpha {
    __device__ __forceinline__ bool reject(uint32_t, float, float) const { return false; }
};

// Traversal stacks. The reference keeps a private uint[64] per work item
// (bvh_traversal.cl:185). Here slot k of a lane lives in LDS at lds[k * stride]
// ([depth][lane], conflict-free when lanes sit at equal depth); SpillStack
// keeps only the first L slots in LDS and the rest in a per-lane global
// region (slot k >= L at spill[(k - L) * spillStride]), which trades a rare
// global access for 64 / L times more waves per CU.
struct LdsStack {
    uint32_t* lds;
    int stride;
    __device__ __forceinline__ void put(int k, uint32_t v) const { lds[k * stride] = v; }
    __device__ __forceinline__ uint32_t get(int k) const { return lds[k * stride]; }
};
template <int L>
struct SpillStack {
    uint32_t* lds;
    int stride;
    uint32_t* spill;
    size_t spillStride;
    __device__ __forceinline__ void put(int k, uint32_t v) const {
        if (k < L) lds[k * stride] = v;
        else spill[(size_t)(k - L) * spillStride] = v;
    }
    __device__ __forceinline__ uint32_t get(int k) const {
        return k < L ? lds[k * stride] : spill[(size_t)(k - L) * spillStride];
    }
};

// rayIntersection (bvh_traversal.cl:182-215). Returns the face or kNone; t,
// b1, b2 describe the accepted hit. An alpha-rejected triangle leaves t
// unchanged (the reference stores t before its alpha test, :87 vs :102-105,
// which lets a masked-out triangle occlude; that is not reproduced). A push
// beyond 64 entries (the reference's array bound) is dropped and reported
// through *overflow.
template <bool STATS, class Alpha = NoAlpha, class Stack = LdsStack>
__device__ __forceinline__ uint32_t closestHit(const GeomScene& g, V3 o, V3 d, float& t,
                                               const Stack& stack, TraceStats& st,
                                               float* b1Out = nullptr, float* b2Out = nullptr,
                                               const Alpha& alpha = Alpha(),
                                               bool* overflow = nullptr) {
    uint32_t best = kNone;
    float bb1 = 0.0f, bb2 = 0.0f;
    auto tryLeaf = [&](uint32_t leafIdx) {
        const float4* lp = reinterpret_cast<const float4*>(g.leaves + leafIdx);
        const float4 mn = lp[0], mx = lp[1];
        if (!boxTest(mn, mx, o, d, t)) return;
        const uint32_t obj = reinterpret_cast<const uint32_t*>(lp + 2)[0];
        if constexpr (STATS) ++st.tests;
        float tt = t, b1, b2;
        if (!triTest(g, o, d, obj, tt, b1, b2)) return;
        if (alpha.reject(obj, b1, b2)) return;
        t = tt;
        best = obj;
        bb1 = b1;
        bb2 = b2;
    };
    if (g.n == 1) {
        tryLeaf(0);
    } else if (g.n > 1) {
        int depth = 0;
        stack.put(0, 0u);
        depth = 1;
        bool over = false;
        while (depth > 0) {
            --depth;
            const uint32_t idx = stack.get(depth);
            const float4* np = reinterpret_cast<const float4*>(g.inodes + idx);
            const float4 mn = np[0], mx = np[1];
            if (!boxTest(mn, mx, o, d, t)) continue;
            const uint4 c = reinterpret_cast<const uint4*>(np)[2];
            if constexpr (STATS) ++st.visits;
            if (c.x & 0xFFu) tryLeaf(c.y);
            else if (depth < kStackDepth) stack.put(depth++, c.y);
            else over = true;
            if (c.x & 0xFF00u) tryLeaf(c.z);
            else if (depth < kStackDepth) stack.put(depth++, c.z);
            else over = true;
        }
        if (overflow && over) *overflow = true;
    }
    if (b1Out) *b1Out = bb1;
    if (b2Out) *b2Out = bb2;
    return best;
}

// Occlusion query with the same visiting order and tests as closestHit with
// t = +inf, stopped at the first accepted triangle with a finite distance:
// true exactly when closestHit would report a hit.
template <class Alpha = NoAlpha, class Stack = LdsStack>
__device__ __forceinline__ bool anyHit(const GeomScene& g, V3 o, V3 d, const Stack& stack,
                                       const Alpha& alpha = Alpha(), bool* overflow = nullptr) {
    float t = INFINITY;
    auto tryLeaf = [&](uint32_t leafIdx) -> bool {
        const float4* lp = reinterpret_cast<const float4*>(g.leaves + leafIdx);
        const float4 mn = lp[0], mx = lp[1];
        if (!boxTest(mn, mx, o, d, t)) return false;
        const uint32_t obj = reinterpret_cast<const uint32_t*>(lp + 2)[0];
        float tt = t, b1, b2;
        if (!triTest(g, o, d, obj, tt, b1, b2)) return false;
        if (alpha.reject(obj, b1, b2)) return false;
        return !isinf(tt);
    };
    if (g.n == 1) return tryLeaf(0);
    if (g.n == 0) return false;
    int depth = 0;
    stack.put(0, 0u);
    depth = 1;
    bool over = false, hit = false;
    while (depth > 0 && !hit) {
        --depth;
        const uint32_t idx = stack.get(depth);
        const float4* np = reinterpret_cast<const float4*>(g.inodes + idx);
        const float4 mn = np[0], mx = np[1];
        if (!boxTest(mn, mx, o, d, t)) continue;
        const uint4 c = reinterpret_cast<const uint4*>(np)[2];
        if (c.x & 0xFFu) hit = tryLeaf(c.y);
        else if (depth < kStackDepth) stack.put(depth++, c.y);
        else over = true;
        if (hit) break;
        if (c.x & 0xFF00u) hit = tryLeaf(c.z);
        else if (depth < kStackDepth) stack.put(depth++, c.z);
        else over = true;
    }
    if (overflow && over) *overflow = true;
    return hit;
}

// ---- ordered traversal -------------------------------------------------------
// Nearer child first with a padded slab test; ties in t go to the smaller face
// index, so the hit is the lexicographic minimum of (t, face) over accepted
// triangles whatever the visiting order. Per ray: reciprocal direction with
// components below 1e-20 in magnitude replaced by +-1e-20; per axis
// t0 = (min - o) * inv, t1 = (max - o) * inv; tmin = max(min(t0, t1) over axes,
// 0), tmax = min(max(t0, t1) over axes); accepted when e = tmin * (1 - 2^-16)
// <= min(tmax * (1 + 2^-16), t), with e kept as the entry distance and a miss
// returned as NaN (every later "e <= t" fails, also for t = +inf). At a node
// whose box passed, leaf children are tested nearer first (child 0 on equal
// entries), then the nearer internal child is entered and the farther pushed;
// a popped node's box is tested again against the current t.
constexpr float kPadDown = 1.0f - 1.0f / 65536.0f;
constexpr float kPadUp = 1.0f + 1.0f / 65536.0f;

__device__ __forceinline__ float safeInv(float d) {
    return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d));
}
__device__ __forceinline__ float slabEnter(float4 mn, float4 mx, V3 o, V3 inv, float tBound) {
    const float x0 = (mn.x - o.x) * inv.x, x1 = (mx.x - o.x) * inv.x;
    const float y0 = (mn.y - o.y) * inv.y, y1 = (mx.y - o.y) * inv.y;
    const float z0 = (mn.z - o.z) * inv.z, z1 = (mx.z - o.z) * inv.z;
    const float tmin = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
    const float tmax = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * kPadUp;
    const float e = tmin * kPadDown;
    return e <= fminf(tmax, tBound) ? e : __builtin_nanf("");
}

// ANY = true: stop at the first accepted triangle with a finite distance
// (occlusion); the return value then only tells hit / miss.
template <bool STATS, bool ANY, class Alpha = NoAlpha, class Stack = LdsStack>
__device__ __forceinline__ uint32_t orderedHit(const GeomScene& g, V3 o, V3 d, float& t, const Stack& stack,
                                               TraceStats& st, float* b1Out = nullptr, float* b2Out = nullptr,
                                               const Alpha& alpha = Alpha(), bool* overflow = nullptr) {
    const V3 inv = {safeInv(d.x), safeInv(d.y), safeInv(d.z)};
    uint32_t best = kNone;
    float bb1 = 0.0f, bb2 = 0.0f;
    bool done = false;
    auto leaf = [&](uint32_t li) {
        const uint32_t obj = reinterpret_cast<const uint32_t*>(g.leaves + li)[8];
        if constexpr (STATS) ++st.tests;
        float tt = t, b1, b2;
        if (!triTest(g, o, d, obj, tt, b1, b2)) return;
        if (tt == t && obj > best) return;
        if (alpha.reject(obj, b1, b2)) return;
        t = tt;
        best = obj;
        bb1 = b1;
        bb2 = b2;
        if (ANY && !isinf(tt)) done = true;
    };
    auto box = [&](bool isLeaf, uint32_t idx, float4& mn, float4& mx) {
        const float4* p = isLeaf ? reinterpret_cast<const float4*>(g.leaves + idx)
                                 : reinterpret_cast<const float4*>(g.inodes + idx);
        mn = p[0];
        mx = p[1];
    };
    if (g.n == 1) {
        float4 mn, mx;
        box(true, 0, mn, mx);
        if (slabEnter(mn, mx, o, inv, t) <= t) leaf(0);
    } else if (g.n > 1) {
        float4 mn, mx;
        box(false, 0, mn, mx);
        if (slabEnter(mn, mx, o, inv, t) <= t) {
            int depth = 0;
            uint32_t node = 0;
            bool over = false;
            while (true) {
                if constexpr (STATS) ++st.visits;
                const uint4 c = reinterpret_cast<const uint4*>(g.inodes + node)[2];
                const bool l0 = (c.x & 0xFFu) != 0, l1 = (c.x & 0xFF00u) != 0;
                float4 mn0, mx0, mn1, mx1;
                box(l0, c.y, mn0, mx0);
                box(l1, c.z, mn1, mx1);
                const float e0 = slabEnter(mn0, mx0, o, inv, t), e1 = slabEnter(mn1, mx1, o, inv, t);
                const bool first1 = e1 < e0;
                if (l0 || l1) {
                    const uint32_t ca = first1 ? c.z : c.y, cb = first1 ? c.y : c.z;
                    const bool la = first1 ? l1 : l0, lb = first1 ? l0 : l1;
                    const float ea = first1 ? e1 : e0, eb = first1 ? e0 : e1;
                    if (la && ea <= t) leaf(ca);
                    if (!done && lb && eb <= t) leaf(cb);
                    if (done) break;
                }
                const bool go0 = !l0 && e0 <= t, go1 = !l1 && e1 <= t;
                if (go0 && go1) {
                    const uint32_t nearC = first1 ? c.z : c.y, farC = first1 ? c.y : c.z;
                    if (depth < kStackDepth) stack.put(depth++, farC);
                    else over = true;
                    node = nearC;
                    continue;
                }
                if (go0 || go1) {
                    node = go0 ? c.y : c.z;
                    continue;
                }
                bool found = false;
                while (depth > 0) {
                    const uint32_t cand = stack.get(--depth);
                    box(false, cand, mn, mx);
                    if (slabEnter(mn, mx, o, inv, t) <= t) {
                        node = cand;
                        found = true;
                        break;
                    }
                }
                if (!found) break;
            }
            if (overflow && over) *overflow = true;
        }
    }
    if (b1Out) *b1Out = bb1;
    if (b2Out) *b2Out = bb2;
    return best;
}

}  // namespace dev
}  // namespace clr
