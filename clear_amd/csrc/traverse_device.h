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

// Traversal layout derived from the drop-in nodes for the ordered walk:
// per internal node i, 4 x float4 = {min of child 0 | ref0, max of child 0 |
// ref1, min of child 1, max of child 1} (refs in the .w lanes as bits: an
// internal child's index, or a leaf child's face index | 0x80000000), and per
// face its three vertices as float4s. One 64-B load per visited node and one
// 48-B load per triangle test, where the drop-in layout needs two dependent
// round trips for each.
struct GeomScene {
    const clr_lbvh_inode* inodes;
    const clr_lbvh_leaf* leaves;
    uint32_t n;
    const clr_float3* verts;
    const clr_face* faces;
    const float4* anodes = nullptr;  // [n - 1][4] (LBVH) or [2n - 1][4] by node (midpoint BVH)
    const float4* atris = nullptr;   // [n][3]
    // CPU midpoint BVH (clr_bvh_node[2n - 1], 3 float4 each). When set, the
    // reference-order walks use it instead of inodes / leaves, and the
    // traversal layout is indexed by its node numbers.
    const float4* bvh = nullptr;
    // 4-wide records for the renderers' walk ([slots][8], slots as anodes):
    // record i holds the up to four grandchildren of internal node i (the
    // children of each internal child, a leaf child itself), as SoA
    // float4s: lo.x, lo.y, lo.z, hi.x, hi.y, hi.z of slots 0..3, then the
    // four references (face | kLeafRef, node, or kNone for an empty slot
    // with an empty box). Unused when null.
    const float4* wnodes = nullptr;
};
constexpr uint32_t kLeafRef = 0x80000000u;
static_assert(sizeof(clr_lbvh_inode) == 48 && sizeof(clr_lbvh_leaf) == 48 && sizeof(clr_bvh_node) == 48,
              "node records are three float4 with the box first");

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

// rayIntersection over the CPU midpoint BVH (bvh_traversal.cl:256-285): pop a
// node, test its box against the current t (boxTest); an internal node pushes
// children[1] then children[0], a leaf (children[1] == 0xFFFFFFFF) tests
// triangle children[0], replacing the hit when tt <= t. ANY: stop at the
// first accepted triangle with a finite distance. Child or face indices out
// of range, more pops than nodes, or a push beyond 64 entries stop the walk
// or drop the push and set *overflow.
template <bool STATS, bool ANY, class Alpha, class Stack>
__device__ __forceinline__ uint32_t bvhWalk(const GeomScene& g, V3 o, V3 d, float& t, const Stack& stack,
                                             TraceStats& st, float* b1Out, float* b2Out, const Alpha& alpha,
                                             bool* overflow) {
    uint32_t best = kNone;
    float bb1 = 0.0f, bb2 = 0.0f;
    const uint32_t numNodes = 2u * g.n - 1u;
    int depth = 0;
    bool over = false;
    if (g.n) {
        stack.put(0, 0u);
        depth = 1;
    }
    uint32_t iters = 0;
    while (depth > 0) {
        if (++iters > numNodes + 1u) {
            over = true;
            break;
        }
        const uint32_t idx = stack.get(--depth);
        const float4* np = g.bvh + 3 * (size_t)idx;
        const float4 mn = np[0], mx = np[1];
        if (!boxTest(mn, mx, o, d, t)) continue;
        const uint2 c = *reinterpret_cast<const uint2*>(np + 2);
        if (c.y != kNone) {
            if constexpr (STATS) ++st.visits;
            if (depth + 2 <= kStackDepth && c.x < numNodes && c.y < numNodes) {
                stack.put(depth++, c.y);
                stack.put(depth++, c.x);
            } else {
                over = true;
            }
        } else {
            if constexpr (STATS) ++st.tests;
            if (c.x >= g.n) {
                over = true;
                continue;
            }
            float tt = t, b1, b2;
            if (!triTest(g, o, d, c.x, tt, b1, b2)) continue;
            if (alpha.reject(c.x, b1, b2)) continue;
            t = tt;
            best = c.x;
            bb1 = b1;
            bb2 = b2;
            if (ANY && !isinf(tt)) break;
        }
    }
    if (overflow && over) *overflow = true;
    if (b1Out) *b1Out = bb1;
    if (b2Out) *b2Out = bb2;
    return best;
}

// rayIntersection (bvh_traversal.cl:182-215). Returns the face or kNone; t,
// b1, b2 describe the accepted hit. An alpha-rejected triangle leaves t
// unchanged (the reference stores t before its alpha test, :87 vs :102-105,
// which lets a masked-out triangle occlude; that is not reproduced). A push
// beyond 64 entries (the reference's array bound) is dropped and reported
// through *overflow.
template <bool STATS, class Alpha = NoAlpha, class Stack = LdsStack>
__device__ __forceinline__ uint32_t lbvhClosest(const GeomScene& g, V3 o, V3 d, float& t,
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
        uint32_t iters = 0;
        while (depth > 0) {
            if (++iters > 2u * g.n + 2u) {  // malformed nodes: stop and flag
                over = true;
                break;
            }
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

// The reference's walk over whichever structure the scene carries: the CPU
// midpoint BVH when g.bvh is set, else the LBVH.
template <bool STATS, class Alpha = NoAlpha, class Stack = LdsStack>
__device__ __forceinline__ uint32_t closestHit(const GeomScene& g, V3 o, V3 d, float& t,
                                               const Stack& stack, TraceStats& st,
                                               float* b1Out = nullptr, float* b2Out = nullptr,
                                               const Alpha& alpha = Alpha(),
                                               bool* overflow = nullptr) {
    if (g.bvh) return bvhWalk<STATS, false>(g, o, d, t, stack, st, b1Out, b2Out, alpha, overflow);
    return lbvhClosest<STATS>(g, o, d, t, stack, st, b1Out, b2Out, alpha, overflow);
}

// Occlusion query with the same visiting order and tests as closestHit with
// t = +inf, stopped at the first accepted triangle with a finite distance:
// true exactly when closestHit would report a hit.
template <class Alpha = NoAlpha, class Stack = LdsStack>
__device__ __forceinline__ bool anyHit(const GeomScene& g, V3 o, V3 d, const Stack& stack,
                                       const Alpha& alpha = Alpha(), bool* overflow = nullptr) {
    float t = INFINITY;
    if (g.bvh) {
        TraceStats st{0u, 0u};
        return bvhWalk<false, true>(g, o, d, t, stack, st, nullptr, nullptr, alpha, overflow) != kNone &&
               !isinf(t);
    }
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
    uint32_t iters = 0;
    while (depth > 0 && !hit) {
        if (++iters > 2u * g.n + 2u) {  // malformed nodes: stop and flag
            over = true;
            break;
        }
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

// Moller-Trumbore on the packed triangle copy (same operations as triTest).
__device__ __forceinline__ bool triTestPacked(const float4* tris, V3 o, V3 d, uint32_t faceIdx, float& t,
                                              float& b1Out, float& b2Out) {
    const float4* tp = tris + 3 * (size_t)faceIdx;
    const float4 a = tp[0], b = tp[1], c = tp[2];
    const V3 p0 = {a.x, a.y, a.z}, p1 = {b.x, b.y, b.z}, p2 = {c.x, c.y, c.z};
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

// Resumable ordered walk over the traversal layout: begin() sets up one ray,
// each step() either visits one internal node (both child boxes, its leaf
// children's triangles nearer first, then enter the nearer internal child and
// push the farther) or pops one candidate and re-tests its box against the
// current t. Same decisions, hits and counts as a walk run to completion, so
// callers can interleave rays (refill idle lanes) without changing results.
// any (set by begin): stop at the first accepted triangle with a finite
// distance (occlusion queries).
// PAIRS (for stacks with a spill region of kStackPairs slots): the farther
// child is pushed with its entry distance, and a visit that enters no child
// pops at once, entering the first candidate whose entry is still <= t. That
// is the same decision as testing its box again (the entry does not depend on
// t), without the box fetch and without a separate pop step, so hits and
// visit / test counts are unchanged.
constexpr int kStackPairs = 128;  // stack slots of a PAIRS walk: 64 (ref, entry) pairs

template <bool STATS, class Alpha, class Stack, bool PAIRS = false>
struct Ordered {
    GeomScene g;  // by value: rebind() switches scenes between rays
    const Stack& stack;
    Alpha alpha;
    V3 o, d, inv;
    float t, b1, b2;
    uint32_t best, node, steps;
    int depth;
    bool done, over, any;
    TraceStats st;

    __device__ __forceinline__ Ordered(const GeomScene& g_, const Stack& s_, const Alpha& a_)
        : g(g_), stack(s_), alpha(a_) {}
    __device__ __forceinline__ void rebind(const GeomScene& g_, const Alpha& a_) {
        g = g_;
        alpha = a_;
    }

    __device__ __forceinline__ void leaf(uint32_t face) {
        if constexpr (STATS) ++st.tests;
        float tt = t, x1, x2;
        if (!triTestPacked(g.atris, o, d, face, tt, x1, x2)) return;
        if (tt == t && face > best) return;
        if (alpha.reject(face, x1, x2)) return;
        t = tt;
        best = face;
        b1 = x1;
        b2 = x2;
        if (any && !isinf(tt)) done = true;
    }

    __device__ __forceinline__ void begin(V3 o_, V3 d_, float tmax, bool anyHit = false) {
        o = o_;
        d = d_;
        inv = {safeInv(d.x), safeInv(d.y), safeInv(d.z)};
        t = tmax;
        b1 = b2 = 0.0f;
        best = kNone;
        node = kNone;
        depth = 0;
        done = false;
        over = false;
        any = anyHit;
        steps = 0;
        st = {0u, 0u};
        if (g.n == 0) {
            done = true;
        } else if (g.n == 1) {
            // The single leaf: LBVH leaf 0 (objIdx) or the midpoint root leaf
            // (children[0]); both hold the box in their first 32 bytes and
            // the face at byte 32.
            const float4* lp = g.bvh ? g.bvh : reinterpret_cast<const float4*>(g.leaves);
            if (slabEnter(lp[0], lp[1], o, inv, t) <= t) leaf(reinterpret_cast<const uint32_t*>(lp + 2)[0]);
            done = true;
        } else {
            const float4* np = g.bvh ? g.bvh : reinterpret_cast<const float4*>(g.inodes);
            if (slabEnter(np[0], np[1], o, inv, t) <= t) node = 0;
            else done = true;
        }
    }

    __device__ __forceinline__ void step() {
        // A tree visits each node at most once and pops at most as often as
        // it pushes: more than 2n steps means malformed nodes; stop and flag.
        if (++steps > 2u * g.n + 2u) {
            done = true;
            over = true;
            return;
        }
        if (node != kNone) {
            if constexpr (STATS) ++st.visits;
            const float4* np = g.anodes + 4 * (size_t)node;
            const float4 lo0 = np[0], hi0 = np[1], lo1 = np[2], hi1 = np[3];
            const uint32_t r0 = __float_as_uint(lo0.w), r1 = __float_as_uint(hi0.w);
            const float e0 = slabEnter(lo0, hi0, o, inv, t), e1 = slabEnter(lo1, hi1, o, inv, t);
            const bool l0 = (r0 & kLeafRef) != 0, l1 = (r1 & kLeafRef) != 0;
            const bool first1 = e1 < e0;
            if (l0 || l1) {
                const uint32_t ra = first1 ? r1 : r0, rb = first1 ? r0 : r1;
                const bool la = first1 ? l1 : l0, lb = first1 ? l0 : l1;
                const float ea = first1 ? e1 : e0, eb = first1 ? e0 : e1;
                if (la && ea <= t) leaf(ra & ~kLeafRef);
                if (!done && lb && eb <= t) leaf(rb & ~kLeafRef);
                if (done) return;
            }
            const bool go0 = !l0 && e0 <= t, go1 = !l1 && e1 <= t;
            if (go0 && go1) {
                if constexpr (PAIRS) {
                    if (depth + 2 <= kStackPairs) {
                        stack.put(depth, first1 ? r0 : r1);
                        stack.put(depth + 1, __float_as_uint(first1 ? e0 : e1));
                        depth += 2;
                    } else {
                        over = true;
                    }
                } else {
                    if (depth < kStackDepth) stack.put(depth++, first1 ? r0 : r1);
                    else over = true;
                }
                node = first1 ? r1 : r0;
            } else if (go0 || go1) {
                node = go0 ? r0 : r1;
            } else {
                node = kNone;
                if constexpr (PAIRS) {
                    while (depth > 0) {
                        depth -= 2;
                        const uint32_t cand = stack.get(depth);
                        if (__uint_as_float(stack.get(depth + 1)) <= t) {
                            node = cand;
                            break;
                        }
                    }
                }
                if (node == kNone && depth == 0) done = true;
            }
            return;
        }
        if constexpr (PAIRS) {  // not reached: PAIRS visits pop inline
            done = true;
            return;
        }
        // The candidate's own box: first 32 bytes of its 48-byte record (LBVH
        // internal node, or midpoint BVH node by node number).
        const uint32_t cand = stack.get(--depth);
        const float4* cp = (g.bvh ? g.bvh : reinterpret_cast<const float4*>(g.inodes)) + 3 * (size_t)cand;
        if (slabEnter(cp[0], cp[1], o, inv, t) <= t) node = cand;
        else if (depth == 0) done = true;
    }
};

// Ordered walk over the 4-wide records (GeomScene::wnodes): a visit loads
// one 128-byte record and tests up to four grandchildren boxes, so the chain
// of dependent node loads is half the binary walk's. Leaf entries are tested
// nearest first at the visit; internal entries within t continue with the
// nearest and push the others (reference, entry distance) farthest first;
// a pop skips entries beyond the current t. A grandchild lies inside its
// parent's box, so testing it alone admits every triangle the binary walk
// would test, and the accepted hit is the same (closest t, then the smaller
// face index, alpha rejections as in the binary walk). CAP: stack slots (two
// per entry); a push beyond them is dropped and flagged.
template <bool STATS, class Alpha, class Stack, int CAP>
struct Wide {
    GeomScene g;  // by value: rebind() switches scenes between rays
    const Stack& stack;
    Alpha alpha;
    V3 o, d, inv;
    float t, b1, b2;
    uint32_t best, node, steps;
    int depth;
    bool done, over, any;
    TraceStats st;

    __device__ __forceinline__ Wide(const GeomScene& g_, const Stack& s_, const Alpha& a_)
        : g(g_), stack(s_), alpha(a_) {}
    __device__ __forceinline__ void rebind(const GeomScene& g_, const Alpha& a_) {
        g = g_;
        alpha = a_;
    }

    __device__ __forceinline__ void leaf(uint32_t face) {
        if constexpr (STATS) ++st.tests;
        float tt = t, x1, x2;
        if (!triTestPacked(g.atris, o, d, face, tt, x1, x2)) return;
        if (tt == t && face > best) return;
        if (alpha.reject(face, x1, x2)) return;
        t = tt;
        best = face;
        b1 = x1;
        b2 = x2;
        if (any && !isinf(tt)) done = true;
    }

    __device__ __forceinline__ void begin(V3 o_, V3 d_, float tmax, bool anyHit = false) {
        o = o_;
        d = d_;
        inv = {safeInv(d.x), safeInv(d.y), safeInv(d.z)};
        t = tmax;
        b1 = b2 = 0.0f;
        best = kNone;
        node = kNone;
        depth = 0;
        done = false;
        over = false;
        any = anyHit;
        steps = 0;
        st = {0u, 0u};
        if (g.n == 0) {
            done = true;
        } else if (g.n == 1) {
            const float4* lp = g.bvh ? g.bvh : reinterpret_cast<const float4*>(g.leaves);
            if (slabEnter(lp[0], lp[1], o, inv, t) <= t) leaf(reinterpret_cast<const uint32_t*>(lp + 2)[0]);
            done = true;
        } else {
            const float4* np = g.bvh ? g.bvh : reinterpret_cast<const float4*>(g.inodes);
            if (slabEnter(np[0], np[1], o, inv, t) <= t) node = 0;
            else done = true;
        }
    }

    __device__ __forceinline__ void step() {
        // Each node is visited at most once: more than 2n steps means
        // malformed records; stop and flag.
        if (++steps > 2u * g.n + 2u) {
            done = true;
            over = true;
            return;
        }
        if constexpr (STATS) ++st.visits;
        const float4* wp = g.wnodes + 8 * (size_t)node;
        const float4 lx = wp[0], ly = wp[1], lz = wp[2], hx = wp[3], hy = wp[4], hz = wp[5], rf = wp[6];
        float k[4] = {slabEnter(make_float4(lx.x, ly.x, lz.x, 0.0f), make_float4(hx.x, hy.x, hz.x, 0.0f), o, inv, t),
                      slabEnter(make_float4(lx.y, ly.y, lz.y, 0.0f), make_float4(hx.y, hy.y, hz.y, 0.0f), o, inv, t),
                      slabEnter(make_float4(lx.z, ly.z, lz.z, 0.0f), make_float4(hx.z, hy.z, hz.z, 0.0f), o, inv, t),
                      slabEnter(make_float4(lx.w, ly.w, lz.w, 0.0f), make_float4(hx.w, hy.w, hz.w, 0.0f), o, inv, t)};
        uint32_t r[4] = {__float_as_uint(rf.x), __float_as_uint(rf.y), __float_as_uint(rf.z), __float_as_uint(rf.w)};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (!(k[c] <= t) || r[c] == kNone) {  // a miss (NaN) or an empty slot
                k[c] = INFINITY;
                r[c] = kNone;
            }
        }
        // Nearest first: a five-exchange sorting network on (entry, ref).
        auto cx = [&](int a, int b) {
            if (k[b] < k[a]) {
                const float tk = k[a];
                k[a] = k[b];
                k[b] = tk;
                const uint32_t tr = r[a];
                r[a] = r[b];
                r[b] = tr;
            }
        };
        cx(0, 1);
        cx(2, 3);
        cx(0, 2);
        cx(1, 3);
        cx(1, 2);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (r[c] != kNone && (r[c] & kLeafRef) && k[c] <= t) {
                leaf(r[c] & ~kLeafRef);
                if (done) return;
            }
        }
        uint32_t next = kNone;
        float enext = 0.0f;
#pragma unroll
        for (int c = 3; c >= 0; --c) {
            if (!(r[c] & kLeafRef) && k[c] <= t) {
                if (next != kNone) push(next, enext);
                next = r[c];
                enext = k[c];
            }
        }
        if (next != kNone) {
            node = next;
            return;
        }
        node = kNone;
        while (depth > 0) {
            depth -= 2;
            const uint32_t cand = stack.get(depth);
            if (__uint_as_float(stack.get(depth + 1)) <= t) {
                node = cand;
                return;
            }
        }
        done = true;
    }

    __device__ __forceinline__ void push(uint32_t ref, float entry) {
        if (depth + 2 <= CAP) {
            stack.put(depth, ref);
            stack.put(depth + 1, __float_as_uint(entry));
            depth += 2;
        } else {
            over = true;
        }
    }
};

// The 4-wide walk run to completion for one ray (as orderedHit).
template <bool STATS, bool ANY, class Alpha = NoAlpha, class Stack = LdsStack, int CAP = kStackDepth>
__device__ __forceinline__ uint32_t wideHit(const GeomScene& g, V3 o, V3 d, float& t, const Stack& stack,
                                            TraceStats& st, float* b1Out = nullptr, float* b2Out = nullptr,
                                            const Alpha& alpha = Alpha(), bool* overflow = nullptr) {
    Wide<STATS, Alpha, Stack, CAP> w(g, stack, alpha);
    w.begin(o, d, t, ANY);
    while (!w.done) w.step();
    t = w.t;
    if (b1Out) *b1Out = w.b1;
    if (b2Out) *b2Out = w.b2;
    if (overflow && w.over) *overflow = true;
    if constexpr (STATS) st = w.st;
    return w.best;
}

// The ordered walk run to completion for one ray. Returns the face (kNone
// when nothing was accepted) and updates t, b1, b2.
template <bool STATS, bool ANY, class Alpha = NoAlpha, class Stack = LdsStack, bool PAIRS = false>
__device__ __forceinline__ uint32_t orderedHit(const GeomScene& g, V3 o, V3 d, float& t, const Stack& stack,
                                               TraceStats& st, float* b1Out = nullptr, float* b2Out = nullptr,
                                               const Alpha& alpha = Alpha(), bool* overflow = nullptr) {
    Ordered<STATS, Alpha, Stack, PAIRS> w(g, stack, alpha);
    w.begin(o, d, t, ANY);
    while (!w.done) w.step();
    t = w.t;
    if (b1Out) *b1Out = w.b1;
    if (b2Out) *b2Out = w.b2;
    if (overflow && w.over) *overflow = true;
    if constexpr (STATS) st = w.st;
    return w.best;
}

}  // namespace dev
}  // namespace clr
