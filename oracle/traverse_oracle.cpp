// CPU oracle for closest-hit traversal — TEST INFRASTRUCTURE ONLY.
//
// Serial restatement of the reference's LBVH traversal
// (OpenCL_PathTracer/OpenCL/bvh_traversal.cl; its CPU twin is
// Equivalent/sim_bvh_traversal.hpp). Loaded by tests/ and bench.py's
// cpu_baseline leg only.
//
// Arithmetic: -ffp-contract=off; dot/cross left to right as written in
// Equivalent/sim_base.hpp:695-703; OpenCL fmax restated as omax (NaN yields
// the other operand). Documented deviation from the reference: face indices
// are 32-bit (the reference passes `ushort faceIdx`, bvh_traversal.cl:43,55,
// which truncates every index above 65535).
//
// Pinned by orc_trace_bruteforce (an O(N) loop of the same triangle test over
// every face, no boxes) in tests/test_oracle_traverse.py.
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../include/clear_amd.h"
#include "ordered_traversal.h"

namespace {

std::atomic<uint64_t> g_overflowedRays{0};

struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float omax(float a, float b) { return (b > a || a != a) ? b : a; }
inline V3 ld(const clr_float3& v) { return {v.x, v.y, v.z}; }

// rayTriangleIntersection (bvh_traversal.cl:54-108), geometric part. The
// alpha-texture rejection (:102-105) needs UVs and textures and is part of
// the render oracle; faces in this entry point carry no alpha texture.
bool triTest(const clr_float3* verts, const clr_face* faces, V3 org, V3 dir, uint32_t faceIdx,
             float* t) {
    const clr_face& f = faces[faceIdx];
    V3 p0 = ld(verts[f.p0]), p1 = ld(verts[f.p1]), p2 = ld(verts[f.p2]);
    V3 edge01 = sub(p1, p0);
    V3 edge02 = sub(p2, p0);
    V3 p = cross(dir, edge02);
    float det = dot(edge01, p);
    if (det > -0.000001f && det < 0.000001f) return false;
    float invDet = 1.0f / det;
    V3 d = sub(org, p0);
    float b1 = dot(d, p) * invDet;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    V3 q = cross(d, edge01);
    float b2 = dot(dir, q) * invDet;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
    float tt = dot(edge02, q) * invDet;
    if (tt < 0.0f || tt > *t) return false;
    *t = tt;
    return true;
}

// rayAABBIntersection (bvh_traversal.cl:160-179), per component as the
// OpenCL vector relational/select semantics define it.
bool boxTest(const float* bmin, const float* bmax, V3 org, V3 dir, float tBound) {
    const float o[3] = {org.x, org.y, org.z};
    const float dv[3] = {dir.x, dir.y, dir.z};
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
            float coord = o[k] + maxt * dv[k];
            bool hit = (coord >= bmin[k] && coord <= bmax[k]) || t[k] == maxt;
            if (!hit) return false;
        }
    }
    return true;
}

}  // namespace

extern "C" {

// rayIntersection over the LBVH (bvh_traversal.cl:182-215): explicit stack of
// 64, node 0 first; pop, test the internal node's box against the current t;
// for child 0 then child 1: push an internal child, or test the leaf box and
// then its triangle. A triangle replaces the hit when tt <= t. The hit flag is
// !isinf(t) for tmax = inf; with a finite tmax a hit is "a triangle accepted".
// numFaces == 1 (no internal node) tests leaf 0 directly.
int orc_trace_closest(const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, uint32_t n,
                      const clr_float3* verts, const clr_face* faces, const float* rays,
                      uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats) {
    for (uint32_t r = 0; r < numRays; ++r) {
        V3 org = {rays[0 * (size_t)numRays + r], rays[1 * (size_t)numRays + r],
                  rays[2 * (size_t)numRays + r]};
        V3 dir = {rays[3 * (size_t)numRays + r], rays[4 * (size_t)numRays + r],
                  rays[5 * (size_t)numRays + r]};
        float t = rays[6 * (size_t)numRays + r];
        uint32_t best = 0xFFFFFFFFu;
        uint32_t visits = 0, tests = 0;
        bool overflowed = false;
        if (n == 1) {
            if (boxTest(leaves[0].min, leaves[0].max, org, dir, t)) {
                ++tests;
                if (triTest(verts, faces, org, dir, leaves[0].objIdx, &t)) best = leaves[0].objIdx;
            }
        } else if (n > 1) {
            uint32_t stack[64];
            uint32_t depth = 0;
            stack[depth++] = 0;
            while (depth > 0) {
                --depth;
                const clr_lbvh_inode& N = inodes[stack[depth]];
                if (!boxTest(N.min, N.max, org, dir, t)) continue;
                ++visits;
                for (int i = 0; i < 2; ++i) {
                    if (!N.isChild[i]) {
                        // A push beyond the 64 entries is dropped, as the GPU walk
                        // drops it (raising its overflow status); counted per ray.
                        if (depth >= 64) {
                            overflowed = true;
                            continue;
                        }
                        stack[depth++] = N.c[i];
                        continue;
                    }
                    const clr_lbvh_leaf& L = leaves[N.c[i]];
                    if (!boxTest(L.min, L.max, org, dir, t)) continue;
                    ++tests;
                    if (triTest(verts, faces, org, dir, L.objIdx, &t)) best = L.objIdx;
                }
            }
        }
        if (overflowed) ++g_overflowedRays;
        bool hit = best != 0xFFFFFFFFu && !std::isinf(t);
        hitFace[r] = hit ? best : 0xFFFFFFFFu;
        hitT[r] = t;
        if (stats) {
            stats[2 * (size_t)r + 0] = visits;
            stats[2 * (size_t)r + 1] = tests;
        }
    }
    return 0;
}

// Rays whose reference walk (orc_trace_closest) dropped a push since the last
// reset.
uint64_t orc_trace_overflows(int reset) {
    const uint64_t v = g_overflowedRays;
    if (reset) g_overflowedRays = 0;
    return v;
}

// rayIntersection over the CPU midpoint BVH (bvh_traversal.cl:256-285):
// explicit stack of 64, node 0 first; pop, test the node's box against the
// current t; an internal node pushes children[1] then children[0] (so
// children[0] is visited first); a leaf (children[1] == 0xFFFFFFFF) tests
// triangle children[0], replacing the hit when tt <= t. stats: (internal
// nodes whose box passed, triangle tests) per ray.
int orc_trace_bvh(const clr_bvh_node* nodes, uint32_t numNodes, const clr_float3* verts,
                  const clr_face* faces, const float* rays, uint32_t numRays, uint32_t* hitFace,
                  float* hitT, uint32_t* stats) {
    for (uint32_t r = 0; r < numRays; ++r) {
        const size_t R = numRays;
        const V3 org = {rays[r], rays[R + r], rays[2 * R + r]};
        const V3 dir = {rays[3 * R + r], rays[4 * R + r], rays[5 * R + r]};
        float t = rays[6 * R + r];
        uint32_t best = 0xFFFFFFFFu, visits = 0, tests = 0;
        uint32_t stack[64];
        uint32_t depth = 0;
        if (numNodes) stack[depth++] = 0;
        while (depth > 0) {
            const clr_bvh_node& N = nodes[stack[--depth]];
            if (!boxTest(N.min, N.max, org, dir, t)) continue;
            if (N.children[1] != 0xFFFFFFFFu) {
                ++visits;
                if (depth + 2 > 64) return 2;  // stack bound exceeded
                if (N.children[0] >= numNodes || N.children[1] >= numNodes) return 3;
                stack[depth++] = N.children[1];
                stack[depth++] = N.children[0];
            } else {
                ++tests;
                if (triTest(verts, faces, org, dir, N.children[0], &t)) best = N.children[0];
            }
        }
        const bool hit = best != 0xFFFFFFFFu && !std::isinf(t);
        hitFace[r] = hit ? best : 0xFFFFFFFFu;
        hitT[r] = t;
        if (stats) {
            stats[2 * (size_t)r + 0] = visits;
            stats[2 * (size_t)r + 1] = tests;
        }
    }
    return 0;
}

// Ordered traversal (ordered_traversal.h) with the geometric triangle test:
// face, t and per-ray (visits, tests); anyHit != 0 stops at the first
// accepted triangle (hitFace then only tells hit / miss).
int orc_trace_ordered(const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, uint32_t n,
                      const clr_float3* verts, const clr_face* faces, const float* rays, uint32_t numRays,
                      int anyHit, uint32_t* hitFace, float* hitT, uint32_t* stats) {
    for (uint32_t r = 0; r < numRays; ++r) {
        const size_t R = numRays;
        const float o[3] = {rays[r], rays[R + r], rays[2 * R + r]};
        const float d[3] = {rays[3 * R + r], rays[4 * R + r], rays[5 * R + r]};
        float t = rays[6 * R + r];
        const V3 org = {o[0], o[1], o[2]}, dir = {d[0], d[1], d[2]};
        auto tri = [&](uint32_t face, float bound, float* tt) {
            float x = bound;
            if (!triTest(verts, faces, org, dir, face, &x)) return false;
            *tt = x;
            return true;
        };
        orc_ordered::Counts cnt;
        bool over = false;
        const uint32_t best = orc_ordered::traverse(inodes, leaves, n, o, d, t, tri, cnt, anyHit != 0, &over);
        if (over) return 2;
        const bool hit = best != 0xFFFFFFFFu && !std::isinf(t);
        hitFace[r] = hit ? best : 0xFFFFFFFFu;
        hitT[r] = t;
        if (stats) {
            stats[2 * (size_t)r + 0] = cnt.visits;
            stats[2 * (size_t)r + 1] = cnt.tests;
        }
    }
    return 0;
}

// The 4-wide form of the ordered traversal (orc_ordered::traverseWide): same
// outputs as orc_trace_ordered, stats count visited records and tests.
int orc_trace_wide(const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, uint32_t n, const clr_float3* verts,
                   const clr_face* faces, const float* rays, uint32_t numRays, int anyHit, uint32_t* hitFace,
                   float* hitT, uint32_t* stats) {
    for (uint32_t r = 0; r < numRays; ++r) {
        const size_t R = numRays;
        const float o[3] = {rays[r], rays[R + r], rays[2 * R + r]};
        const float d[3] = {rays[3 * R + r], rays[4 * R + r], rays[5 * R + r]};
        float t = rays[6 * R + r];
        const V3 org = {o[0], o[1], o[2]}, dir = {d[0], d[1], d[2]};
        auto tri = [&](uint32_t face, float bound, float* tt) {
            float x = bound;
            if (!triTest(verts, faces, org, dir, face, &x)) return false;
            *tt = x;
            return true;
        };
        orc_ordered::Counts cnt;
        bool over = false;
        const uint32_t best = orc_ordered::traverseWide(inodes, leaves, n, o, d, t, tri, cnt, anyHit != 0, &over);
        if (over) return 2;
        const bool hit = best != 0xFFFFFFFFu && !std::isinf(t);
        hitFace[r] = hit ? best : 0xFFFFFFFFu;
        hitT[r] = t;
        if (stats) {
            stats[2 * (size_t)r + 0] = cnt.visits;
            stats[2 * (size_t)r + 1] = cnt.tests;
        }
    }
    return 0;
}

// Brute force with the ordered traversal's rule: the lexicographic minimum of
// (t, face) over all faces the triangle test accepts.
int orc_trace_bruteforce_minid(const clr_float3* verts, const clr_face* faces, uint32_t n, const float* rays,
                               uint32_t numRays, uint32_t* hitFace, float* hitT) {
    for (uint32_t r = 0; r < numRays; ++r) {
        const size_t R = numRays;
        const V3 org = {rays[r], rays[R + r], rays[2 * R + r]};
        const V3 dir = {rays[3 * R + r], rays[4 * R + r], rays[5 * R + r]};
        float t = rays[6 * R + r];
        uint32_t best = 0xFFFFFFFFu;
        for (uint32_t f = 0; f < n; ++f) {
            float x = t;
            if (!triTest(verts, faces, org, dir, f, &x)) continue;
            if (x == t && f > best) continue;
            t = x;
            best = f;
        }
        const bool hit = best != 0xFFFFFFFFu && !std::isinf(t);
        hitFace[r] = hit ? best : 0xFFFFFFFFu;
        hitT[r] = t;
    }
    return 0;
}

// Brute-force closest hit: the same triangle test over every face in index
// order, no boxes. nTied[r] = number of faces whose t equals the final t
// exactly (ties are broken by test order, so IDs are only comparable when
// nTied == 1).
int orc_trace_bruteforce(const clr_float3* verts, const clr_face* faces, uint32_t n,
                         const float* rays, uint32_t numRays, uint32_t* hitFace, float* hitT,
                         uint32_t* nTied) {
    for (uint32_t r = 0; r < numRays; ++r) {
        V3 org = {rays[0 * (size_t)numRays + r], rays[1 * (size_t)numRays + r],
                  rays[2 * (size_t)numRays + r]};
        V3 dir = {rays[3 * (size_t)numRays + r], rays[4 * (size_t)numRays + r],
                  rays[5 * (size_t)numRays + r]};
        float t0 = rays[6 * (size_t)numRays + r];
        float t = t0;
        uint32_t best = 0xFFFFFFFFu;
        for (uint32_t f = 0; f < n; ++f)
            if (triTest(verts, faces, org, dir, f, &t)) best = f;
        uint32_t ties = 0;
        if (best != 0xFFFFFFFFu) {
            for (uint32_t f = 0; f < n; ++f) {
                float tf = t;
                if (triTest(verts, faces, org, dir, f, &tf) && tf == t) ++ties;
            }
        }
        bool hit = best != 0xFFFFFFFFu && !std::isinf(t);
        hitFace[r] = hit ? best : 0xFFFFFFFFu;
        hitT[r] = t;
        if (nTied) nTied[r] = ties;
    }
    return 0;
}

}  // extern "C"
