// CPU oracle for the LBVH build — TEST INFRASTRUCTURE ONLY.
//
// A serial restatement of the reference's OpenCL build
// (OpenCL_PathTracer/OpenCL/lbvh_construction.cl, driven by
// OpenCL_PathTracer/LBVHBuilder.cpp:160-375). Only tests/, the smoke check and
// bench.py's cpu_baseline leg may load it; the product never does.
//
// Arithmetic rules (fixed so that the GPU path can match bit for bit):
//  * compiled with -ffp-contract=off, no fast-math: IEEE +,-,*,/ only;
//  * OpenCL fmin/fmax are restated as "NaN yields the other operand, ties keep
//    the left operand" (omin/omax), which removes the +-0 ambiguity;
//  * convert_uint_rtz of NaN or a non-positive value is 0 (OpenCL leaves
//    out-of-range conversions undefined).
//
// Independent pins live in lbvh_pins.cpp (stable sort, top-down radix tree,
// per-range folds) and in tests/test_oracle_lbvh.py (closed forms).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/clear_amd.h"

namespace {

struct F3 {
    float x, y, z;
};

inline float omin(float a, float b) { return (b < a || a != a) ? b : a; }
inline float omax(float a, float b) { return (b > a || a != a) ? b : a; }
inline F3 vmin(F3 a, F3 b) { return {omin(a.x, b.x), omin(a.y, b.y), omin(a.z, b.z)}; }
inline F3 vmax(F3 a, F3 b) { return {omax(a.x, b.x), omax(a.y, b.y), omax(a.z, b.z)}; }

struct AABB {
    F3 min, max, center;
};

inline uint32_t clz32(uint32_t v) { return v ? (uint32_t)__builtin_clz(v) : 32u; }

inline F3 load3(const clr_float3& v) { return {v.x, v.y, v.z}; }

// calcAABBs (lbvh_construction.cl:81-96): box.min = fmin(p0, fmin(p1, p2)),
// box.max likewise, center = (min + max) * 0.5f.
AABB tri_aabb(F3 p0, F3 p1, F3 p2) {
    AABB b;
    b.min = vmin(p0, vmin(p1, p2));
    b.max = vmax(p0, vmax(p1, p2));
    b.center = {(b.min.x + b.max.x) * 0.5f, (b.min.y + b.max.y) * 0.5f,
                (b.min.z + b.max.z) * 0.5f};
    return b;
}

// min(convert_uint_rtz(v), numDivs - 1) (lbvh_construction.cl:179).
inline uint32_t quantize(float v, uint32_t numDivs) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return numDivs - 1u;
    uint32_t q = (uint32_t)v;
    return q < numDivs - 1u ? q : numDivs - 1u;
}

// _numCommonBits (lbvh_construction.cl:322-332). Returns -1 outside [0, n).
int numCommonBits(const uint32_t* mc, const uint32_t* indices, uint32_t n, uint32_t bits,
                  const uint32_t* mci, int64_t i0, int64_t i1) {
    if (i1 < 0 || i1 >= (int64_t)n) return -1;
    uint32_t j = indices[i1];
    int lx = (int)(clz32(mci[0] ^ mc[3 * j + 0]) - (32u - bits));
    int ly = (int)(clz32(mci[1] ^ mc[3 * j + 1]) - (32u - bits));
    int lz = (int)(clz32(mci[2] ^ mc[3 * j + 2]) - (32u - bits));
    int ret = std::min(lz * 3 + 0, std::min(ly * 3 + 1, lx * 3 + 2));
    if ((uint32_t)ret != bits * 3u) return ret;
    return ret + (int)clz32((uint32_t)((int32_t)i0 ^ (int32_t)i1));
}

inline int sign8(int v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }

}  // namespace

extern "C" {

// Per-triangle AABBs from vertices + faces (calcAABBs). out: float[n][9]
// (min xyz, max xyz, center xyz).
int orc_tri_aabbs(const clr_float3* verts, const clr_face* faces, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        AABB b = tri_aabb(load3(verts[faces[i].p0]), load3(verts[faces[i].p1]),
                          load3(verts[faces[i].p2]));
        float* o = out + 9 * (size_t)i;
        o[0] = b.min.x; o[1] = b.min.y; o[2] = b.min.z;
        o[3] = b.max.x; o[4] = b.max.y; o[5] = b.max.z;
        o[6] = b.center.x; o[7] = b.center.y; o[8] = b.center.z;
    }
    return 0;
}

// Full build. Outputs: internal nodes [n-1], leaves [n]; optionally the
// per-primitive Morton triples (uint32 [n][3], input order), the sorted
// indices (uint32 [n]) and the scene AABB (float[6]: min xyz, max xyz).
int orc_lbvh_build(const clr_float3* verts, const clr_face* faces, uint32_t n, uint32_t bits,
                   clr_lbvh_inode* inodes, clr_lbvh_leaf* leaves, uint32_t* mortonOut,
                   uint32_t* sortedOut, float* sceneOut) {
    if (n == 0 || bits < 1 || bits > 21) return 1;

    // calcAABBs
    std::vector<AABB> aabbs(n);
    for (uint32_t i = 0; i < n; ++i)
        aabbs[i] = tri_aabb(load3(verts[faces[i].p0]), load3(verts[faces[i].p1]),
                            load3(verts[faces[i].p2]));

    // unifyAABBs (lbvh_construction.cl:105-162): min/max fold of the triangle
    // boxes, identity = +-INFINITY (the padding of the last block).
    F3 smin = {INFINITY, INFINITY, INFINITY};
    F3 smax = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i) {
        smin = vmin(smin, aabbs[i].min);
        smax = vmax(smax, aabbs[i].max);
    }
    if (sceneOut) {
        sceneOut[0] = smin.x; sceneOut[1] = smin.y; sceneOut[2] = smin.z;
        sceneOut[3] = smax.x; sceneOut[4] = smax.y; sceneOut[5] = smax.z;
    }
    // Host-side size (LBVHBuilder.cpp:236-240).
    F3 ssize = {smax.x - smin.x, smax.y - smin.y, smax.z - smin.z};

    // calcMortonCodes (lbvh_construction.cl:169-181), numDivs = 1 << bits
    // (LBVHBuilder.cpp:244).
    const uint32_t numDivs = 1u << bits;
    const float fDivs = (float)numDivs;
    std::vector<uint32_t> mc(3 * (size_t)n);
    std::vector<uint32_t> indices(n), shadow(n);
    for (uint32_t i = 0; i < n; ++i) {
        F3 c = aabbs[i].center;
        float nx = (c.x - smin.x) / ssize.x;
        float ny = (c.y - smin.y) / ssize.y;
        float nz = (c.z - smin.z) / ssize.z;
        mc[3 * i + 0] = quantize(nx * fDivs, numDivs);
        mc[3 * i + 1] = quantize(ny * fDivs, numDivs);
        mc[3 * i + 2] = quantize(nz * fDivs, numDivs);
        indices[i] = i;
    }
    if (mortonOut) std::memcpy(mortonOut, mc.data(), mc.size() * sizeof(uint32_t));

    // Radix sort (LBVHBuilder.cpp:254-288): one pass per bit level, LSB first.
    // Each pass is blockwiseSort (3 stable 1-bit splits, x then y then z,
    // lbvh_construction.cl:231-272) + calcBlockwiseHistograms + global scan +
    // globalScatter (:279-316). Together they are a stable counting sort of the
    // index array by the 3-bit digit x + 2y + 4z; restated directly.
    for (uint32_t bit = 0; bit < bits; ++bit) {
        uint32_t count[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        std::vector<uint8_t> digit(n);
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t idx = indices[i];
            uint8_t d = (uint8_t)(((mc[3 * idx + 0] >> bit) & 1u) |
                                  (((mc[3 * idx + 1] >> bit) & 1u) << 1) |
                                  (((mc[3 * idx + 2] >> bit) & 1u) << 2));
            digit[i] = d;
            ++count[d];
        }
        uint32_t offset[8];
        uint32_t run = 0;
        for (int d = 0; d < 8; ++d) { offset[d] = run; run += count[d]; }
        for (uint32_t i = 0; i < n; ++i) shadow[offset[digit[i]]++] = indices[i];
        indices.swap(shadow);
    }
    if (sortedOut) std::memcpy(sortedOut, indices.data(), n * sizeof(uint32_t));

    // constructBinaryRadixTree (lbvh_construction.cl:342-410).
    std::vector<uint32_t> parents(2 * (size_t)n, 0xFFFFFFFFu);
    std::vector<uint32_t> counters(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        clr_lbvh_leaf& L = leaves[i];
        std::memset(&L, 0, sizeof(L));
        uint32_t obj = indices[i];
        L.min[0] = aabbs[obj].min.x; L.min[1] = aabbs[obj].min.y; L.min[2] = aabbs[obj].min.z;
        L.max[0] = aabbs[obj].max.x; L.max[1] = aabbs[obj].max.y; L.max[2] = aabbs[obj].max.z;
        L.objIdx = obj;
    }
    for (uint32_t i = 0; i + 1 < n; ++i) {
        const uint32_t* mci = &mc[3 * (size_t)indices[i]];
#define NCB(a, b) numCommonBits(mc.data(), indices.data(), n, bits, mci, (a), (b))
        int commonR = NCB(i, (int64_t)i + 1);
        int commonL = NCB(i, (int64_t)i - 1);
        int d = sign8(commonR - commonL);
        int minNumComBits = d > 0 ? commonL : commonR;
        uint64_t lMax = 2;
        while (NCB(i, (int64_t)i + (int64_t)lMax * d) > minNumComBits) lMax <<= 1;
        uint64_t l = 0;
        for (uint64_t t = lMax >> 1; t >= 1; t >>= 1)
            if (NCB(i, (int64_t)i + (int64_t)(l + t) * d) > minNumComBits) l += t;
        int64_t otherEnd = (int64_t)i + (int64_t)l * d;
        int numComBitsNode = NCB(i, otherEnd);
        uint64_t splitPos = 0;
        float scale = 0.5f;  // float split search, as written (:383-391)
        while (true) {
            uint64_t t = (uint64_t)std::ceil((float)l * scale);
            if (NCB(i, (int64_t)i + (int64_t)(splitPos + t) * d) > numComBitsNode) splitPos += t;
            if (t <= 1) break;
            scale *= 0.5f;
        }
#undef NCB
        int64_t split = (int64_t)i + (int64_t)splitPos * d + std::min(d, 0);
        clr_lbvh_inode& N = inodes[i];
        std::memset(&N, 0, sizeof(N));
        int64_t lo = std::min<int64_t>(i, otherEnd), hi = std::max<int64_t>(i, otherEnd);
        bool leftIsLeaf = lo == split;
        bool rightIsLeaf = hi == split + 1;
        N.isChild[0] = leftIsLeaf;
        N.isChild[1] = rightIsLeaf;
        N.c[0] = (uint32_t)split;
        N.c[1] = (uint32_t)(split + 1);
        parents[(leftIsLeaf ? 0 : n) + (size_t)split] = i;
        parents[(rightIsLeaf ? 0 : n) + (size_t)split + 1] = i;
    }

    // calcNodeAABBs (lbvh_construction.cl:415-492): every leaf climbs; the
    // first arrival at a node stops, the second merges fmin(left, right). The
    // merge order is fixed, so the serial arrival order below gives the same
    // boxes as any parallel interleaving.
    if (n >= 2) {
        for (uint32_t g = 0; g < n; ++g) {
            F3 mn = {leaves[g].min[0], leaves[g].min[1], leaves[g].min[2]};
            F3 mx = {leaves[g].max[0], leaves[g].max[1], leaves[g].max[2]};
            uint32_t self = g;
            bool selfIsLeaf = true;
            uint32_t tgt = parents[g];
            while (true) {
                if (counters[tgt]++ == 0) break;
                clr_lbvh_inode& T = inodes[tgt];
                auto childMin = [&](int k) -> F3 {
                    uint32_t c = T.c[k];
                    if (T.isChild[k]) return {leaves[c].min[0], leaves[c].min[1], leaves[c].min[2]};
                    return {inodes[c].min[0], inodes[c].min[1], inodes[c].min[2]};
                };
                auto childMax = [&](int k) -> F3 {
                    uint32_t c = T.c[k];
                    if (T.isChild[k]) return {leaves[c].max[0], leaves[c].max[1], leaves[c].max[2]};
                    return {inodes[c].max[0], inodes[c].max[1], inodes[c].max[2]};
                };
                bool selfIsLeft = selfIsLeaf ? (T.c[0] == self && T.isChild[0]) : (T.c[0] == self);
                if (selfIsLeft) {
                    mn = vmin(mn, childMin(1));
                    mx = vmax(mx, childMax(1));
                } else {
                    mn = vmin(childMin(0), mn);
                    mx = vmax(childMax(0), mx);
                }
                T.min[0] = mn.x; T.min[1] = mn.y; T.min[2] = mn.z;
                T.max[0] = mx.x; T.max[1] = mx.y; T.max[2] = mx.z;
                if (tgt == 0) break;
                self = tgt;
                selfIsLeaf = false;
                tgt = parents[n + tgt];
            }
        }
    }
    return 0;
}

}  // extern "C"
