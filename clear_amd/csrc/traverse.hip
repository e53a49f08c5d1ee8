// Closest-hit traversal of the LBVH for gfx950.
//
// Replaces rayIntersection (OpenCL_PathTracer/OpenCL/bvh_traversal.cl:182-215)
// with the same visiting order and tests, so hits agree bit for bit:
//   * pop a node, test its own box (rayAABBIntersection, :160-179) against the
//     current t; for child 0 then child 1: push an internal child, or test the
//     leaf box and then the triangle (Moller-Trumbore, :54-108);
//   * a triangle replaces the hit when tt <= t (ties go to the last tested).
// Layout choices for the hardware: one ray per lane, SoA ray planes read
// coalesced, the 64-entry stack lives in LDS as [depth][lane] (16 KiB per
// wave, conflict-free at equal depth) instead of per-lane scratch, and node /
// vertex records are fetched as 16-B vectors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <string>

#include "internal.h"
#include "traverse_device.h"

namespace clr {
namespace {

constexpr int kTraceBlock = 64;

template <bool STATS>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(
    const clr_lbvh_inode* __restrict__ inodes, const clr_lbvh_leaf* __restrict__ leaves,
    uint32_t n, const clr_float3* __restrict__ verts, const clr_face* __restrict__ faces,
    const float* __restrict__ rays, uint32_t numRays, uint32_t* __restrict__ hitFace,
    float* __restrict__ hitT, uint32_t* __restrict__ stats, uint32_t* __restrict__ status,
    const float4* __restrict__ bvh) {
    __shared__ uint32_t s_stack[dev::kStackDepth * kTraceBlock];
    const uint32_t r = blockIdx.x * kTraceBlock + threadIdx.x;
    if (r >= numRays) return;
    const size_t R = numRays;
    dev::V3 o = {rays[r], rays[R + r], rays[2 * R + r]};
    dev::V3 d = {rays[3 * R + r], rays[4 * R + r], rays[5 * R + r]};
    float t = rays[6 * R + r];
    dev::GeomScene g{inodes, leaves, n, verts, faces, nullptr, nullptr, bvh};
    dev::TraceStats st{0u, 0u};
    bool over = false;
    uint32_t best = dev::closestHit<STATS>(g, o, d, t, dev::LdsStack{s_stack + threadIdx.x, kTraceBlock},
                                           st, nullptr, nullptr, dev::NoAlpha(), &over);
    if (over) atomicOr(status, kStatusStackOverflow);
    const bool hit = best != kEmpty && !isinf(t);
    hitFace[r] = hit ? best : kEmpty;
    hitT[r] = t;
    if constexpr (STATS) {
        stats[2 * (size_t)r] = st.visits;
        stats[2 * (size_t)r + 1] = st.tests;
    }
}

// Traversal layout (see GeomScene): node i gets both children's boxes and
// references, face f its three vertices.
__global__ void k_accel_build(const clr_lbvh_inode* __restrict__ inodes, const clr_lbvh_leaf* __restrict__ leaves,
                              uint32_t n, const clr_float3* __restrict__ verts, const clr_face* __restrict__ faces,
                              float4* __restrict__ anodes, float4* __restrict__ atris) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 < n) {
        const uint4 c = reinterpret_cast<const uint4*>(inodes + i)[2];
        float4 box[2][2];
        uint32_t ref[2];
        for (int k = 0; k < 2; ++k) {
            const bool leaf = ((c.x >> (8 * k)) & 0xFFu) != 0;
            const uint32_t ci = k ? c.z : c.y;
            const float4* p = leaf ? reinterpret_cast<const float4*>(leaves + ci)
                                   : reinterpret_cast<const float4*>(inodes + ci);
            box[k][0] = p[0];
            box[k][1] = p[1];
            ref[k] = leaf ? (leaves[ci].objIdx | dev::kLeafRef) : ci;
        }
        float4* o = anodes + 4 * (size_t)i;
        o[0] = make_float4(box[0][0].x, box[0][0].y, box[0][0].z, __uint_as_float(ref[0]));
        o[1] = make_float4(box[0][1].x, box[0][1].y, box[0][1].z, __uint_as_float(ref[1]));
        o[2] = make_float4(box[1][0].x, box[1][0].y, box[1][0].z, 0.0f);
        o[3] = make_float4(box[1][1].x, box[1][1].y, box[1][1].z, 0.0f);
    }
    if (i < n) {
        const uint32_t* f = reinterpret_cast<const uint32_t*>(faces + i);
        float4* o = atris + 3 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            const float4 v = *reinterpret_cast<const float4*>(verts + f[k]);
            o[k] = make_float4(v.x, v.y, v.z, 0.0f);
        }
    }
}

// Traversal layout of a CPU midpoint BVH, indexed by node number: internal
// node i (children[1] != 0xFFFFFFFF) gets both children's boxes, a leaf child
// as its face | kLeafRef, an internal child as its node number; leaf node
// slots stay unused. Faces as in k_accel_build.
__global__ void k_accel_build_bvh(const float4* __restrict__ bvh, uint32_t n, const clr_float3* __restrict__ verts,
                                  const clr_face* __restrict__ faces, float4* __restrict__ anodes,
                                  float4* __restrict__ atris) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t numNodes = 2 * n - 1;
    if (i < numNodes) {
        const uint2 c = *reinterpret_cast<const uint2*>(bvh + 3 * (size_t)i + 2);
        if (c.y != kEmpty && c.x < numNodes && c.y < numNodes) {
            float4 box[2][2];
            uint32_t ref[2];
            for (int k = 0; k < 2; ++k) {
                const uint32_t ci = k ? c.y : c.x;
                const float4* p = bvh + 3 * (size_t)ci;
                box[k][0] = p[0];
                box[k][1] = p[1];
                const uint2 cc = *reinterpret_cast<const uint2*>(p + 2);
                ref[k] = cc.y == kEmpty ? (cc.x | dev::kLeafRef) : ci;
            }
            float4* o = anodes + 4 * (size_t)i;
            o[0] = make_float4(box[0][0].x, box[0][0].y, box[0][0].z, __uint_as_float(ref[0]));
            o[1] = make_float4(box[0][1].x, box[0][1].y, box[0][1].z, __uint_as_float(ref[1]));
            o[2] = make_float4(box[1][0].x, box[1][0].y, box[1][0].z, 0.0f);
            o[3] = make_float4(box[1][1].x, box[1][1].y, box[1][1].z, 0.0f);
        }
    }
    if (i < n) {
        const uint32_t* f = reinterpret_cast<const uint32_t*>(faces + i);
        float4* o = atris + 3 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            const float4 v = *reinterpret_cast<const float4*>(verts + f[k]);
            o[k] = make_float4(v.x, v.y, v.z, 0.0f);
        }
    }
}

// 4-wide records (GeomScene::wnodes) of the LBVH: record i holds the
// children of each internal child of node i, or a leaf child itself, in slot
// order left.left, left.right, right.left, right.right (a leaf child takes
// the first of its two slots; the other stays empty).
struct WideSlots {
    float lo[3][4], hi[3][4];
    uint32_t ref[4];
    __device__ __forceinline__ void clear() {
        for (int c = 0; c < 4; ++c) {
            for (int a = 0; a < 3; ++a) {
                lo[a][c] = INFINITY;
                hi[a][c] = -INFINITY;
            }
            ref[c] = dev::kNone;
        }
    }
    __device__ __forceinline__ void put(int c, float4 l, float4 h, uint32_t r) {
        lo[0][c] = l.x;
        lo[1][c] = l.y;
        lo[2][c] = l.z;
        hi[0][c] = h.x;
        hi[1][c] = h.y;
        hi[2][c] = h.z;
        ref[c] = r;
    }
    __device__ __forceinline__ void store(float4* o) const {
        for (int a = 0; a < 3; ++a) {
            o[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
            o[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
        }
        o[6] = make_float4(__uint_as_float(ref[0]), __uint_as_float(ref[1]), __uint_as_float(ref[2]),
                           __uint_as_float(ref[3]));
        o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
};

__global__ void k_accel_wide(const clr_lbvh_inode* __restrict__ inodes, const clr_lbvh_leaf* __restrict__ leaves,
                             uint32_t n, float4* __restrict__ wnodes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 >= n) return;
    WideSlots w;
    w.clear();
    const uint4 c = reinterpret_cast<const uint4*>(inodes + i)[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const bool leaf = ((c.x >> (8 * k)) & 0xFFu) != 0;
        const uint32_t ci = k ? c.z : c.y;
        if (leaf) {
            if (ci >= n) continue;  // malformed nodes leave the slot empty
            const float4* p = reinterpret_cast<const float4*>(leaves + ci);
            w.put(2 * k, p[0], p[1], leaves[ci].objIdx | dev::kLeafRef);
            continue;
        }
        if (ci + 1 >= n) continue;
        const uint4 cc = reinterpret_cast<const uint4*>(inodes + ci)[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const bool gl = ((cc.x >> (8 * m)) & 0xFFu) != 0;
            const uint32_t gi = m ? cc.z : cc.y;
            if (gl ? gi >= n : gi + 1 >= n) continue;
            const float4* p = gl ? reinterpret_cast<const float4*>(leaves + gi)
                                 : reinterpret_cast<const float4*>(inodes + gi);
            w.put(2 * k + m, p[0], p[1], gl ? (leaves[gi].objIdx | dev::kLeafRef) : gi);
        }
    }
    w.store(wnodes + 8 * (size_t)i);
}

// The same records for a CPU midpoint BVH, by node number (leaf node slots
// stay unused).
__global__ void k_accel_wide_bvh(const float4* __restrict__ bvh, uint32_t n, float4* __restrict__ wnodes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t numNodes = 2 * n - 1;
    if (i >= numNodes) return;
    const uint2 c = *reinterpret_cast<const uint2*>(bvh + 3 * (size_t)i + 2);
    if (c.y == kEmpty || c.x >= numNodes || c.y >= numNodes) return;
    WideSlots w;
    w.clear();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint32_t ci = k ? c.y : c.x;
        const float4* p = bvh + 3 * (size_t)ci;
        const uint2 cc = *reinterpret_cast<const uint2*>(p + 2);
        if (cc.y == kEmpty) {  // a leaf: its face
            w.put(2 * k, p[0], p[1], cc.x | dev::kLeafRef);
            continue;
        }
        if (cc.x >= numNodes || cc.y >= numNodes) continue;  // malformed: the binary walk flags it
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const uint32_t gi = m ? cc.y : cc.x;
            const float4* q = bvh + 3 * (size_t)gi;
            const uint2 gc = *reinterpret_cast<const uint2*>(q + 2);
            w.put(2 * k + m, q[0], q[1], gc.y == kEmpty ? (gc.x | dev::kLeafRef) : gi);
        }
    }
    w.store(wnodes + 8 * (size_t)i);
}

// Ordered traversal, grid-stride over the rays with a fixed grid (waves per
// CU x CUs). Each lane walks two rays at once, stepping them in turn, so one
// ray's node fetch is in flight while the other's boxes and triangles are
// tested (the walk is bound by dependent load latency). Each walker's stack
// keeps kOrdLds / 2 slots in LDS ([slot][lane]) and the rest in its part of
// `spill`.
constexpr int kOrdLds = 32;
constexpr int kOrdBlock = 128;
constexpr int kOrdHalf = kOrdLds / 2;

template <bool STATS, bool WIDE>
__global__ __launch_bounds__(kOrdBlock) void k_trace_ordered(
    const clr_lbvh_inode* __restrict__ inodes, const clr_lbvh_leaf* __restrict__ leaves, uint32_t n,
    const clr_float3* __restrict__ verts, const clr_face* __restrict__ faces, const float* __restrict__ rays,
    uint32_t numRays, uint32_t* __restrict__ hitFace, float* __restrict__ hitT, uint32_t* __restrict__ stats,
    uint32_t* __restrict__ status, uint32_t* __restrict__ spill, const float4* __restrict__ anodes,
    const float4* __restrict__ atris, const float4* __restrict__ bvh, const float4* __restrict__ wnodes) {
    __shared__ uint32_t s_stack[kOrdLds * kOrdBlock];
    const uint32_t gid = blockIdx.x * kOrdBlock + threadIdx.x;
    const uint32_t lanes = gridDim.x * kOrdBlock;
    using S = dev::SpillStack<kOrdHalf>;
    const S stk0{s_stack + threadIdx.x, kOrdBlock, spill + gid, lanes};
    const S stk1{s_stack + kOrdHalf * kOrdBlock + threadIdx.x, kOrdBlock,
                 spill + (size_t)(dev::kStackPairs - kOrdHalf) * lanes + gid, lanes};
    const dev::GeomScene g{inodes, leaves, n, verts, faces, anodes, atris, bvh, wnodes};
    using W = std::conditional_t<WIDE, dev::Wide<STATS, dev::NoAlpha, S, dev::kStackPairs>,
                                 dev::Ordered<STATS, dev::NoAlpha, S, true>>;
    W w0(g, stk0, dev::NoAlpha()), w1(g, stk1, dev::NoAlpha());
    const size_t R = numRays;
    bool over = false;
    uint32_t next = gid, r0 = kEmpty, r1 = kEmpty;  // this lane's rays: gid, gid + lanes, ...
    auto start = [&](W& w, uint32_t& r) {
        if (next < numRays) {
            r = next;
            next += lanes;
            const dev::V3 o = {rays[r], rays[R + r], rays[2 * R + r]};
            const dev::V3 d = {rays[3 * R + r], rays[4 * R + r], rays[5 * R + r]};
            w.begin(o, d, rays[6 * R + r], false);
        } else {
            r = kEmpty;
            w.done = true;
        }
    };
    auto finish = [&](const W& w, uint32_t r) {
        const bool hit = w.best != kEmpty && !isinf(w.t);
        hitFace[r] = hit ? w.best : kEmpty;
        hitT[r] = w.t;
        if (w.over) over = true;
        if constexpr (STATS) {
            stats[2 * (size_t)r] = w.st.visits;
            stats[2 * (size_t)r + 1] = w.st.tests;
        }
    };
    start(w0, r0);
    start(w1, r1);
    while (r0 != kEmpty || r1 != kEmpty) {
        if (r0 != kEmpty) {
            if (!w0.done) w0.step();
            if (w0.done) {
                finish(w0, r0);
                start(w0, r0);
            }
        }
        if (r1 != kEmpty) {
            if (!w1.done) w1.step();
            if (w1.done) {
                finish(w1, r1);
                start(w1, r1);
            }
        }
    }
    if (over) atomicOr(status, kStatusStackOverflow);
}

}  // namespace

size_t accelNodeBytes(uint32_t n, bool midpoint) {
    const size_t slots = midpoint ? 2 * (size_t)n - 1 : (size_t)std::max(n, 2u) - 1;
    return (slots * 64 + 255) & ~size_t(255);
}

size_t accelWideBytes(uint32_t n, bool midpoint) { return 2 * accelNodeBytes(n, midpoint); }

void launchAccelWide(const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh, uint32_t n,
                     float4* wnodes, hipStream_t stream) {
    if (n < 2) return;
    if (bvh)
        k_accel_wide_bvh<<<(2 * n - 1 + 255) / 256, 256, 0, stream>>>(reinterpret_cast<const float4*>(bvh), n,
                                                                       wnodes);
    else
        k_accel_wide<<<(n + 255) / 256, 256, 0, stream>>>(inodes, leaves, n, wnodes);
}

int buildAccel(clr_ctx* ctx, const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh,
               uint32_t n, const clr_float3* vertices, const clr_face* faces, const float4** anodes,
               const float4** atris, hipStream_t stream, const float4** wnodes) {
    const size_t nodeBytes = accelNodeBytes(n, bvh != nullptr);
    const size_t triBytes = ((size_t)n * 48 + 255) & ~size_t(255);
    const size_t need = nodeBytes + triBytes + (wnodes ? accelWideBytes(n, bvh != nullptr) : 0);
    if (ctx->accelBytes < need) {
        int rc = hipCheck(ctx, hipDeviceSynchronize(), "accel regrow sync");
        if (rc) return rc;
        if (ctx->accel) (void)hipFree(ctx->accel);
        ctx->accel = nullptr;
        ctx->accelBytes = 0;
        hipError_t e = hipMalloc(&ctx->accel, need);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "traversal layout of " + std::to_string(need) + " bytes");
        }
        ctx->accelBytes = need;
    }
    float4* an = static_cast<float4*>(ctx->accel);
    float4* at = reinterpret_cast<float4*>(static_cast<char*>(ctx->accel) + nodeBytes);
    launchAccelBuild(inodes, leaves, bvh, n, vertices, faces, an, at, stream);
    *anodes = an;
    *atris = at;
    if (wnodes) {
        float4* wn = reinterpret_cast<float4*>(static_cast<char*>(ctx->accel) + nodeBytes + triBytes);
        launchAccelWide(inodes, leaves, bvh, n, wn, stream);
        *wnodes = wn;
    }
    return hipCheck(ctx, hipGetLastError(), "traversal layout launch");
}

void launchAccelBuild(const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh,
                      uint32_t n, const clr_float3* vertices, const clr_face* faces, float4* anodes, float4* atris,
                      hipStream_t stream) {
    if (bvh)
        k_accel_build_bvh<<<(2 * n - 1 + 255) / 256, 256, 0, stream>>>(reinterpret_cast<const float4*>(bvh), n,
                                                                        vertices, faces, anodes, atris);
    else
        k_accel_build<<<(n + 255) / 256, 256, 0, stream>>>(inodes, leaves, n, vertices, faces, anodes, atris);
}

int traceOrdered(clr_ctx* ctx, const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh,
                 uint32_t n, const clr_float3* vertices, const clr_face* faces, const float* rays, uint32_t numRays,
                 uint32_t* hitFace, float* hitT, uint32_t* stats, hipStream_t stream, bool wide) {
    if (numRays == 0) return CLR_OK;
    int rc = ensureStatus(ctx);
    if (rc) return rc;
    int cus = 0;
    rc = hipCheck(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "CU count");
    if (rc) return rc;
    // 16 waves per CU (LDS: 16 KiB per 128-thread block).
    const uint32_t grid = std::max(1u, std::min((uint32_t)cus * 8u, (numRays + kOrdBlock - 1) / kOrdBlock));
    const size_t lanes = (size_t)grid * kOrdBlock;
    rc = ensureTraceSpill(ctx, lanes * 2 * (dev::kStackPairs - kOrdHalf) * sizeof(uint32_t));
    if (rc) return rc;
    const float4 *anodes = nullptr, *atris = nullptr, *wnodes = nullptr;
    rc = buildAccel(ctx, inodes, leaves, bvh, n, vertices, faces, &anodes, &atris, stream, wide ? &wnodes : nullptr);
    if (rc) return rc;
    const float4* b4 = reinterpret_cast<const float4*>(bvh);
    auto launch = [&](auto kernel) {
        kernel<<<grid, kOrdBlock, 0, stream>>>(inodes, leaves, n, vertices, faces, rays, numRays, hitFace, hitT, stats,
                                               ctx->devStatus, ctx->traceSpill, anodes, atris, b4, wnodes);
    };
    if (wide) {
        if (stats) launch(k_trace_ordered<true, true>);
        else launch(k_trace_ordered<false, true>);
    } else {
        if (stats) launch(k_trace_ordered<true, false>);
        else launch(k_trace_ordered<false, false>);
    }
    return hipCheck(ctx, hipGetLastError(), "trace launch");
}

int traceClosest(clr_ctx* ctx, const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh,
                 uint32_t n, const clr_float3* vertices, const clr_face* faces, const float* rays,
                 uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats,
                 hipStream_t stream) {
    if (numRays == 0) return CLR_OK;
    int rc = ensureStatus(ctx);
    if (rc) return rc;
    const unsigned grid = (numRays + kTraceBlock - 1) / kTraceBlock;
    const float4* b4 = reinterpret_cast<const float4*>(bvh);
    if (stats)
        k_trace_closest<true><<<grid, kTraceBlock, 0, stream>>>(
            inodes, leaves, n, vertices, faces, rays, numRays, hitFace, hitT, stats, ctx->devStatus, b4);
    else
        k_trace_closest<false><<<grid, kTraceBlock, 0, stream>>>(
            inodes, leaves, n, vertices, faces, rays, numRays, hitFace, hitT, stats, ctx->devStatus, b4);
    return hipCheck(ctx, hipGetLastError(), "trace launch");
}

}  // namespace clr
