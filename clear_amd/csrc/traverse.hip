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
    float* __restrict__ hitT, uint32_t* __restrict__ stats, uint32_t* __restrict__ status) {
    __shared__ uint32_t s_stack[dev::kStackDepth * kTraceBlock];
    const uint32_t r = blockIdx.x * kTraceBlock + threadIdx.x;
    if (r >= numRays) return;
    const size_t R = numRays;
    dev::V3 o = {rays[r], rays[R + r], rays[2 * R + r]};
    dev::V3 d = {rays[3 * R + r], rays[4 * R + r], rays[5 * R + r]};
    float t = rays[6 * R + r];
    dev::GeomScene g{inodes, leaves, n, verts, faces};
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

}  // namespace

int traceClosest(clr_ctx* ctx, const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves,
                 uint32_t n, const clr_float3* vertices, const clr_face* faces, const float* rays,
                 uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats,
                 hipStream_t stream) {
    if (numRays == 0) return CLR_OK;
    int rc = ensureStatus(ctx);
    if (rc) return rc;
    const unsigned grid = (numRays + kTraceBlock - 1) / kTraceBlock;
    if (stats)
        k_trace_closest<true><<<grid, kTraceBlock, 0, stream>>>(
            inodes, leaves, n, vertices, faces, rays, numRays, hitFace, hitT, stats, ctx->devStatus);
    else
        k_trace_closest<false><<<grid, kTraceBlock, 0, stream>>>(
            inodes, leaves, n, vertices, faces, rays, numRays, hitFace, hitT, stats, ctx->devStatus);
    return hipCheck(ctx, hipGetLastError(), "trace launch");
}

}  // namespace clr
