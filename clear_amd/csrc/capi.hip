// C-ABI entry points of libclear_amd (declared in include/clear_amd.h).
//
// Each entry point validates its arguments, makes the context's device
// current for the duration of the call, enqueues work on the caller's stream
// and returns a status; nothing here throws across the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "internal.h"

namespace clr {

int fail(clr_ctx* ctx, int status, const std::string& msg) {
    if (ctx) ctx->lastError = msg;
    return status;
}

int hipCheck(clr_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return CLR_OK;
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(ctx, CLR_ERR_HIP, m);
}

int ensurePool(clr_ctx* ctx, size_t bytes) {
    if (ctx->poolBytes >= bytes) return CLR_OK;
    // The pool may be in use by work still queued on any stream.
    int rc = hipCheck(ctx, hipDeviceSynchronize(), "pool regrow sync");
    if (rc) return rc;
    if (ctx->pool) {
        (void)hipFree(ctx->pool);
        ctx->pool = nullptr;
        ctx->poolBytes = 0;
    }
    size_t want = bytes + bytes / 8;  // headroom against small regrowths
    hipError_t e = hipMalloc(&ctx->pool, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->pool = nullptr;
        return fail(ctx, CLR_ERR_OUT_OF_MEMORY,
                    "scratch pool of " + std::to_string(want) + " bytes: " + hipGetErrorString(e));
    }
    ctx->poolBytes = want;
    return CLR_OK;
}

int ensureExchange(clr_ctx* ctx, size_t nodes) {
    if (nodes == 0) nodes = 1;
    if (ctx->xchgNodes >= nodes) return CLR_OK;
    int rc = hipCheck(ctx, hipDeviceSynchronize(), "exchange regrow sync");
    if (rc) return rc;
    const size_t want = nodes + nodes / 8;
    if (ctx->xchg) (void)hipFree(ctx->xchg);
    ctx->xchg = nullptr;
    ctx->xchgNodes = 0;
    hipError_t e = hipMalloc(&ctx->xchg, want * 6 * sizeof(uint64_t));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->xchg = nullptr;
        return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "hierarchy exchange buffer");
    }
    ctx->xchgNodes = want;
    fillZero(reinterpret_cast<uint32_t*>(ctx->xchg), ctx->xchgNodes * 12, nullptr);
    return hipCheck(ctx, hipDeviceSynchronize(), "exchange clear");
}

int ensureRenderScratch(clr_ctx* ctx, uint32_t numFaces) {
    if (!ctx->renderWords) {
        hipError_t e = hipMalloc(&ctx->renderWords, 64 * sizeof(uint32_t));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ctx->renderWords = nullptr;
            return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "render words");
        }
    }
    if (ctx->faceLightCap < numFaces || !ctx->faceLight) {
        int rc = hipCheck(ctx, hipDeviceSynchronize(), "render scratch regrow sync");
        if (rc) return rc;
        if (ctx->faceLight) (void)hipFree(ctx->faceLight);
        ctx->faceLight = nullptr;
        ctx->faceLightCap = 0;
        size_t want = std::max<size_t>(numFaces, 1);
        hipError_t e = hipMalloc(&ctx->faceLight, want * sizeof(uint32_t));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ctx->faceLight = nullptr;
            return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "face light map");
        }
        ctx->faceLightCap = want;
    }
    return CLR_OK;
}

int ensureTraceSpill(clr_ctx* ctx, size_t bytes) {
    if (ctx->traceSpillBytes >= bytes) return CLR_OK;
    int rc = hipCheck(ctx, hipDeviceSynchronize(), "trace spill regrow sync");
    if (rc) return rc;
    if (ctx->traceSpill) (void)hipFree(ctx->traceSpill);
    ctx->traceSpill = nullptr;
    ctx->traceSpillBytes = 0;
    hipError_t e = hipMalloc(&ctx->traceSpill, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->traceSpill = nullptr;
        return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "trace spill");
    }
    ctx->traceSpillBytes = bytes;
    return CLR_OK;
}

int ensureStatus(clr_ctx* ctx) {
    if (ctx->devStatus) return CLR_OK;
    hipError_t e = hipMalloc(&ctx->devStatus, 256);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->devStatus = nullptr;
        return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "device status word");
    }
    fillZero(ctx->devStatus, 64, nullptr);
    return hipCheck(ctx, hipDeviceSynchronize(), "status clear");
}

}  // namespace clr

using namespace clr;

extern "C" {

const char* clr_error_string(int status) {
    switch (status) {
        case CLR_OK: return "ok";
        case CLR_ERR_INVALID_ARG: return "invalid argument";
        case CLR_ERR_OUT_OF_MEMORY: return "out of device memory";
        case CLR_ERR_HIP: return "HIP runtime error";
        case CLR_ERR_UNSUPPORTED: return "unsupported parameter";
        case CLR_ERR_NO_DEVICE: return "no usable device";
        default: return "unknown status";
    }
}

int clr_create(clr_ctx** out, int device) {
    if (!out) return CLR_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        (void)hipGetLastError();
        return CLR_ERR_NO_DEVICE;
    }
    clr_ctx* ctx = new (std::nothrow) clr_ctx();
    if (!ctx) return CLR_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    DeviceScope ds(device);
    if (ds.err != hipSuccess) {
        delete ctx;
        return CLR_ERR_NO_DEVICE;
    }
    for (auto& row : ctx->events)
        for (auto& e : row)
            if (hipEventCreate(&e) != hipSuccess) {
                delete ctx;
                return CLR_ERR_HIP;
            }
    if (hipDeviceGetAttribute(&ctx->numCUs, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
        (void)hipGetLastError();
        ctx->numCUs = 0;
    }
    // The status and build-epoch words exist from the start, so no entry
    // point allocates them while a caller's stream is being captured.
    if (ensureStatus(ctx) != CLR_OK) {
        clr_destroy(ctx);
        return CLR_ERR_OUT_OF_MEMORY;
    }
    // Device address of the no-allocation build's static array; a failure
    // here only disables that entry point (it then returns the error).
    if (lbvhNoAllocInit(ctx) != CLR_OK) ctx->noAllocScratch = nullptr;
    *out = ctx;
    return CLR_OK;
}

int clr_destroy(clr_ctx* ctx) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    {
        DeviceScope ds(ctx->device);
        (void)hipDeviceSynchronize();
        if (ctx->pool) (void)hipFree(ctx->pool);
        if (ctx->xchg) (void)hipFree(ctx->xchg);
        if (ctx->devStatus) (void)hipFree(ctx->devStatus);
        if (ctx->faceLight) (void)hipFree(ctx->faceLight);
        if (ctx->renderWords) (void)hipFree(ctx->renderWords);
        if (ctx->wfPool) (void)hipFree(ctx->wfPool);
        if (ctx->traceSpill) (void)hipFree(ctx->traceSpill);
        if (ctx->accel) (void)hipFree(ctx->accel);
        if (ctx->adjTable) (void)hipFree(ctx->adjTable);
        if (ctx->batch) (void)hipFree(ctx->batch);
        destroyWfState(ctx);
        for (auto& row : ctx->events)
            for (auto& e : row)
                if (e) (void)hipEventDestroy(e);
    }
    delete ctx;
    return CLR_OK;
}

const char* clr_last_error(const clr_ctx* ctx) { return ctx ? ctx->lastError.c_str() : ""; }

int clr_scratch_bytes(const clr_ctx* ctx, uint64_t* bytes) {
    if (!ctx || !bytes) return CLR_ERR_INVALID_ARG;
    *bytes = ctx->poolBytes;
    return CLR_OK;
}

int clr_device_status(clr_ctx* ctx, uint32_t* flags, int clear) {
    if (!ctx || !flags) return CLR_ERR_INVALID_ARG;
    *flags = 0;
    if (!ctx->devStatus) return CLR_OK;
    DeviceScope ds(ctx->device);
    int rc = hipCheck(ctx, hipDeviceSynchronize(), "status sync");
    if (rc) return rc;
    rc = hipCheck(ctx, hipMemcpy(flags, ctx->devStatus, 4, hipMemcpyDeviceToHost), "status read");
    if (rc) return rc;
    if (clear) {
        fillZero(ctx->devStatus, 1, nullptr);
        rc = hipCheck(ctx, hipDeviceSynchronize(), "status clear");
    }
    return rc;
}

int clr_set_profiling(clr_ctx* ctx, int enable) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    ctx->profiling = enable != 0;
    ctx->profiledBuilds = 0;
    return CLR_OK;
}

int clr_lbvh_last_profile(clr_ctx* ctx, float* ms, int maxStages, int* numStages) {
    if (!ctx || !ms || !numStages || maxStages < 0) return CLR_ERR_INVALID_ARG;
    const int builds = ctx->profiledBuilds;
    if (builds == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "no profiled build recorded");
    DeviceScope ds(ctx->device);
    const int stages = clr_ctx::kStageEvents - 1;
    *numStages = stages;
    for (int k = 0; k < stages && k < maxStages; ++k) ms[k] = 0.0f;
    for (int b = 0; b < builds; ++b) {
        int rc = hipCheck(ctx, hipEventSynchronize(ctx->events[b][stages]), "profile sync");
        if (rc) return rc;
        for (int k = 0; k < stages && k < maxStages; ++k) {
            float v = 0.0f;
            rc = hipCheck(ctx, hipEventElapsedTime(&v, ctx->events[b][k], ctx->events[b][k + 1]),
                          "profile elapsed");
            if (rc) return rc;
            ms[k] += v / (float)builds;
        }
    }
    return CLR_OK;
}

static int checkBuildArgs(clr_ctx* ctx, uint32_t n, uint32_t bits, const void* inodes,
                          const void* leaves) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (n == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "numFaces must be >= 1");
    if (n >= (1u << 30)) return fail(ctx, CLR_ERR_UNSUPPORTED, "numFaces must be < 2^30");
    if (bits < 1 || bits > 21)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "numBitsPerDim must be in [1, 21]");
    if (!leaves || (n > 1 && !inodes))
        return fail(ctx, CLR_ERR_INVALID_ARG, "null node buffer");
    if (!aligned16(leaves) || (inodes && !aligned16(inodes)))
        return fail(ctx, CLR_ERR_INVALID_ARG, "node buffers must be 16-byte aligned");
    return CLR_OK;
}

int clr_lbvh_build(clr_ctx* ctx, const clr_float3* vertices, const clr_face* faces,
                   uint32_t numFaces, uint32_t numBitsPerDim, clr_lbvh_inode* internalNodes,
                   clr_lbvh_leaf* leafNodes, void* stream) {
    int rc = checkBuildArgs(ctx, numFaces, numBitsPerDim, internalNodes, leafNodes);
    if (rc) return rc;
    if (!vertices || !faces) return fail(ctx, CLR_ERR_INVALID_ARG, "null vertices/faces");
    if (!aligned16(vertices)) return fail(ctx, CLR_ERR_INVALID_ARG, "vertices must be 16-B aligned");
    if ((reinterpret_cast<uintptr_t>(faces) & 7u) != 0)
        return fail(ctx, CLR_ERR_INVALID_ARG, "faces must be 8-B aligned");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    LbvhInput in;
    in.vertices = vertices;
    in.faces = faces;
    return lbvhBuild(ctx, in, numFaces, numBitsPerDim, internalNodes, leafNodes,
                     static_cast<hipStream_t>(stream));
}

int clr_lbvh_build_soup(clr_ctx* ctx, const float* tris, uint32_t numTris, uint32_t numBitsPerDim,
                        clr_lbvh_inode* internalNodes, clr_lbvh_leaf* leafNodes, void* stream) {
    int rc = checkBuildArgs(ctx, numTris, numBitsPerDim, internalNodes, leafNodes);
    if (rc) return rc;
    if (!tris) return fail(ctx, CLR_ERR_INVALID_ARG, "null triangle soup");
    if ((reinterpret_cast<uintptr_t>(tris) & 3u) != 0)
        return fail(ctx, CLR_ERR_INVALID_ARG, "soup must be 4-B aligned");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    LbvhInput in;
    in.tris = tris;
    return lbvhBuild(ctx, in, numTris, numBitsPerDim, internalNodes, leafNodes,
                     static_cast<hipStream_t>(stream));
}

int clr_lbvh_build_soup_noalloc(clr_ctx* ctx, const float* tris, uint32_t numTris, uint32_t numBitsPerDim,
                                clr_lbvh_inode* internalNodes, clr_lbvh_leaf* leafNodes, void* stream) {
    int rc = checkBuildArgs(ctx, numTris, numBitsPerDim, internalNodes, leafNodes);
    if (rc) return rc;
    if (!tris) return fail(ctx, CLR_ERR_INVALID_ARG, "null triangle soup");
    if ((reinterpret_cast<uintptr_t>(tris) & 3u) != 0)
        return fail(ctx, CLR_ERR_INVALID_ARG, "soup must be 4-B aligned");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return lbvhBuildSoupNoAlloc(ctx, tris, numTris, numBitsPerDim, internalNodes, leafNodes,
                                static_cast<hipStream_t>(stream));
}

uint32_t clr_lbvh_noalloc_max_tris(void) { return lbvhNoAllocMaxTris(); }

int clr_lbvh_reserve(clr_ctx* ctx, uint32_t maxFaces, uint32_t numBitsPerDim) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (maxFaces == 0 || maxFaces >= (1u << 30) || numBitsPerDim < 1 || numBitsPerDim > 21)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "reserve size out of range");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    const int rc = ensurePool(ctx, lbvhScratchBytes(maxFaces, numBitsPerDim));
    return rc ? rc : ensureExchange(ctx, maxFaces);
}

int clr_lbvh_reserve_segmented(clr_ctx* ctx, uint32_t maxTris, uint32_t numSegments, uint32_t numBitsPerDim) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (maxTris == 0 || maxTris >= (1u << 30) || numSegments == 0 || numSegments > maxTris || numBitsPerDim < 1 ||
        numBitsPerDim > 10)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "reserve size out of range");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    const int rc = ensurePool(ctx, lbvhSegScratchBytes(maxTris, numSegments, numBitsPerDim));
    return rc ? rc : ensureExchange(ctx, maxTris);
}

int clr_lbvh_build_segmented(clr_ctx* ctx, const float* tris, const uint32_t* segOffsets,
                             uint32_t numSegments, uint32_t numTris, uint32_t numBitsPerDim,
                             clr_lbvh_inode* internalNodes, clr_lbvh_leaf* leafNodes,
                             void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (!tris || !segOffsets || numSegments == 0 || !internalNodes || !leafNodes)
        return fail(ctx, CLR_ERR_INVALID_ARG, "null argument or zero segments");
    if (numBitsPerDim < 1 || numBitsPerDim > 10)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "segmented build supports numBitsPerDim 1..10");
    if (numTris < numSegments) return fail(ctx, CLR_ERR_INVALID_ARG, "every segment needs >= 1 triangle");
    if (numTris >= (1u << 30)) return fail(ctx, CLR_ERR_UNSUPPORTED, "total triangles must be < 2^30");
    if (!aligned16(internalNodes) || !aligned16(leafNodes))
        return fail(ctx, CLR_ERR_INVALID_ARG, "node buffers must be 16-byte aligned");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return lbvhBuildSegmented(ctx, tris, segOffsets, numSegments, numTris, numBitsPerDim, internalNodes,
                              leafNodes, static_cast<hipStream_t>(stream));
}

int clr_trace_closest(clr_ctx* ctx, const clr_lbvh_inode* internalNodes,
                      const clr_lbvh_leaf* leafNodes, uint32_t numFaces,
                      const clr_float3* vertices, const clr_face* faces, const float* rays,
                      uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats,
                      void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (numFaces == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "numFaces must be >= 1");
    if (!leafNodes || (numFaces > 1 && !internalNodes) || !vertices || !faces)
        return fail(ctx, CLR_ERR_INVALID_ARG, "null scene buffer");
    if (numRays > 0 && (!rays || !hitFace || !hitT))
        return fail(ctx, CLR_ERR_INVALID_ARG, "null ray buffer");
    if (!aligned16(leafNodes) || (internalNodes && !aligned16(internalNodes)) ||
        !aligned16(vertices))
        return fail(ctx, CLR_ERR_INVALID_ARG, "nodes and vertices must be 16-byte aligned");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return traceClosest(ctx, internalNodes, leafNodes, nullptr, numFaces, vertices, faces, rays, numRays,
                        hitFace, hitT, stats, static_cast<hipStream_t>(stream));
}

int clr_trace_closest_ordered(clr_ctx* ctx, const clr_lbvh_inode* internalNodes,
                              const clr_lbvh_leaf* leafNodes, uint32_t numFaces,
                              const clr_float3* vertices, const clr_face* faces, const float* rays,
                              uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats,
                              void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (numFaces == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "numFaces must be >= 1");
    if (!leafNodes || (numFaces > 1 && !internalNodes) || !vertices || !faces)
        return fail(ctx, CLR_ERR_INVALID_ARG, "null scene buffer");
    if (numRays > 0 && (!rays || !hitFace || !hitT))
        return fail(ctx, CLR_ERR_INVALID_ARG, "null ray buffer");
    if (!aligned16(leafNodes) || (internalNodes && !aligned16(internalNodes)) ||
        !aligned16(vertices))
        return fail(ctx, CLR_ERR_INVALID_ARG, "nodes and vertices must be 16-byte aligned");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return traceOrdered(ctx, internalNodes, leafNodes, nullptr, numFaces, vertices, faces, rays, numRays,
                        hitFace, hitT, stats, static_cast<hipStream_t>(stream));
}

int clr_trace_closest_wide(clr_ctx* ctx, const clr_lbvh_inode* internalNodes, const clr_lbvh_leaf* leafNodes,
                           uint32_t numFaces, const clr_float3* vertices, const clr_face* faces, const float* rays,
                           uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats, void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (numFaces == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "numFaces must be >= 1");
    if (!leafNodes || (numFaces > 1 && !internalNodes) || !vertices || !faces)
        return fail(ctx, CLR_ERR_INVALID_ARG, "null scene buffer");
    if (numRays > 0 && (!rays || !hitFace || !hitT))
        return fail(ctx, CLR_ERR_INVALID_ARG, "null ray buffer");
    if (!aligned16(leafNodes) || (internalNodes && !aligned16(internalNodes)) ||
        !aligned16(vertices))
        return fail(ctx, CLR_ERR_INVALID_ARG, "nodes and vertices must be 16-byte aligned");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return traceOrdered(ctx, internalNodes, leafNodes, nullptr, numFaces, vertices, faces, rays, numRays,
                        hitFace, hitT, stats, static_cast<hipStream_t>(stream), true);
}

int clr_trace_closest_bvh(clr_ctx* ctx, const clr_bvh_node* nodes, uint32_t numFaces, const clr_float3* vertices,
                          const clr_face* faces, const float* rays, uint32_t numRays, uint32_t* hitFace,
                          float* hitT, uint32_t* stats, int ordered, void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (numFaces == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "numFaces must be >= 1");
    if (numFaces >= (1u << 31)) return fail(ctx, CLR_ERR_UNSUPPORTED, "numFaces must be < 2^31");
    if (!nodes || !vertices || !faces) return fail(ctx, CLR_ERR_INVALID_ARG, "null scene buffer");
    if (numRays > 0 && (!rays || !hitFace || !hitT))
        return fail(ctx, CLR_ERR_INVALID_ARG, "null ray buffer");
    if (!aligned16(nodes) || !aligned16(vertices))
        return fail(ctx, CLR_ERR_INVALID_ARG, "nodes and vertices must be 16-byte aligned");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (ordered)
        return traceOrdered(ctx, nullptr, nullptr, nodes, numFaces, vertices, faces, rays, numRays, hitFace, hitT,
                            stats, st);
    return traceClosest(ctx, nullptr, nullptr, nodes, numFaces, vertices, faces, rays, numRays, hitFace, hitT,
                        stats, st);
}

}  // extern "C"

extern "C" {

int clr_render(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params,
               uint32_t* randStates, clr_float3* pixels, int mode, void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (!randStates || !pixels) return fail(ctx, CLR_ERR_INVALID_ARG, "null render argument");
    int rc = checkRenderArgs(ctx, scene, params);
    if (rc) return rc;
    if (!aligned16(randStates) || !aligned16(pixels))
        return fail(ctx, CLR_ERR_INVALID_ARG, "buffers must be 16-byte aligned");
    if (mode == CLR_RENDER_WAVEFRONT && params->maxDepth > 255)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "wavefront renderer supports maxDepth <= 255");
    if (params->samplesPerPixel == 0) return CLR_OK;
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (mode == CLR_RENDER_MEGAKERNEL) return renderMegakernel(ctx, *scene, *params, randStates, pixels, st);
    if (mode == CLR_RENDER_WAVEFRONT) return renderWavefront(ctx, *scene, *params, randStates, pixels, st);
    return fail(ctx, CLR_ERR_INVALID_ARG, "unknown render mode");
}

int clr_bvh_visualizer(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params,
                       clr_float3* pixels, void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (!pixels || !params) return fail(ctx, CLR_ERR_INVALID_ARG, "null visualizer argument");
    clr_render_params rp = *params;
    rp.maxDepth = 1;  // not used here
    int rc = checkRenderArgs(ctx, scene, &rp);
    if (rc) return rc;
    if (!aligned16(pixels)) return fail(ctx, CLR_ERR_INVALID_ARG, "buffers must be 16-byte aligned");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return renderBvhVisualizer(ctx, *scene, rp, pixels, static_cast<hipStream_t>(stream));
}

int clr_render_batch(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                     const clr_render_params* params, uint32_t* randStates, clr_float3* pixels,
                     void* stream) {
    return clr_render_batch_mode(ctx, scenes, numScenes, params, randStates, pixels, CLR_RENDER_WAVEFRONT, stream);
}

int clr_render_batch_mode(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                          const clr_render_params* params, uint32_t* randStates, clr_float3* pixels, int mode,
                          void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (mode != CLR_RENDER_MEGAKERNEL && mode != CLR_RENDER_WAVEFRONT)
        return fail(ctx, CLR_ERR_INVALID_ARG, "unknown render mode");
    if (!scenes || !randStates || !pixels || !params) return fail(ctx, CLR_ERR_INVALID_ARG, "null render argument");
    if (numScenes == 0) return CLR_OK;
    for (uint32_t s = 0; s < numScenes; ++s) {
        int rc = checkRenderArgs(ctx, scenes + s, params);
        if (rc) return rc;
    }
    if (!aligned16(randStates) || !aligned16(pixels))
        return fail(ctx, CLR_ERR_INVALID_ARG, "buffers must be 16-byte aligned");
    if (mode == CLR_RENDER_WAVEFRONT && params->maxDepth > 255)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "wavefront renderer supports maxDepth <= 255");
    if (params->flags & CLR_RENDER_REFERENCE_TRAVERSAL)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "batched render uses the ordered traversal only");
    const uint64_t tiles = (uint64_t)((params->width + 7) / 8) * ((params->height + 7) / 8) * 64;
    if ((uint64_t)params->width * params->height * numScenes >= (1ull << 31) || tiles * numScenes >= (1ull << 31))
        return fail(ctx, CLR_ERR_UNSUPPORTED, "batch too large (scenes x pixels >= 2^31)");
    if (params->samplesPerPixel == 0) return CLR_OK;
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    if (mode == CLR_RENDER_MEGAKERNEL)
        return renderMegakernelBatch(ctx, scenes, numScenes, *params, randStates, pixels,
                                     static_cast<hipStream_t>(stream));
    return renderWavefrontBatch(ctx, scenes, numScenes, *params, randStates, pixels, static_cast<hipStream_t>(stream));
}

int clr_render_adjoint(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params,
                       const clr_diff_params* diff, const uint32_t* randStates, const clr_float3* dL_dimage,
                       float* gAlbedo, float* gRoughness, float* gEmission, void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    int rc = checkRenderArgs(ctx, scene, params);
    if (rc) return rc;
    if (!diff || !randStates || !dL_dimage) return fail(ctx, CLR_ERR_INVALID_ARG, "null adjoint argument");
    if (!aligned16(randStates) || !aligned16(dL_dimage))
        return fail(ctx, CLR_ERR_INVALID_ARG, "buffers must be 16-byte aligned");
    if (diff->numMaterials > 256) return fail(ctx, CLR_ERR_UNSUPPORTED, "at most 256 differentiable materials");
    if (diff->numMaterials && (!diff->matOffsets || !gAlbedo || !gRoughness))
        return fail(ctx, CLR_ERR_INVALID_ARG, "null material list or gradient output");
    for (uint32_t k = 0; k < diff->numMaterials; ++k)
        if (diff->matOffsets[k] > 0xFFFFu) return fail(ctx, CLR_ERR_INVALID_ARG, "material offset beyond 16 bits");
    if (diff->lightProp != 0xFFFFFFFFu && (diff->lightProp > 0xFFFFu || !gEmission))
        return fail(ctx, CLR_ERR_INVALID_ARG, "emitter offset beyond 16 bits or null emission gradient");
    if (params->maxDepth > 15) return fail(ctx, CLR_ERR_UNSUPPORTED, "adjoint supports maxDepth <= 15");
    if (params->samplesPerPixel == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "samplesPerPixel must be >= 1");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return renderAdjoint(ctx, *scene, *params, *diff, randStates, dL_dimage, gAlbedo, gRoughness,
                         diff->lightProp != 0xFFFFFFFFu ? gEmission : nullptr, static_cast<hipStream_t>(stream));
}

int clr_wavefront_get_config(clr_ctx* ctx, clr_wavefront_config* config) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (!config) return fail(ctx, CLR_ERR_INVALID_ARG, "null config");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return wavefrontGetConfig(ctx, config);
}

int clr_wavefront_set_config(clr_ctx* ctx, const clr_wavefront_config* config) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (!config) return fail(ctx, CLR_ERR_INVALID_ARG, "null config");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return wavefrontSetConfig(ctx, config);
}

int clr_wavefront_autotune(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params,
                           clr_wavefront_config* config, void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    int rc = checkRenderArgs(ctx, scene, params);
    if (rc) return rc;
    if (params->maxDepth > 255) return fail(ctx, CLR_ERR_UNSUPPORTED, "wavefront renderer supports maxDepth <= 255");
    if (params->samplesPerPixel == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "autotune needs samplesPerPixel >= 1");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return wavefrontAutotune(ctx, scene, 1, false, *params, config, static_cast<hipStream_t>(stream));
}

int clr_wavefront_autotune_batch(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                                 const clr_render_params* params, clr_wavefront_config* config, void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (!scenes || !params || numScenes == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "null or empty batch");
    for (uint32_t s = 0; s < numScenes; ++s) {
        int rc = checkRenderArgs(ctx, scenes + s, params);
        if (rc) return rc;
    }
    if (params->maxDepth > 255) return fail(ctx, CLR_ERR_UNSUPPORTED, "wavefront renderer supports maxDepth <= 255");
    if (params->flags & CLR_RENDER_REFERENCE_TRAVERSAL)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "batched render uses the ordered traversal only");
    if (params->samplesPerPixel == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "autotune needs samplesPerPixel >= 1");
    const uint64_t tiles = (uint64_t)((params->width + 7) / 8) * ((params->height + 7) / 8) * 64;
    if ((uint64_t)params->width * params->height * numScenes >= (1ull << 31) || tiles * numScenes >= (1ull << 31))
        return fail(ctx, CLR_ERR_UNSUPPORTED, "batch too large (scenes x pixels >= 2^31)");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return wavefrontAutotune(ctx, scenes, numScenes, true, *params, config, static_cast<hipStream_t>(stream));
}

}  // extern "C"

namespace clr {
int checkRenderArgs(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params) {
    if (!scene || !params) return fail(ctx, CLR_ERR_INVALID_ARG, "null render argument");
    const clr_scene_dev& s = *scene;
    const bool accelOk = s.bvhNodes ? aligned16(s.bvhNodes) : (s.leafNodes && (s.numFaces == 1 || s.internalNodes));
    if (s.numFaces == 0 || !s.vertices || !s.faces || !accelOk || !s.materialsData || !s.texturesData ||
        !s.otherResources || (s.numLights && !s.lights))
        return fail(ctx, CLR_ERR_INVALID_ARG, "incomplete scene");
    if (s.numFaces >= (1u << 31)) return fail(ctx, CLR_ERR_UNSUPPORTED, "numFaces must be < 2^31");
    if (s.numLights == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "scene has no light");
    if (params->width == 0 || params->height == 0 || params->maxDepth == 0)
        return fail(ctx, CLR_ERR_INVALID_ARG, "empty image or zero bounce cap");
    if ((uint64_t)params->width * params->height >= (1ull << 31))
        return fail(ctx, CLR_ERR_UNSUPPORTED, "image too large");
    if ((params->flags & CLR_RENDER_LIGHT_SAMPLING_ONLY) && (params->flags & CLR_RENDER_BSDF_SAMPLING_ONLY))
        return fail(ctx, CLR_ERR_INVALID_ARG, "LIGHT_SAMPLING_ONLY and BSDF_SAMPLING_ONLY exclude each other");
    if (!aligned16(s.vertices) || (s.leafNodes && !aligned16(s.leafNodes)) ||
        (s.internalNodes && !aligned16(s.internalNodes)) ||
        !aligned16(s.otherResources) || !aligned16(s.texturesData) || !aligned16(s.materialsData))
        return fail(ctx, CLR_ERR_INVALID_ARG, "buffers must be 16-byte aligned");
    return CLR_OK;
}
}  // namespace clr

extern "C" {

int clr_seed_pixel_states(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t* randStates,
                          void* stream) {
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (!randStates || !aligned16(randStates)) return fail(ctx, CLR_ERR_INVALID_ARG, "state buffer");
    if ((uint64_t)width * height >= (1ull << 31)) return fail(ctx, CLR_ERR_UNSUPPORTED, "image too large");
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return seedPixelStates(ctx, width, height, randStates, static_cast<hipStream_t>(stream));
}

}  // extern "C"
