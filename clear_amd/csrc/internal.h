// Internal declarations shared by the HIP translation units of libclear_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/clear_amd.h"

struct clr_ctx {
    int device = 0;
    std::string lastError;
    // Scratch pool owned by the context, regrown on demand (the reference's
    // m_bufferGenericPool, LBVHBuilder.cpp:130).
    void* pool = nullptr;
    size_t poolBytes = 0;
    // Profiling of the LBVH build stages.
    bool profiling = false;
    hipEvent_t events[8] = {};
    int numEvents = 0;
    bool haveProfile = false;
    hipStream_t profStream = nullptr;
};

namespace clr {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// Records `msg` as the context's last error and returns `status`.
int fail(clr_ctx* ctx, int status, const std::string& msg);
// Checks a HIP status; on failure records it and returns CLR_ERR_HIP.
int hipCheck(clr_ctx* ctx, hipError_t e, const char* what);
// Makes sure the pool holds at least `bytes`; synchronises the device when it
// has to regrow.
int ensurePool(clr_ctx* ctx, size_t bytes);

// Bump allocator over the pool (256-B aligned sub-buffers).
struct Carve {
    char* base;
    size_t off = 0;
    explicit Carve(void* b) : base(static_cast<char*>(b)) {}
    template <typename T>
    T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += count * sizeof(T);
        return p;
    }
};

// ---- LBVH build (lbvh.hip) ----
struct LbvhInput {
    const float* tris = nullptr;           // soup [n][9], or
    const clr_float3* vertices = nullptr;  // drop-in vertices + faces
    const clr_face* faces = nullptr;
};
size_t lbvhScratchBytes(uint32_t n, uint32_t bitsPerDim);
int lbvhBuild(clr_ctx* ctx, const LbvhInput& in, uint32_t n, uint32_t bitsPerDim,
              clr_lbvh_inode* inodes, clr_lbvh_leaf* leaves, hipStream_t stream);
size_t lbvhSegScratchBytes(uint32_t totalTris, uint32_t numSegments, uint32_t bitsPerDim);
int lbvhBuildSegmented(clr_ctx* ctx, const float* tris, const uint32_t* hostSegOffsets,
                       uint32_t numSegments, uint32_t bitsPerDim, clr_lbvh_inode* inodes,
                       clr_lbvh_leaf* leaves, hipStream_t stream);

// ---- traversal (traverse.hip) ----
int traceClosest(clr_ctx* ctx, const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves,
                 uint32_t n, const clr_float3* vertices, const clr_face* faces, const float* rays,
                 uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats,
                 hipStream_t stream);

}  // namespace clr
