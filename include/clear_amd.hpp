/*
 * clear_amd.hpp — C++ wrappers over the C-ABI (clear_amd.h) that keep the
 * calling shape of the reference's host code: the LBVHBuilder class
 * (OpenCL_PathTracer/LBVHBuilder.h:32-75) and the render loop of
 * OpenCL_PathTracer/main.cpp:475-494, with failures raised as exceptions as
 * the reference's cl::Error is (main.cpp:17, 545-551). Header-only; link
 * libclear_amd.so. Streams are hipStream_t values passed as void*
 * (nullptr = the default stream), so this header needs no HIP headers.
 */
#ifndef CLEAR_AMD_HPP
#define CLEAR_AMD_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "clear_amd.h"

namespace clr {

// A failed call: the C-ABI status (CLR_ERR_*) and its message.
class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(clr_ctx* ctx, int rc, const char* call) {
    if (rc == CLR_OK) return;
    std::string msg = std::string(call) + ": " + clr_error_string(rc);
    if (ctx && clr_last_error(ctx)[0]) msg += std::string(": ") + clr_last_error(ctx);
    throw Error(rc, msg);
}

// Owns a clr_ctx on one device (for this path, the reference's cl::Context /
// cl::Device pair). One host thread and one stream at a time.
class Context {
public:
    explicit Context(int device = 0) { check(nullptr, clr_create(&ctx_, device), "clr_create"); }
    ~Context() {
        if (ctx_) clr_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    clr_ctx* get() const { return ctx_; }
    // Device-side failure flags since the last clear (clr_device_status); synchronises.
    uint32_t deviceStatus(bool clear = true) {
        uint32_t flags = 0;
        check(ctx_, clr_device_status(ctx_, &flags, clear ? 1 : 0), "clr_device_status");
        return flags;
    }

private:
    clr_ctx* ctx_ = nullptr;
};

// LBVHBuilder (LBVHBuilder.h:32-75). The constructor reserves scratch for
// numFaces triangles (the reference compiles its kernels and sizes its pool
// there, LBVHBuilder.cpp:19-48); perform enqueues one build on `queue` with
// the reference's arguments: vertices (cl_float3[]), faces (Face[numFaces]),
// numBitsPerDim, and the caller-owned node outputs (InternalNode[numFaces-1],
// LeafNode[numFaces]). With profiling, per-stage GPU times of the builds are
// recorded and lastProfile() returns their mean in ms (the reference prints
// event times, LBVHBuilder.cpp:173-349).
class LBVHBuilder {
public:
    LBVHBuilder(Context& ctx, uint32_t numFaces, uint32_t numBitsPerDim = 10) : ctx_(ctx) {
        if (numFaces) check(ctx_.get(), clr_lbvh_reserve(ctx_.get(), numFaces, numBitsPerDim), "clr_lbvh_reserve");
    }
    void perform(void* queue, const clr_float3* vertices, const clr_face* faces, uint32_t numFaces,
                 uint32_t numBitsPerDim, clr_lbvh_inode* internalNodes, clr_lbvh_leaf* leafNodes,
                 bool profiling = false) {
        if (profiling != profiling_) {
            check(ctx_.get(), clr_set_profiling(ctx_.get(), profiling ? 1 : 0), "clr_set_profiling");
            profiling_ = profiling;
        }
        check(ctx_.get(),
              clr_lbvh_build(ctx_.get(), vertices, faces, numFaces, numBitsPerDim, internalNodes, leafNodes, queue),
              "clr_lbvh_build");
    }
    // The same build from a triangle soup (tris[numTris][9]).
    void performSoup(void* queue, const float* tris, uint32_t numTris, uint32_t numBitsPerDim,
                     clr_lbvh_inode* internalNodes, clr_lbvh_leaf* leafNodes) {
        check(ctx_.get(), clr_lbvh_build_soup(ctx_.get(), tris, numTris, numBitsPerDim, internalNodes, leafNodes, queue),
              "clr_lbvh_build_soup");
    }
    // Mean per-stage times (ms) of the profiled builds; waits for them.
    std::vector<float> lastProfile() {
        float ms[8];
        int n = 0;
        check(ctx_.get(), clr_lbvh_last_profile(ctx_.get(), ms, 8, &n), "clr_lbvh_last_profile");
        return std::vector<float>(ms, ms + n);
    }

private:
    Context& ctx_;
    bool profiling_ = false;
};

// Per-pixel seeding of main.cpp:93-106 (clr_seed_pixel_states).
inline void seedPixelStates(Context& ctx, uint32_t width, uint32_t height, uint32_t* randStates,
                            void* queue = nullptr) {
    check(ctx.get(), clr_seed_pixel_states(ctx.get(), width, height, randStates, queue), "clr_seed_pixel_states");
}

// The render loop of main.cpp:475-494 as one call: adds params.samplesPerPixel
// samples per pixel into `pixels` (undivided, as the reference accumulates)
// and advances `randStates` (clr_render).
inline void render(Context& ctx, const clr_scene_dev& scene, const clr_render_params& params, uint32_t* randStates,
                   clr_float3* pixels, int mode = CLR_RENDER_MEGAKERNEL, void* queue = nullptr) {
    check(ctx.get(), clr_render(ctx.get(), &scene, &params, randStates, pixels, mode, queue), "clr_render");
}

}  // namespace clr

#endif /* CLEAR_AMD_HPP */
