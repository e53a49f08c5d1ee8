// Internal declarations shared by the HIP translation units of libclear_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/clear_amd.h"

namespace clr {
struct WfState;
namespace dev {
struct RScene;
}
}

struct clr_ctx {
    int device = 0;
    int numCUs = 0;  // compute units of the device (launch sizing)
    std::string lastError;
    // Scratch pool owned by the context, regrown on demand (the reference's
    // m_bufferGenericPool, LBVHBuilder.cpp:130).
    void* pool = nullptr;
    size_t poolBytes = 0;
    // Profiling of the LBVH build stages: one event per stage boundary for up
    // to kMaxProfiledBuilds builds since profiling was (re)enabled.
    static constexpr int kMaxProfiledBuilds = 64;
    static constexpr int kStageEvents = 6;
    bool profiling = false;
    hipEvent_t events[kMaxProfiledBuilds][kStageEvents] = {};
    int profiledBuilds = 0;
    // Box hand-off buffer of the bottom-up hierarchy pass: six 64-bit words
    // (float bits | build epoch) per internal node. Only ever written with
    // epoch-tagged words, zeroed when (re)allocated and when the epoch wraps,
    // so a stale word can never carry the current epoch.
    uint64_t* xchg = nullptr;
    size_t xchgNodes = 0;
    // 64 device words, allocated by clr_create: word kWordStatus holds the
    // device-side failure flags (bounded spin exceeded, stack overflow), read
    // back by clr_device_status; word kWordEpoch the LBVH build epoch, which
    // the build's kernels advance and read on the build's stream.
    uint32_t* devStatus = nullptr;
    // Render scratch: face -> first light index, and 64 small device words
    // (word 0: first light at infinity).
    uint32_t* faceLight = nullptr;
    size_t faceLightCap = 0;
    uint32_t* renderWords = nullptr;
    // Wavefront path-state pool, configuration and captured graph.
    void* wfPool = nullptr;
    size_t wfPoolBytes = 0;
    clr::WfState* wf = nullptr;
    // Batched rendering: scene table, traversal layouts, face-light maps.
    void* batch = nullptr;
    size_t batchBytes = 0;
    // Adjoint: matPtr -> parameter table (65536 x uint16) + 256 offsets.
    void* adjTable = nullptr;
    // Traversal layout of the ordered walk (rebuilt by every call that uses it).
    void* accel = nullptr;
    size_t accelBytes = 0;
    // Traversal-stack overflow region of the ordered trace kernel.
    uint32_t* traceSpill = nullptr;
    size_t traceSpillBytes = 0;
    // Device address of this device's static working array of
    // clr_lbvh_build_soup_noalloc (looked up by clr_create; not owned).
    void* noAllocScratch = nullptr;
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

// Bits of the device status word.
constexpr uint32_t kStatusLookbackSpin = 1u << 0;
constexpr uint32_t kStatusExchangeSpin = 1u << 1;
constexpr uint32_t kStatusStackOverflow = 1u << 2;
constexpr uint32_t kStatusHierarchy = 1u << 3;
constexpr uint32_t kStatusSegments = 1u << 4;
// Upper bound on any device-side wait loop (iterations of one atomic load).
constexpr uint32_t kSpinLimit = 1u << 22;

// Grows the box hand-off buffer to `nodes` entries (synchronous when it has
// to reallocate; the new buffer is zeroed).
int ensureExchange(clr_ctx* ctx, size_t nodes);
// Allocates the device status words if clr_create has not (synchronous).
int ensureStatus(clr_ctx* ctx);
// Indices into clr_ctx::devStatus.
constexpr uint32_t kWordStatus = 0;
constexpr uint32_t kWordEpoch = 1;
constexpr uint32_t kWordMegaTiles = 2;  // next 8x8 tile of a megakernel render

// Bump allocator over the pool (256-B aligned sub-buffers).
// Makes the context's device current for the duration of a C-ABI call and
// restores the caller's device afterwards.
struct DeviceScope {
    int prev = -1;
    hipError_t err = hipSuccess;
    explicit DeviceScope(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) err = hipSetDevice(dev);
    }
    ~DeviceScope() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

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
// Fills `count` 32-bit words with zero on `stream` (kernel, not hipMemset).
void fillZero(uint32_t* words, size_t count, hipStream_t stream);
size_t lbvhSegScratchBytes(uint32_t totalTris, uint32_t numSegments, uint32_t bitsPerDim);
// The build that allocates nothing (clr_lbvh_build_soup_noalloc): looks up
// this device's static working array once per device (with the device
// current) and records it in ctx; the build itself enqueues kernels only.
int lbvhNoAllocInit(clr_ctx* ctx);
uint32_t lbvhNoAllocMaxTris();
int lbvhBuildSoupNoAlloc(clr_ctx* ctx, const float* tris, uint32_t numTris, uint32_t bitsPerDim,
                         clr_lbvh_inode* inodes, clr_lbvh_leaf* leaves, hipStream_t stream);
// segOffsets: device array of numSegments + 1 offsets, ending at numTris.
int lbvhBuildSegmented(clr_ctx* ctx, const float* tris, const uint32_t* segOffsets,
                       uint32_t numSegments, uint32_t numTris, uint32_t bitsPerDim, clr_lbvh_inode* inodes,
                       clr_lbvh_leaf* leaves, hipStream_t stream);

// ---- rendering (render.hip, wavefront.hip) ----
int ensureRenderScratch(clr_ctx* ctx, uint32_t numFaces);
int renderMegakernel(clr_ctx* ctx, const clr_scene_dev& sd, const clr_render_params& rp,
                     uint32_t* states, clr_float3* pixels, hipStream_t stream);
int renderWavefront(clr_ctx* ctx, const clr_scene_dev& sd, const clr_render_params& rp,
                    uint32_t* states, clr_float3* pixels, hipStream_t stream);
// Batches (clr_render_batch, clr_render_batch_mode): the shared set-up and
// the two renderers.
int prepareBatchScenes(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes, hipStream_t stream,
                       dev::RScene** tableOut, dev::RScene* firstOut);
int renderWavefrontBatch(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                         const clr_render_params& rp, uint32_t* states, clr_float3* pixels, hipStream_t stream);
int renderMegakernelBatch(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                          const clr_render_params& rp, uint32_t* states, clr_float3* pixels, hipStream_t stream);
// Whether `stream` is being captured into a graph (by the caller).
int streamCapturing(clr_ctx* ctx, hipStream_t stream, bool& capturing);
int seedPixelStates(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t* states,
                    hipStream_t stream);
int renderAdjoint(clr_ctx* ctx, const clr_scene_dev& sd, const clr_render_params& rp, const clr_diff_params& dp,
                  const uint32_t* states, const clr_float3* dLdI, float* gAlb, float* gSig, float* gEm,
                  hipStream_t stream);
int wavefrontGetConfig(clr_ctx* ctx, clr_wavefront_config* cfg);
int wavefrontSetConfig(clr_ctx* ctx, const clr_wavefront_config* cfg);
int wavefrontAutotune(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes, bool batch,
                      const clr_render_params& rp, clr_wavefront_config* out, hipStream_t stream);
void destroyWfState(clr_ctx* ctx);
int renderBvhVisualizer(clr_ctx* ctx, const clr_scene_dev& sd, const clr_render_params& rp, clr_float3* pixels,
                        hipStream_t stream);
// Argument checks shared by clr_render and the autotuner.
int checkRenderArgs(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params);

// ---- traversal (traverse.hip) ----
int ensureTraceSpill(clr_ctx* ctx, size_t bytes);
// Launches the traversal-layout build into caller-given arrays (from the
// LBVH pair, or from a CPU midpoint BVH when bvh is not null).
void launchAccelBuild(const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh,
                      uint32_t n, const clr_float3* vertices, const clr_face* faces, float4* anodes, float4* atris,
                      hipStream_t stream);
// Bytes of the traversal layout's node array (256-B aligned).
size_t accelNodeBytes(uint32_t n, bool midpoint);
// The renderers' 4-wide records (128 B per node slot) and their build.
size_t accelWideBytes(uint32_t n, bool midpoint);
void launchAccelWide(const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh, uint32_t n,
                     float4* wnodes, hipStream_t stream);
// Face -> first light map and first environment light of one scene.
void launchFaceLights(const clr_light_info* lights, uint32_t numLights, uint32_t numFaces, uint32_t* faceLight,
                      int* envLight, hipStream_t stream);
// Builds the ordered walk's traversal layout from the drop-in nodes on `stream`
// into context memory and returns its two arrays, and with `wnodes` the
// renderers' 4-wide records too.
int buildAccel(clr_ctx* ctx, const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh,
               uint32_t n, const clr_float3* vertices, const clr_face* faces, const float4** anodes,
               const float4** atris, hipStream_t stream, const float4** wnodes = nullptr);
int traceOrdered(clr_ctx* ctx, const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh,
                 uint32_t n, const clr_float3* vertices, const clr_face* faces, const float* rays, uint32_t numRays,
                 uint32_t* hitFace, float* hitT, uint32_t* stats, hipStream_t stream, bool wide = false);
int traceClosest(clr_ctx* ctx, const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, const clr_bvh_node* bvh,
                 uint32_t n, const clr_float3* vertices, const clr_face* faces, const float* rays,
                 uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats,
                 hipStream_t stream);

}  // namespace clr
