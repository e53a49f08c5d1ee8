/*
 * clear_amd.h — C-ABI of the MI355X (gfx950) LBVH + path-tracer hot path.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns an
 * int status (CLR_OK == 0) and never throws. Device pointers are caller-owned
 * and must live on the context's device. Work is enqueued on `stream`
 * (a hipStream_t passed as void*; NULL = the default stream) and is
 * stream-ordered: no entry point synchronises the host unless its comment says
 * so. A clr_ctx is externally synchronised: one host thread at a time, and
 * its scratch (LBVH pool, exchange words, render and wavefront state) is shared
 * by all its calls, so work enqueued through one context must stay on one
 * stream, or calls on different streams must be ordered by the caller (use
 * one context per concurrent stream).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   OpenCL_PathTracer/LBVHBuilder.h:32-75, LBVHBuilder.cpp:160-375  (LBVHBuilder)
 *   OpenCL_PathTracer/OpenCL/bvh_traversal.cl:182-215                (rayIntersection)
 *   OpenCL_PathTracer/OpenCL/pathtracer.cl:25-197                    (pathtracing)
 *   OpenCL_PathTracer/main.cpp:475-494                               (render loop)
 */
#ifndef CLEAR_AMD_H
#define CLEAR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    CLR_OK = 0,
    CLR_ERR_INVALID_ARG = 1,   /* null pointer, bad size, bad enum          */
    CLR_ERR_OUT_OF_MEMORY = 2, /* scratch pool could not grow               */
    CLR_ERR_HIP = 3,           /* a HIP runtime call failed                 */
    CLR_ERR_UNSUPPORTED = 4,   /* parameter outside the supported range     */
    CLR_ERR_NO_DEVICE = 5      /* no usable gfx950 device at that ordinal   */
};

/* ---- data ABI (byte-compatible with the reference host structs) --------- */

/* cl_float3: 16 bytes, w unused. Vertices, normals and tangents use it
 * (Scene.h:31-33, uploaded as-is at main.cpp:342). */
typedef struct clr_float3 {
    float x, y, z, w;
} clr_float3;

/* Face, 56 bytes (Face.hpp:15-73; scene.cl:22-29). UINT32_MAX = attribute
 * absent; matPtr/lightPtr are byte offsets into materialsData, 0xFFFF = none. */
typedef struct clr_face {
    uint32_t p0, p1, p2;
    uint32_t vn0, vn1, vn2;
    uint32_t vt0, vt1, vt2;
    uint32_t uv0, uv1, uv2;
    uint32_t alphaTexPtr;
    uint16_t matPtr, lightPtr;
} clr_face;

/* LBVH::InternalNode, 48 bytes (LBVHBuilder.h:17-22). Node 0 is the root.
 * isChild[k] != 0 means child k is a leaf index, else an internal index. */
typedef struct clr_lbvh_inode {
    float min[4];
    float max[4];
    uint8_t isChild[2];
    uint8_t dum0[2];
    uint32_t c[2];
    uint8_t dum1[4];
} clr_lbvh_inode;

/* LBVH::LeafNode, 48 bytes (LBVHBuilder.h:25-29). */
typedef struct clr_lbvh_leaf {
    float min[4];
    float max[4];
    uint32_t objIdx;
    uint8_t dum0[12];
} clr_lbvh_leaf;

/* BVHNode of the CPU midpoint BVH, 48 bytes (BVH.hpp:61-68;
 * bvh_traversal.cl:33-37). Node 0 is the root; a leaf has children[1] ==
 * 0xFFFFFFFF and children[0] = its face index. A tree over n faces has
 * 2n - 1 nodes. */
typedef struct clr_bvh_node {
    float min[4];
    float max[4];
    uint32_t children[2];
    uint32_t dum0[2];
} clr_bvh_node;

/* LightInfo, 8 bytes (Scene.h:23-27; scene.cl:32-35). */
typedef struct clr_light_info {
    uint8_t atInfinity;
    uint8_t dum0[3];
    uint32_t reference;
} clr_light_info;

/* ---- context ------------------------------------------------------------ */
typedef struct clr_ctx clr_ctx;

/* Creates a context on HIP device `device` (replaces the LBVHBuilder ctor,
 * LBVHBuilder.cpp:19-48, minus the runtime compile: kernels are prebuilt). */
int clr_create(clr_ctx** ctx, int device);
int clr_destroy(clr_ctx* ctx);
/* Static string for a status code. */
const char* clr_error_string(int status);
/* Detail of the last failure on this context ("" when none). */
const char* clr_last_error(const clr_ctx* ctx);
/* Bytes currently held by the context's scratch pool. */
int clr_scratch_bytes(const clr_ctx* ctx, uint64_t* bytes);
/* Device-side failure flags raised by kernels since the last clear: bit 0 a
 * radix-sort look-back wait, bit 1 a hierarchy hand-off wait exceeded its
 * bound (the result is then wrong), bit 2 a traversal stack overflowed, bit 3
 * the hierarchy met a node with more than two children (unsorted keys), bit 4
 * a segmented build rejected its scene offsets.
 * Synchronises the device. No reference counterpart: the reference has no
 * failure detection (its atomic refit has no bound, lbvh_construction.cl:465). */
int clr_device_status(clr_ctx* ctx, uint32_t* flags, int clear);

/* ---- LBVH build ---------------------------------------------------------- */

/* Drop-in for LBVHBuilder::perform (LBVHBuilder.cpp:160-375).
 *   vertices      clr_float3[]    (device, 16-B stride, as buf_vertices)
 *   faces         clr_face[numFaces] (device)
 *   numBitsPerDim quantisation bits per axis, 1..21 (reference main uses 20;
 *                 10 gives the 30-bit Morton code)
 *   internalNodes clr_lbvh_inode[numFaces-1] (device, written)
 *   leafNodes     clr_lbvh_leaf[numFaces]    (device, written)
 * Unlike the reference there is no host read-back of the scene bounds inside
 * the build (LBVHBuilder.cpp:228-229): the call is fully stream-ordered and
 * may be captured into a hipGraph once the scratch pool is large enough
 * (see clr_lbvh_reserve); every replay is a build of the inputs' contents at
 * that time (the hierarchy's hand-off tags advance on the device).
 * numFaces == 1 writes leafNodes[0] only. */
int clr_lbvh_build(clr_ctx* ctx, const clr_float3* vertices, const clr_face* faces,
                   uint32_t numFaces, uint32_t numBitsPerDim,
                   clr_lbvh_inode* internalNodes, clr_lbvh_leaf* leafNodes, void* stream);

/* Same build from a triangle soup: tris[numTris][9] = p0.xyz p1.xyz p2.xyz
 * (device, 36-B records). Identical output to clr_lbvh_build on the
 * equivalent vertices/faces (vertex 3i+k = corner k of triangle i). */
int clr_lbvh_build_soup(clr_ctx* ctx, const float* tris, uint32_t numTris,
                        uint32_t numBitsPerDim, clr_lbvh_inode* internalNodes,
                        clr_lbvh_leaf* leafNodes, void* stream);

/* clr_lbvh_build_soup that allocates nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under hipGraph stream capture, tight serving loops, harnesses that forbid
 * hipMalloc). Instead of the context's scratch pool it works in a static
 * device array of 104 MiB + 256 B that every device loading the library
 * holds until the process ends, sized for numTris <=
 * clr_lbvh_noalloc_max_tris() (1,048,576) and numBitsPerDim 1..10; larger
 * requests return CLR_ERR_UNSUPPORTED. The call enqueues kernels only: no
 * hipMalloc, no hipFree, no host synchronisation. Every call on one device
 * shares that device's array whatever its context, so calls that overlap on
 * one device (other streams, contexts or threads) race and must be ordered by
 * the caller. Output identical to clr_lbvh_build_soup. */
int clr_lbvh_build_soup_noalloc(clr_ctx* ctx, const float* tris, uint32_t numTris,
                                uint32_t numBitsPerDim, clr_lbvh_inode* internalNodes,
                                clr_lbvh_leaf* leafNodes, void* stream);
/* Largest numTris clr_lbvh_build_soup_noalloc accepts. */
uint32_t clr_lbvh_noalloc_max_tris(void);

/* Pre-grows the scratch pool and the hierarchy's exchange buffer for builds
 * of up to maxFaces triangles (mirrors the capacity argument of the
 * LBVHBuilder ctor). Synchronous. After it, builds of up to maxFaces
 * triangles allocate nothing and may be captured into a hipGraph. */
int clr_lbvh_reserve(clr_ctx* ctx, uint32_t maxFaces, uint32_t numBitsPerDim);

/* Segmented build: numSegments independent scenes in one set of launches.
 * segOffsets[numSegments+1] (DEVICE) are prefix offsets into the soup (in
 * triangles): scene s owns tris[segOffsets[s] .. segOffsets[s+1]), every
 * scene holds >= 1 triangle, segOffsets[0] = 0 and segOffsets[numSegments] =
 * numTris (the total, passed so that the host never reads the offsets back).
 * Outputs are concatenated: scene s writes leafNodes[segOffsets[s] ...] and
 * internalNodes[segOffsets[s] - s ...] (numTris_s - 1 nodes), with child
 * indices local to the scene, exactly as clr_lbvh_build_soup would write them
 * for that scene alone. The offsets are checked on the device: offsets that
 * break these rules raise bit 4 of clr_device_status and nothing is written.
 * Stream-ordered with no host synchronisation once the scratch pool holds
 * the build (a first call, or clr_lbvh_reserve_segmented, grows it), so the
 * call may be captured into a hipGraph and replayed. numBitsPerDim 1..10,
 * numTris < 2^30. No counterpart in the reference, which builds one scene per
 * LBVHBuilder::perform (LBVHBuilder.cpp:160-375). */
int clr_lbvh_build_segmented(clr_ctx* ctx, const float* tris, const uint32_t* segOffsets,
                             uint32_t numSegments, uint32_t numTris, uint32_t numBitsPerDim,
                             clr_lbvh_inode* internalNodes, clr_lbvh_leaf* leafNodes,
                             void* stream);

/* Pre-grows the scratch pool for segmented builds of up to maxTris triangles
 * in numSegments scenes. Synchronous. */
int clr_lbvh_reserve_segmented(clr_ctx* ctx, uint32_t maxTris, uint32_t numSegments, uint32_t numBitsPerDim);

/* Stage profiling of LBVH builds. clr_set_profiling(ctx, 1) starts recording
 * HIP events at the stage boundaries of every following build (up to 64
 * builds, on the build's own stream); clr_lbvh_last_profile returns the mean
 * per-stage GPU time in ms over the recorded builds, stages =
 * [bounds, morton+histograms, radix sort, block-local hierarchy (k_hier_local),
 * cross-block hierarchy (k_hier_global)]. It waits for the
 * recorded builds to finish. (Replaces the profiling prints of perform,
 * LBVHBuilder.cpp:173-349.) */
int clr_set_profiling(clr_ctx* ctx, int enable);
int clr_lbvh_last_profile(clr_ctx* ctx, float* msPerStage, int maxStages, int* numStages);

/* ---- closest-hit traversal ----------------------------------------------- */

/* Closest hit of numRays rays against an LBVH (semantics of rayIntersection,
 * bvh_traversal.cl:182-215: t starts at tmax, a triangle replaces the current
 * hit when its t <= current t, Moller-Trumbore with |det| < 1e-6 culling).
 *   rays   device float[7][numRays] SoA planes: ox, oy, oz, dx, dy, dz, tmax
 *   hitFace device uint32_t[numRays]: face index or 0xFFFFFFFF on a miss
 *   hitT    device float[numRays]: hit distance (tmax on a miss)
 *   stats   optional device uint32_t[2*numRays]: internal-node visits and
 *           triangle tests per ray (NULL to skip)
 * Face indices are 32-bit (the reference truncates to 16 bits,
 * bvh_traversal.cl:43,55). */
int clr_trace_closest(clr_ctx* ctx, const clr_lbvh_inode* internalNodes,
                      const clr_lbvh_leaf* leafNodes, uint32_t numFaces,
                      const clr_float3* vertices, const clr_face* faces,
                      const float* rays, uint32_t numRays, uint32_t* hitFace, float* hitT,
                      uint32_t* stats, void* stream);

/* Same arguments and outputs as clr_trace_closest, with the ordered
 * traversal: at each node both children's boxes are slab-tested (reciprocal
 * direction, entry padded by 1 - 2^-16, exit by 1 + 2^-16, so no box holding
 * an accepted intersection is rejected), the nearer child is entered first,
 * and a triangle replaces the hit when its t is smaller, or equal with a
 * smaller face index. The hit is therefore the lexicographic minimum of
 * (t, face) over all faces the reference's triangle test accepts,
 * independent of visiting order; it differs from clr_trace_closest only when
 * two faces tie in t exactly (the reference keeps the last one its walk
 * tests) or a ray grazes a box within rounding. stats count entered internal
 * nodes and triangle tests of this traversal. No counterpart in the
 * reference; the renderers use its 4-wide form (clr_trace_closest_wide). */
int clr_trace_closest_ordered(clr_ctx* ctx, const clr_lbvh_inode* internalNodes,
                              const clr_lbvh_leaf* leafNodes, uint32_t numFaces,
                              const clr_float3* vertices, const clr_face* faces,
                              const float* rays, uint32_t numRays, uint32_t* hitFace, float* hitT,
                              uint32_t* stats, void* stream);

/* Same arguments, outputs and hits as clr_trace_closest_ordered, with the
 * 4-wide form of the ordered traversal the renderers use: the LBVH's
 * internal nodes are regrouped into 128-byte records of up to four
 * grandchildren (the children of each internal child, a leaf child itself),
 * a visit slab-tests all four, tests the leaf entries nearest first, enters
 * the nearest internal one and pushes the others with their entry distance
 * (a pop skips entries beyond the current t). Every grandchild lies inside
 * its parent's box, so the accepted (t, face) is the ordered traversal's;
 * stats count visited records (about half the binary walk's node visits)
 * and triangle tests. No counterpart in the reference. */
int clr_trace_closest_wide(clr_ctx* ctx, const clr_lbvh_inode* internalNodes, const clr_lbvh_leaf* leafNodes,
                           uint32_t numFaces, const clr_float3* vertices, const clr_face* faces, const float* rays,
                           uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats, void* stream);

/* Closest hit over the CPU midpoint BVH (nodes: device clr_bvh_node[2 *
 * numFaces - 1], e.g. from clr_bvh_build_midpoint). ordered == 0: the
 * reference's walk over those nodes (rayIntersection without __USE_LBVH,
 * bvh_traversal.cl:256-285: pop a node, test its box against the current t,
 * push children[1] then children[0], test a leaf's triangle replacing the hit
 * when tt <= t); ordered != 0: the ordered traversal of
 * clr_trace_closest_ordered on these nodes, whose hits equal that call's on
 * any BVH of the same faces. Other arguments as clr_trace_closest. */
int clr_trace_closest_bvh(clr_ctx* ctx, const clr_bvh_node* nodes, uint32_t numFaces,
                          const clr_float3* vertices, const clr_face* faces, const float* rays,
                          uint32_t numRays, uint32_t* hitFace, float* hitT, uint32_t* stats,
                          int ordered, void* stream);

/* ---- path tracing ---------------------------------------------------------- */

/* Device-resident scene: the arguments of the reference kernel `pathtracing`
 * (pathtracer.cl:25-33) plus sizes. Empty normal / tangent / uv arrays may be
 * 1-element dummies as in the reference (main.cpp:450-464). */
typedef struct clr_scene_dev {
    const clr_float3* vertices;
    const clr_float3* normals;
    const clr_float3* tangents;
    const float* uvs;                 /* float2 per uv */
    const clr_face* faces;
    const clr_light_info* lights;
    const uint8_t* materialsData;
    const uint8_t* texturesData;
    const uint8_t* otherResources;    /* [0..2]: camera, environment, light-power offsets */
    const clr_lbvh_inode* internalNodes;
    const clr_lbvh_leaf* leafNodes;
    uint32_t numFaces;
    uint32_t numLights;
    /* The reference's non-LBVH build (pathtracer.cl:25-33 with BVHNodes in
     * place of the LBVH pair): when not NULL, rendering traverses this CPU
     * midpoint BVH (2 * numFaces - 1 nodes) and the LBVH pointers are not
     * read. */
    const clr_bvh_node* bvhNodes;
} clr_scene_dev;

typedef struct clr_render_params {
    uint32_t width, height;   /* image size; must equal the camera record's */
    uint32_t samplesPerPixel; /* camera paths per pixel in this call */
    uint32_t maxDepth;        /* bounce cap: a path stops after maxDepth bounces
                                 (the reference's cap is ray.depth > 99, i.e. 100,
                                 pathtracer.cl:166) */
    uint32_t flags;           /* CLR_RENDER_* flags below */
} clr_render_params;

enum { CLR_RENDER_MEGAKERNEL = 0, CLR_RENDER_WAVEFRONT = 1 };

/* clr_render_params.flags. By default rays use the ordered traversal of
 * clr_trace_closest_ordered (nearer child first, ties to the smaller face
 * index); this flag selects the reference's fixed-order walk and box test
 * instead (bvh_traversal.cl:160-215). The two differ only where two
 * triangles are hit at exactly the same distance or a ray grazes a box
 * within rounding. */
#define CLR_RENDER_REFERENCE_TRAVERSAL 1u
/* Turns Russian roulette off (paths stop only at maxDepth or when the BSDF
 * sample fails): the estimator is then a smooth function of the material and
 * emission parameters for fixed random states, as finite-difference checks
 * with common random numbers need. Not in the reference (pathtracer.cl:167-172
 * always applies it). */
#define CLR_RENDER_NO_RR 2u
/* Sampling strategy. By default every diffuse or glossy vertex combines a
 * light sample and the BSDF sample's emission hit with the power heuristic
 * (pathtracer.cl:106-128,154-163). These flags keep one strategy alone, for
 * unbiasedness checks (all three converge to the same image):
 *   LIGHT_SAMPLING_ONLY: light samples weigh 1; emission (area or
 *     environment) reached by a diffuse or glossy BSDF sample weighs 0.
 *   BSDF_SAMPLING_ONLY: no light sample is added (its three random numbers
 *     are still drawn, so the paths are the same); emission reached by a BSDF
 *     sample weighs 1.
 * Emission seen by the camera or through a specular bounce weighs 1 in every
 * mode. Setting both flags is an invalid argument. Not in the reference. */
#define CLR_RENDER_LIGHT_SAMPLING_ONLY 4u
#define CLR_RENDER_BSDF_SAMPLING_ONLY 8u

/* Adds samplesPerPixel camera paths per pixel into `pixels` (clr_float3 per
 * pixel, row-major, not divided by the sample count: the caller scales, as
 * post_processing.cl:20-26 does) and advances the per-pixel xorshift128
 * states `randStates` (4 x uint32 per pixel). Replaces the host loop of
 * main.cpp:475-494 launching `pathtracing` per tile and iteration. Both
 * modes consume each pixel's random stream in the reference's order and give
 * the same image up to floating-point evaluation order. */
int clr_render(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params,
               uint32_t* randStates, clr_float3* pixels, int mode, void* stream);

/* Renders numScenes scenes with the same image size and params in one
 * wavefront pass (CLR_RENDER_WAVEFRONT semantics, default ordered traversal):
 * scenes[] is a HOST array of device scenes; randStates and pixels hold the
 * scenes' images back to back (scene s: states[s*W*H*4 ...], pixels[s*W*H
 * ...]). Every scene's image equals what clr_render(..., WAVEFRONT) gives
 * for that scene alone. The reference renders one scene per process
 * (main.cpp:475-494); this is the multi-scene form of that loop, paired with
 * clr_lbvh_build_segmented. numScenes * W * H < 2^31;
 * CLR_RENDER_REFERENCE_TRAVERSAL is not supported here.
 * The call may be captured into a hipGraph on `stream` (together with the
 * segmented build, the state seeding and the image clear of a frame) once
 * an eager call with the same scene count, sizes and params has sized the
 * context's buffers; a capture that would have to grow them fails with
 * CLR_ERR_INVALID_ARG. Under capture the scene table is written by kernels
 * whose arguments carry it, and all samplesPerPixel * (maxDepth + 1) + 1
 * iterations are recorded as plain launches (the eager call stops early by
 * reading a path count back). The graph stays valid while the context's
 * batch and wavefront buffers are not regrown. */
int clr_render_batch(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                     const clr_render_params* params, uint32_t* randStates, clr_float3* pixels,
                     void* stream);

/* clr_render_batch with a choice of renderer: mode CLR_RENDER_WAVEFRONT is
 * clr_render_batch; CLR_RENDER_MEGAKERNEL renders the batch with persistent
 * megakernel waves (ordered traversal) that take 8x8 tiles of all the scenes
 * from one counter. Both give every scene the image clr_render(...,
 * MEGAKERNEL or WAVEFRONT) gives it alone, bit for bit. Same arguments,
 * limits and capture rule as clr_render_batch (an eager call with the same
 * scene count, sizes and params sizes the buffers first), except that the
 * megakernel form takes any maxDepth. Unknown modes return
 * CLR_ERR_INVALID_ARG. Like clr_render_batch, the multi-scene form of
 * main.cpp:475-494. */
int clr_render_batch_mode(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                          const clr_render_params* params, uint32_t* randStates, clr_float3* pixels, int mode,
                          void* stream);

/* bvh_visualizer (bvh_visualizer.cl:25-84), a traversal-cost view: for each
 * pixel one ray through its centre from the lens centre; pixels += the base
 * colour of the surface hit (getBaseColor_fs, reflection.cl:730-772), then
 * += 0.025 * the ray's box hits. With params->flags &
 * CLR_RENDER_REFERENCE_TRAVERSAL the count is the reference's
 * (rayIntersectionVis, bvh_traversal.cl:216-319: internal and leaf boxes
 * passed on the LBVH, node boxes passed on scene->bvhNodes); otherwise it is
 * the ordered traversal's entered nodes plus triangle tests. Uses
 * params->width / height / flags; samplesPerPixel and maxDepth are ignored. */
int clr_bvh_visualizer(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params,
                       clr_float3* pixels, void* stream);

/* ---- post-processing (post_processing.cl, BMPExporter.cpp) ----------------
 * Device buffers of clr_float3 per pixel (row-major), stream-ordered. src and
 * output must not overlap. The reference's gaussianScattering and bloom add
 * with float atomics in arrival order; these add the same terms per output
 * pixel in a fixed order (deterministic; equal to the reference up to the
 * order of float additions). */

/* clear (post_processing.cl:16-18): buffer = 0. */
int clr_post_clear(clr_ctx* ctx, uint32_t width, uint32_t height, clr_float3* buffer, void* stream);

/* scaling (post_processing.cl:20-26): output = src * (1 / spp). */
int clr_post_scaling(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t spp, const clr_float3* src,
                     clr_float3* output, void* stream);

/* gaussianScattering (post_processing.cl:47-66): output += 9x9 Gaussian of
 * src * (1 / spp), taps beyond the image clamped onto its border. */
int clr_post_gaussian_scattering(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t spp,
                                 const clr_float3* src, clr_float3* output, void* stream);

/* bloom (post_processing.cl:70-101): per channel, v = src * (1 / spp); a
 * value above 30 adds 30 to its pixel and exp(-|d|) * (v - 30) to the pixels
 * at offsets d of the square of radius min(7, (int)(v - 30)) (clamped onto
 * the border); other values add v to their pixel. output += the result. */
int clr_post_bloom(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t spp, const clr_float3* src,
                   clr_float3* output, void* stream);

/* toneMapping (post_processing.cl:103-117): channels saturated at 30, luminance
 * Y = 0.2989 R + 0.5866 G + 0.1144 B, scale (1 - exp(-Y)) / Y (1 for Y = 0),
 * byte = trunc(clamp(255 * (scale * c)^(1/2.2), 0, 255)), written B, G, R into
 * bottom-up rows of rowBytes (>= 3 * width) bytes of the device buffer dst;
 * the row padding is written as 0. Luminance and scale are computed in float. */
int clr_post_tone_mapping(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t rowBytes, const clr_float3* src,
                          uint8_t* dst, void* stream);

/* Bytes per BMP row as the reference computes them: 3 * width + width % 4
 * (main.cpp:521, BMPExporter.cpp:15). */
uint32_t clr_bmp_row_bytes(uint32_t width);

/* saveBMP (BMPExporter.cpp:14-53): writes a 24-bit uncompressed BMP (54-byte
 * header, 1 pixel per metre) of HOST pixels in the tone-mapped layout. */
int clr_save_bmp(const char* filename, const void* pixels, uint32_t width, uint32_t height);

/* ---- differentiation ------------------------------------------------------
 * No counterpart in the reference. Differentiable parameters: the albedo
 * (RGB) and roughness sigma of materials made of one Oren-Nayar / Lambert
 * element (reflection.cl:239-255; albedo and sigma are constant textures the
 * caller may rewrite in texturesData between calls), and the emission M (RGB)
 * of one diffuse emitter. */
typedef struct clr_diff_params {
    uint32_t numMaterials;       /* at most 256 */
    const uint32_t* matOffsets;  /* HOST array: their byte offsets in materialsData (faces' matPtr) */
    uint32_t lightProp;          /* the emitter's light-property offset (faces' lightPtr), or
                                    0xFFFFFFFF for none */
} clr_diff_params;

/* Path-replay adjoint of clr_render (megakernel order, same flags). With
 * `randStates` the per-pixel states a forward clr_render call started from
 * (not modified here) and the same params, writes the gradients of
 *     sum_p  dL_dimage[p] . (pixels added by that call)[p] / samplesPerPixel
 * with respect to every listed material's albedo (gAlbedo[m][3]) and
 * roughness (gRoughness[m]) and the emitter's emission (gEmission[3]).
 * Sampling decisions are replayed, not differentiated (Russian roulette
 * weights included; CLR_RENDER_NO_RR makes the estimate smooth in the
 * parameters). Each sample is traced once: the radiance accumulated before
 * each differentiable vertex is recorded and later terms are recovered as
 * L - prefix. maxDepth <= 15. Gradient buffers are device memory, overwritten. */
int clr_render_adjoint(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params,
                       const clr_diff_params* diff, const uint32_t* randStates, const clr_float3* dL_dimage,
                       float* gAlbedo, float* gRoughness, float* gEmission, void* stream);

/* Seeds per-pixel states: seed = pixelIndex ^ 0x9E3779B9 expanded as
 * s_k = seed = 1812433253 * (seed ^ (seed >> 30)) + k, k = 0..3
 * (the per-pixel expansion of main.cpp:101-104, unsigned, no discards). */
int clr_seed_pixel_states(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t* randStates,
                          void* stream);

/* ---- wavefront renderer tuning -------------------------------------------
 * No counterpart in the reference, which launches one megakernel per tile
 * (main.cpp:475-494). CLR_RENDER_WAVEFRONT runs the integrator as queue-driven
 * kernels (extend, shade, miss, shadow) over one path slot per pixel in
 * flight, replayed from a captured hipGraph of itersPerGraph iterations. */
typedef struct clr_wavefront_config {
    uint32_t traceBlock;      /* threads per block of the extend / shadow kernels: 64, 128, 256 */
    uint32_t traceWavesPerCU; /* their grid: waves per CU, 1..128 (above 32, more waves than a CU holds) */
    uint32_t shadeBlock;      /* threads per block of the shade / miss kernels: 64, 128, 256 */
    uint32_t shadeWavesPerCU; /* their grid: waves per CU, 1..128 */
    uint32_t itersPerGraph;   /* iterations per captured graph: even, 2..64 */
    uint32_t checkEvery;      /* graph launches between early-stop checks (a host sync);
                                 0: never, the render stays fully asynchronous */
    uint32_t maxSlots;        /* path slots (pixels in flight); larger images run in chunks */
    uint32_t useGraph;        /* 1: replay the captured graph, 0: launch kernels directly */
    uint32_t refillMin;       /* traversal kernels hand new rays to a wave's idle lanes once at
                                 least this many are idle (1..64; 64: only when all are) */
    uint32_t ldsDepth;        /* traversal stack words per walker kept in LDS (16 or 32; the
                                 rest of the 128 spill to a per-lane global region) */
    uint32_t raysPerLane;     /* rays each lane of the ordered extend / shadow kernels walks at
                                 once (1 or 2): 2 overlaps one ray's node fetch with the other's
                                 tests (incoherent rays), 1 keeps more waves resident */
} clr_wavefront_config;

int clr_wavefront_get_config(clr_ctx* ctx, clr_wavefront_config* config);
int clr_wavefront_set_config(clr_ctx* ctx, const clr_wavefront_config* config);
/* Times candidate configurations on this scene and image size (renders of
 * `params->samplesPerPixel` samples into scratch copies; the caller's image
 * and states are not touched), keeps the fastest in the context and returns
 * it in `config` (may be NULL). Synchronises `stream`. */
int clr_wavefront_autotune(clr_ctx* ctx, const clr_scene_dev* scene, const clr_render_params* params,
                           clr_wavefront_config* config, void* stream);

/* clr_wavefront_autotune for clr_render_batch: the same search timed on
 * batched renders of scenes[0..numScenes) (HOST array). */
int clr_wavefront_autotune_batch(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                                 const clr_render_params* params, clr_wavefront_config* config, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CLEAR_AMD_H */
