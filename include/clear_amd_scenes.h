/*
 * clear_amd_scenes.h — host-side generators for the benchmark scenes.
 *
 * These have no single counterpart in the reference, whose scene is built by
 * hand in main.cpp's buildScene (OpenCL_PathTracer/main.cpp:46-238) from OBJ
 * assets; the benchmark scenes here are procedural so that the CPU oracle and
 * the GPU consume the same bytes. All writes go to HOST memory. Randomness
 * comes from the reference's xorshift128 host generator (XORShift.cpp:12-29:
 * seed expansion 1812433253 * (s ^ (s >> 30)) + i over a signed int, then 50
 * discarded draws).
 */
#ifndef CLEAR_AMD_SCENES_H
#define CLEAR_AMD_SCENES_H

#include <stdint.h>

#include "clear_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Gaussian-cluster triangle soup: numClusters centres uniform in [0,1)^3;
 * each triangle picks a cluster uniformly, its centroid is the centre plus
 * N(0, sigma^2) per axis (Box-Muller), and two edge vectors have directions
 * uniform on the sphere and lengths uniform in [eMin, eMax):
 * v0 = c - (e1 + e2) / 3, v1 = v0 + e1, v2 = v0 + e2.
 * out: float[numTris][9]. */
int clr_gen_cluster_soup(uint32_t numTris, int seed, uint32_t numClusters, float sigma,
                         float eMin, float eMax, float* out);

/* Random rays: origins uniform in [0,1)^3, directions uniform on the sphere
 * (z = 1 - 2u, phi = 2 pi v), tmax = +inf. out: float[7][numRays] SoA planes
 * (ox, oy, oz, dx, dy, dz, tmax). */
int clr_gen_random_rays(uint32_t numRays, int seed, float* out);

/* The CPU midpoint BVH of Scene::build (BVH.hpp:94-252, built when the
 * reference renders without __USE_LBVH): one leaf per face with the box of
 * its three vertices, recursive split of the node's box at the centre of its
 * widest axis by Hoare partition of leaf centres, nodes in depth-first
 * creation order. HOST memory: vertices (16-B float3), faces, out[2n - 1]. */
int clr_bvh_build_midpoint(const clr_float3* vertices, const clr_face* faces, uint32_t numFaces,
                           clr_bvh_node* out);

/* ---- render scenes ------------------------------------------------------
 * Built on the host with the drop-in Scene / MaterialCreator API
 * (include/clear_amd_scene.hpp) and exposed as the reference's arrays and
 * byte blobs; upload them to the device and fill a clr_scene_dev. */
typedef struct clr_host_scene clr_host_scene;

enum {
    /* Closed Cornell box [-1,1]^3, 0.3-radius icosphere (subdivision 5:
     * 20,480 triangles) with vertex normals, 2-triangle ceiling light of
     * emission 15, all Lambert (white 0.75, red/green 0.6), pinhole camera. */
    CLR_SCENE_CORNELL = 0,
    /* Coverage scene: every BxDF (Oren-Nayar, specular reflection and
     * transmission, Ward, Ashikhmin-Shirley, Beckmann coat), checkerboard,
     * bump, RGBA8, gray8 alpha and FP32 textures, normal maps, alpha
     * cut-outs, an image-based environment and a thin lens. */
    CLR_SCENE_SHOWCASE = 1,
    /* White-furnace check: an icosphere of Lambert albedo `albedo` inside a
     * uniform environment of radiance 1; the expected image is the camera
     * weight times 1 (background) or albedo (sphere). */
    CLR_SCENE_FURNACE = 2,
    /* Outdoor renderer workload: a detail^2-vertex fBm heightfield (default
     * 1024^2: 2,093,058 triangles) with a 512^2 RGB8 normal map plus 200
     * subdivision-4 icospheres (1,024,000 triangles), all Lambert under a
     * Beckmann coat (alpha U[0.05,0.5], dielectric eta 1.5), a 1024x512 FP32
     * sky with a sun of peak 5e4, thin lens (default radius 0.02; a negative
     * lensRadius gives a pinhole), 1920x1080 by default. */
    CLR_SCENE_TERRAIN = 3,
    /* Differentiation scene: the closed Cornell room with a 200,000-triangle
     * UV sphere whose face f uses Oren-Nayar material f mod 64 (albedo
     * U[0.2,0.8]^3, roughness U[0.1,0.6] from `seed`, 4 when 0) and the
     * emission-15 ceiling light; 512x512. clr_scene_diff_params lists the
     * differentiable parameters. */
    CLR_SCENE_ADJOINT = 4,
    /* Batch-render scene: numTriangles triangles of the Gaussian-cluster soup
     * of clr_gen_cluster_soup (64 clusters, sigma 0.02, edges 1e-3..1e-2,
     * `seed`) in the unit cube, triangle i using material i mod 8 (four
     * Lambert, four Lambert under a Beckmann coat of alpha U[0.05,0.5], eta
     * 1.5; albedo U[0.3,0.9]^3 from the host RNG), plus a 2-triangle area
     * light (emission 10) facing down above the cube (faces numTriangles and
     * numTriangles+1) and a white Lambert floor quad under it (faces
     * numTriangles+2, +3); pinhole camera, 128x128 by default. */
    CLR_SCENE_CLUSTERS = 5,
    /* Direct-lighting check: a Lambert floor of albedo `albedo` (0 = 0.5)
     * over [-2,2]^2 at y = 0 facing up (faces 0, 1), no environment, and
     * diffuse emitters of emission (4, 2, 1) that are also white Lambert
     * surfaces: detail 0 (default) a quad [-0.1,0.7] x [-0.45,0.05] at
     * y = 0.6 facing down (faces 2, 3); detail 1 that quad plus the triangle
     * (-0.9,0.6,0.1), (-0.3,0.6,0.5), (-0.8,0.6,0.7) in its plane, facing
     * down (face 4); detail 2 an equilateral triangle of circumradius 0.5
     * centred at (0, 0.6, 0) facing along (1, -1, 0) / sqrt 2 (face 2), so
     * the floor at x < -0.6 lies behind it. Pinhole camera at (0, 2.5, 2.5)
     * looking at the origin, 45 degrees, 32x32 by default. At maxDepth 1 the
     * image is the floor's direct lighting, which Lambert's polygon formula
     * gives in closed form. */
    CLR_SCENE_AREA_LIGHTS = 6
};

enum {
    CLR_ARRAY_VERTICES = 0, CLR_ARRAY_NORMALS, CLR_ARRAY_TANGENTS, CLR_ARRAY_UVS, CLR_ARRAY_FACES,
    CLR_ARRAY_LIGHTS, CLR_ARRAY_MATERIALS, CLR_ARRAY_TEXTURES, CLR_ARRAY_OTHERS, CLR_ARRAY_BVH_NODES
};

typedef struct clr_scene_params {
    uint32_t width, height;   /* image size (0 = scene default) */
    uint32_t detail;          /* sphere subdivision level (0 = scene default) */
    int seed;                 /* procedural content seed */
    float lensRadius;         /* thin-lens radius (0 = scene default: pinhole except TERRAIN) */
    float focusDistance;      /* thin-lens focus distance (0 = scene default) */
    uint32_t buildCpuBvh;     /* also build the reference's CPU midpoint BVH */
    float albedo;             /* furnace sphere albedo */
    float fovY;               /* vertical field of view in degrees (0 = scene default) */
    uint32_t numTriangles;    /* CLR_SCENE_CLUSTERS soup size (0 = 10,000) */
} clr_scene_params;

int clr_scene_build(int kind, const clr_scene_params* params, clr_host_scene** out);
int clr_scene_free(clr_host_scene* scene);
int clr_scene_info(const clr_host_scene* scene, uint32_t* numFaces, uint32_t* numLights,
                   uint32_t* width, uint32_t* height);
/* Differentiable parameters of a CLR_SCENE_ADJOINT scene: the number of
 * materials, and per material its byte offset in materialsData (the faces'
 * matPtr) and the texturesData byte offsets of its albedo (3 floats) and
 * roughness sigma (1 float); the emitter's light-property offset (the faces'
 * lightPtr) and the texturesData offset of its emission (3 floats). Arrays
 * may be NULL; when given, *numMaterials is their capacity on entry. */
int clr_scene_diff_params(const clr_host_scene* scene, uint32_t* numMaterials, uint32_t* matOffsets,
                          uint32_t* albedoOffsets, uint32_t* sigmaOffsets, uint32_t* lightProp,
                          uint32_t* emissionOffset);

/* Pointer and byte size of one host array (CLR_ARRAY_*); valid until free. */
int clr_scene_array(const clr_host_scene* scene, int which, const void** data, uint64_t* bytes);

/* ---- asset ingest --------------------------------------------------------
 * Drop-ins for loadImage (ImageLoader.h:24, ImageLoader.cpp:208-232) and
 * loadModel (ModelLoader.hpp:15, ModelLoader.cpp:156-508); the C++ API is
 * include/clear_amd_loaders.hpp. */
enum { CLR_COLOR_RGB8 = 0, CLR_COLOR_RGBA8 = 1, CLR_COLOR_RGBA16F = 2, CLR_COLOR_GRAY8 = 3 };

/* Decodes a PNG or OpenEXR file: RGB8 as 4 bytes per pixel (the 4th 0xFF),
 * RGBA8, RGBA16F (8 bytes per pixel) or gray8, rows top to bottom; PNG samples
 * go through the reference's libpng gamma setting (gammaCorrection 0: file
 * gamma undone; 1: samples as stored). With pixels == NULL only *bytes (the
 * size) is returned; otherwise capacity must be at least that. The message
 * of a failure is in clr_scene_last_error(). */
int clr_image_load(const char* path, int gammaCorrection, uint32_t* width, uint32_t* height,
                   uint32_t* colorType, void* pixels, uint64_t capacity, uint64_t* bytes);

/* A scene from files: each OBJ model (with its MTL materials and their image
 * textures) loaded by loadModel; model k emits emission[3k..3k+2] when that is
 * nonzero (a diffuse light property given as the reference's lightOverride).
 * envPath (OpenEXR, or NULL) becomes an image-based environment scaled by
 * envMultiplier. The camera looks at the centre of the models' bounding box
 * along -z from where the box's bounding sphere fills the field of view
 * (params: image size, default 256x256; fovY, default 45; lens). */
int clr_scene_build_files(const char* const* objPaths, const float* emission, uint32_t numModels,
                          const char* envPath, float envMultiplier, const clr_scene_params* params,
                          clr_host_scene** out);

/* Message of the last failed clr_image_load / clr_scene_build_files of this
 * thread ("" when none). */
const char* clr_scene_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CLEAR_AMD_SCENES_H */
