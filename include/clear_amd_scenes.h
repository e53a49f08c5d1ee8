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

#ifdef __cplusplus
}
#endif

#endif /* CLEAR_AMD_SCENES_H */
