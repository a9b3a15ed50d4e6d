/*
 * rt_workload.h - seeded synthetic scenes and rays for the benchmark
 * workloads (host functions of librt_hip.so; no device needed).
 *
 * The reference ships no scenes or ray sets for its query path (its only
 * scene is the hand-placed NielsScene, test/scene/niels_scene.cpp:10-56), so
 * every workload here is generated from a seed with a PCG32 generator
 * (O'Neill 2014, "pcg32_random_r") and Box-Muller normals computed in double
 * precision, then rounded once to FP32.  Triangles are written as
 * tri_xyz[n][9] = p0 p1 p2; rays as rt_ray (include/rt/rt_c.h).
 */
#ifndef IGX_RT_WORKLOAD_H
#define IGX_RT_WORKLOAD_H

#include <stdint.h>
#include "rt_c.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The seed of workload k is RTW_SEED_BASE + k. */
#define RTW_SEED_BASE 0x5EED0000ull

/* One PCG32 draw for (seed, stream) after `skip` draws, for tests. */
uint32_t rtw_pcg32_nth(uint64_t seed, uint64_t stream, uint64_t skip);

/* Clustered scene: nclusters Gaussian clusters of tris_per_cluster triangles.
 * Cluster centre c ~ U[-range, range]^3, triangle anchor m = c + N(0,sigma)^3,
 * vertices p_j = m + N(0, edge_sigma)^3.  Cluster k uses PCG stream k.
 * Writes nclusters*tris_per_cluster*9 floats; returns the triangle count. */
uint32_t rtw_scene_clusters(uint64_t seed, uint32_t nclusters,
                            uint32_t tris_per_cluster, float range,
                            float sigma, float edge_sigma, float *tri_xyz);

/* Rays with origins uniform in [lo, hi] and directions uniform on the sphere
 * (z = 1 - 2 xi1, phi = 2 pi xi2).  tmax = RT_NO_HIT_T when tmax_hi <= 0,
 * else U[tmax_lo, tmax_hi].  prev_id = RT_MISS.  Ray i uses PCG stream
 * 0x40000000 + i, so any sub-range can be generated independently. */
void rtw_rays_uniform(uint64_t seed, const float lo[3], const float hi[3],
                      uint64_t first, uint32_t n, float tmax_lo,
                      float tmax_hi, rt_ray *rays);

/* Small mixed scene of 4,096 triangles: 64 axis-aligned cubes (768 tris;
 * centres U[-10,10]^3, edge length U[0.5,2]), a 32x32-quad ground grid at
 * y = 0 over [-16,16]^2 (2,048 tris; vertex heights N(0,0.05)) and 1,280
 * loose triangles (centroid U[-10,10]^3, edge vectors N(0,0.3)^3).
 * Returns 4096. */
uint32_t rtw_scene_mixed4k(uint64_t seed, float *tri_xyz);

/* Pinhole primary rays, one per pixel centre, row-major from the top-left
 * pixel: eye, look-at target, world up (0,1,0), vertical fov in degrees.
 * Directions normalised; tmax = RT_NO_HIT_T; prev_id = RT_MISS. */
void rtw_rays_pinhole(uint32_t width, uint32_t height, const float eye[3],
                      const float target[3], float vfov_deg, rt_ray *rays);

/* Axis-aligned bounds of n triangles. */
void rtw_bounds(const float *tri_xyz, uint32_t n, float lo[3], float hi[3]);

/* Displaced icosphere: 20 * 4^subdiv triangles, radius r, each shared vertex
 * moved along its normal by disp * N(0,1).  Returns the triangle count. */
uint32_t rtw_scene_icosphere(uint64_t seed, uint32_t subdiv, float radius,
                             float disp, float *tri_xyz);

#ifdef __cplusplus
}
#endif
#endif /* IGX_RT_WORKLOAD_H */
