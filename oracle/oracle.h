/*
 * oracle.h - CPU restatement of the reference's ray-query algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (igx_raytracing_amd/,
 * include/) may link, load or call this library.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py use it, and
 * only as the checker or the timed CPU baseline.
 *
 * What it restates (citations are reference files):
 *   orc_mt_f32                res/shaders/primitive.glsl:239-284 (FP32, no
 *                             contraction, IEEE divide, in source order)
 *   orc_closest_brute         res/shaders/trace.glsl:9-66 + raygen.comp:39-44
 *   orc_occluded_brute        res/shaders/trace.glsl:70-98
 *   orc_sphere, orc_cube,     res/shaders/primitive.glsl:173-210, :286-333,
 *   orc_plane                 :212-237 (analytic.c), with the Cephes
 *                             single-precision asin/atan (orc_asinf,
 *                             orc_atanf) for the sphere's uv
 *   orc_*_brute_ana           the same scans over the mixed object-index
 *                             space: triangles, spheres, cubes, planes
 *                             (trace.glsl:22-46, :77-95)
 *   orc_mt_f64                the same formula in FP64 (tie adjudicator)
 *   orc_*_brute_guard         the scans without FP32 artifacts (out-of-box
 *                             or NaN accepts), the bound a BVH must meet
 *   orc_lbvh_binary           the GPU LBVH builder's binary tree (lbvh.c)
 *   orc_raygen                raygen.comp: camera rays, scan, packing (callers.c)
 *   orc_pathtrace             the wavefront path tracer, path by path (pathtrace.c)
 *   orc_walk                  the traversal of the GPU kernels over the same
 *                             node bytes, in the same visiting order, with
 *                             the (t, lowest id) rule; counts node fetches
 *                             and triangle tests (roofline bytes)
 *
 * Pinning (see tests/test_oracle.py): hand-derived known-answer triangles,
 * geometric reconstruction o + t d == (1-u-v) p0 + u p1 + v p2 checked in
 * FP64, agreement with the FP64 twin, brute force == BVH walk, occluded ==
 * (closest.t < tmax), ray-order independence.
 *
 * Rays are 8 floats: org.xyz, tmax, dir.xyz, prev_id bits.  Hits are 4 words:
 * id, t bits, u bits, v bits.
 */
#ifndef IGX_RT_ORACLE_H
#define IGX_RT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Returns 1 if the u/v tests pass (then *t, *u, *v are set), else 0. */
int orc_mt_f32(const float o[3], const float d[3], const float p0[3],
               const float p1[3], const float p2[3], float *t, float *u,
               float *v);
int orc_mt_f64(const double o[3], const double d[3], const double p0[3],
               const double p1[3], const double p2[3], double *t, double *u,
               double *v);

/* Brute force over input-order triangles tri[n][9].  nan_flag (optional,
 * one byte per ray) is set where a NaN t was accepted (the reference's
 * strict `t >= hitT` reject lets NaN through and it then poisons hitT). */
void orc_closest_brute(const float *tri, uint32_t ntri, const float *rays,
                       uint32_t nrays, uint32_t *hits, uint8_t *nan_flag,
                       int nthreads);
/* The same scans without the pairs FP32 accepts as artifacts of rounding: a
 * NaN t, or a hit point o + t d (FP64 from the FP32 t) more than `guard`
 * outside the triangle's box.  Near-parallel rays make them: with a
 * denominator a near its rounding error, MT can accept a pair whose residual
 * |o + t d - (p0 + u e1 + v e2)| is of the order of |o - p0|.  A BVH reaches
 * every pair whose hit point lies within its box padding, so its result
 * lies between these scans and the unguarded ones. */
void orc_closest_brute_guard(const float *tri, uint32_t ntri, const float *rays, uint32_t nrays, uint32_t *hits,
                             double guard, int nthreads);
void orc_occluded_brute_guard(const float *tri, uint32_t ntri, const float *rays, uint32_t nrays, uint8_t *flags,
                              double guard, int nthreads);
void orc_occluded_brute(const float *tri, uint32_t ntri, const float *rays,
                        uint32_t nrays, uint8_t *flags, int nthreads);

/* Analytic primitives (trace.glsl:30-46): spheres[ns][4] = centre, radius
 * (the radius is squared at primitive.glsl:180); cubes[nc][6] = start.xyz,
 * end.xyz (Cube{xy0, z0_x1, yz1}, primitive.glsl:40-44); planes[np][4] =
 * normal.xyz, offset.  Object ids continue after the triangles. */
typedef struct orc_analytic {
    const float *spheres;
    uint32_t num_spheres;
    const float *cubes;
    uint32_t num_cubes;
    const float *planes;
    uint32_t num_planes;
} orc_analytic;

float orc_asinf(float x);
float orc_atanf(float x);
/* One primitive test: returns 1 when accepted, and then sets *best_t, uv
 * and n (the geometry normal; n may be NULL). */
int orc_sphere(const float o[3], const float d[3], const float s[4], uint32_t obj, uint32_t prev, float *best_t,
               float uv[2], float n[3]);
int orc_cube(const float o[3], const float d[3], const float c[6], uint32_t obj, uint32_t prev, float *best_t,
             float uv[2], float n[3]);
int orc_plane(const float o[3], const float d[3], const float pl[4], uint32_t obj, uint32_t prev, float *best_t,
              float uv[2], float n[3]);
/* The analytic loops, continuing from the triangle loop's state. */
void orc_analytic_scan(const orc_analytic *a, uint32_t base, const float o[3], const float d[3], uint32_t prev,
                       float *hit_t, uint32_t *obj, float *hu, float *hv, float n[3], int *nan_seen);
void orc_closest_brute_ana(const float *tri, uint32_t ntri, const orc_analytic *ana, const float *rays,
                           uint32_t nrays, uint32_t *hits, uint8_t *nan_flag, int nthreads);
void orc_occluded_brute_ana(const float *tri, uint32_t ntri, const orc_analytic *ana, const float *rays,
                            uint32_t nrays, uint8_t *flags, int nthreads);

/* Traversal of a built scene: node bytes of width W (2, 4, 8), triangle
 * records (48 B: p0 id | e1 0 | e2 0), root reference.  occlusion = 0 writes
 * hits (4 words per ray); 1 writes flags.  counts (optional) gets 2 words per
 * ray: node fetches, triangle tests. */
void orc_walk(int width, const void *nodes, const void *tris, uint32_t root,
              const float *rays, uint32_t nrays, int occlusion,
              uint32_t *hits, uint8_t *flags, uint32_t *counts, int nthreads);

/* Binary LBVH (lbvh.c): outputs as rt_lbvh_binary in include/rt/rt_c.h. */
int orc_lbvh_binary(const float *tri, uint32_t n, uint32_t *morton, uint32_t *order, uint32_t *left,
                    uint32_t *right, float *boxes);

/* raygen.comp restated (callers.c): cam = eye, p0, p1, p2 (3 floats each),
 * jitter (2).  dir_object and uv_normal are float4 images [h][w]; uv_normal
 * is written on hits only.  normals may be NULL. */
void orc_raygen(const float *tri, uint32_t ntri, const uint32_t *normals, const float cam[14], uint32_t width,
                uint32_t height, float *dir_object, float *uv_normal, int nthreads);
/* The same over the mixed object-index space; an analytic hit packs its
 * geometry normal (trace.glsl:58-59). */
void orc_raygen_ana(const float *tri, uint32_t ntri, const uint32_t *normals, const orc_analytic *ana,
                    const float cam[14], uint32_t width, uint32_t height, float *dir_object, float *uv_normal,
                    int nthreads);

/* calculatePrimary's inputs (res/shaders/camera.glsl:14-44,113-138):
 * projection 0 screen, 1 omni, 2 omni stereo TB, 3 stereo TB, 4 omni stereo
 * LR, 5 stereo LR; rand_jitter 1 takes the pixel offset from
 * rand(loc + seed), else from jitter. */
typedef struct orc_camera {
    float eye[3], p0[3], p1[3], p2[3], p3[3], p4[3], p5[3];
    float ipd;
    uint32_t projection, rand_jitter;
    float seed[2], jitter[2];
} orc_camera;
/* Primary rays [height][width] (8 floats each, tmax FLT_MAX, no prev). */
void orc_primary_rays(const orc_camera *cam, uint32_t width, uint32_t height, float *rays);
void orc_raygen_primary(const float *tri, uint32_t ntri, const uint32_t *normals, const orc_analytic *ana,
                        const orc_camera *cam, uint32_t width, uint32_t height, float *dir_object, float *uv_normal,
                        int nthreads);

/* The reference's Light (primitive.glsl:46-54). */
typedef struct orc_light {
    float pos[3];
    uint32_t rad_origin;
    uint32_t dir[2];
    uint32_t color_type[2];
} orc_light;
/* getDirToLight (light.glsl:98-133): unit direction from the light sample to
 * pos, and the distance to a point light's centre (-1: directional). */
void orc_dir_to_light(const orc_light *L, const float pos[3], float rx, float ry, float l[3], float *dist);
/* nv_all.shadow.comp:50-125: the shadow ray of every (sample, pixel) and
 * whether the reference traces it. */
void orc_shadow_rays(const float *dir_object, const float eye[3], const orc_light *L, const float seed[2],
                     uint32_t width, uint32_t height, uint32_t samples, float *rays, uint8_t *trace);
/* nv_all.shadow.comp:50-143: the tile ballots (a tile without hits is 0). */
void orc_shadow_pass(const float *tri, uint32_t ntri, const orc_analytic *ana, const float *dir_object,
                     const float eye[3], const orc_light *L, const float seed[2], uint32_t width, uint32_t height,
                     uint32_t samples, uint64_t *tiles, int nthreads);

/* The wavefront path tracer restated path by path (pathtrace.c); counts
 * receives the rays traced at each depth 0..max_depth. */
void orc_pathtrace(const float *tri, uint32_t ntri, const float cam[14], uint32_t width, uint32_t height, uint32_t spp,
                   uint32_t max_depth, float rr, float albedo, uint64_t seed, float *image, uint32_t *counts,
                   int nthreads);

/* The image values of the listed pixels only (values[k] for pixels[k], a
 * pixel index y * width + x): the same paths and sums as orc_pathtrace. */
void orc_pathtrace_pixels(const float *tri, uint32_t ntri, const float cam[14], uint32_t width, uint32_t height,
                          uint32_t spp, uint32_t max_depth, float rr, float albedo, uint64_t seed,
                          const uint32_t *pixels, uint32_t npix, float *values, int nthreads);

/* Seconds of wall clock (monotonic). */
double orc_now(void);

#ifdef __cplusplus
}
#endif
#endif
