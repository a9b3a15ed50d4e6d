/*
 * rt_c.h - C ABI of the MI355X ray-query engine (closest-hit and occlusion
 * queries over triangle scenes).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (RT_OK = 0, negative = error class) and never throws.  The
 * library owns scene memory (triangles, BVH nodes) behind opaque handles; the
 * caller owns ray, hit, flag and mask buffers, which are DEVICE pointers
 * unless a function says otherwise.
 *
 * Query semantics (per ray, all FP32, IEEE, no contraction):
 *   Moller-Trumbore exactly as res/shaders/primitive.glsl:239-284, i.e. no
 *   epsilon (the `abs(a) < 0` test at :258 is dead), inclusive u/v bounds
 *   (:265,271), reject t <= 0 (strict tmin = 0), reject id == prev_id (:276).
 *   closest:  the (t, id)-lexicographic minimum over accepted triangles with
 *             t < min(tmax, FLT_MAX): the lowest triangle index wins a tie in
 *             t, as the in-order scan of trace.glsl:25-27 with its strict
 *             `t >= hit.hitT` reject does.  Miss: id = RT_MISS,
 *             t = RT_NO_HIT_T (FLT_MAX), u = v = 0 (trace.glsl:14-16).
 *   occluded: exists an accepted triangle with t < min(tmax, FLT_MAX)
 *             (trace.glsl:97 `hit.hitT < maxDist`); a NaN tmax is never
 *             occluded.
 *
 * Threading: a context may be used from several host threads; each stream is
 * ordered by HIP.  Scenes are immutable after creation and may be traced
 * concurrently from any number of streams.
 */
#ifndef IGX_RT_C_H
#define IGX_RT_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_MISS 0xFFFFFFFFu          /* noRayHit, primitive.glsl:7 */
#define RT_NO_HIT_T 3.4028235e38f    /* noHit (FLT_MAX), primitive.glsl:6 */

/* Status codes. */
enum {
    RT_OK = 0,
    RT_ERR_INVALID = -1,     /* bad argument, size or handle */
    RT_ERR_OOM = -2,         /* device or host allocation failed */
    RT_ERR_HIP = -3,         /* a HIP runtime call failed */
    RT_ERR_UNSUPPORTED = -4, /* option combination not built */
    RT_ERR_ADMISSION = -5,   /* serving: request refused by the memory budget */
    RT_ERR_CANCELLED = -6,   /* serving: request cancelled or past deadline */
    RT_ERR_OVERFLOW = -7     /* a device-side bounded structure overflowed */
};

/* One ray, 32 bytes.  Replaces the GLSL `Ray{vec3 pos; vec3 dir;}`
 * (primitive.glsl:23-26) plus the `maxDist` and `prevHit` arguments of
 * traceOcclusion / traceGeometry (trace.glsl:9,70).  tmax is exclusive; use
 * RT_NO_HIT_T for the reference's closest-hit (it has no tmax). */
typedef struct rt_ray {
    float org[3];
    float tmax;
    float dir[3];
    uint32_t prev_id; /* triangle skipped by the query; RT_MISS for none */
} rt_ray;

/* One closest-hit result, 16 bytes: the graded subset of the GLSL `Hit`
 * (primitive.glsl:9-21): object, hitT, uv. */
typedef struct rt_hit {
    uint32_t id;
    float t;
    float u; /* weight of p1 */
    float v; /* weight of p2 */
} rt_hit;

typedef struct rt_ctx_s *rt_ctx;
typedef struct rt_scene_s *rt_scene;
typedef struct rt_bvh_s *rt_bvh;

enum { RT_BUILD_SAH_CPU = 0, RT_BUILD_LBVH_GPU = 1 };
enum { RT_NODE_BVH2 = 2, RT_NODE_BVH4 = 4, RT_NODE_BVH8 = 8 };

typedef struct rt_build_opts {
    uint32_t builder;     /* RT_BUILD_* */
    uint32_t node_width;  /* RT_NODE_BVH2 / BVH4 / BVH8 */
    uint32_t max_leaf;    /* most triangles per leaf, 1..8 (default 4); both
                             builders make a leaf of a subtree this small only
                             when its SAH leaf cost does not exceed splitting */
    uint32_t sah_bins;    /* binned-SAH bins per axis, 4..256 (default 32) */
    float pad_rel;        /* conservative box padding, relative to the scene's
                             largest |coordinate| (default 2^-16) */
    uint32_t keep_source; /* keep the input-order triangles on the device for
                             the brute-force entry points (default 1) */
    uint32_t presplit;    /* host SAH only: most pieces a large triangle is
                             cut into before the build, 1..64 (default 1 = no
                             splitting).  A triangle whose box is larger than
                             a quarter of the mean (half surface area) is
                             clipped at the midpoint of its box's longest
                             axis, recursively, until its pieces are that
                             small; the pieces' tight boxes enter the SAH
                             build and the triangle's record appears in each
                             leaf holding a piece.
                             Results do not change (a triangle tested twice
                             gives the same (t, id)); scenes of long, thin,
                             overlapping triangles trace faster. */
    uint32_t reserved;
} rt_build_opts;

/* Fills opts with the defaults listed above (binned SAH on the host, BVH4). */
void rt_build_opts_default(rt_build_opts *opts);

typedef struct rt_scene_info {
    uint32_t num_tris;
    uint32_t node_width;
    uint32_t num_nodes;
    uint32_t root_ref;
    uint32_t max_depth;   /* nodes on the longest root-to-leaf path */
    uint32_t max_stack;   /* traversal-stack entries a query can need */
    uint32_t num_leaves;
    uint32_t node_bytes;  /* bytes per node (64 or 128) */
    uint64_t nodes_total_bytes;
    uint64_t tris_total_bytes;
    float bounds_lo[3];
    float bounds_hi[3];
    double build_ms;      /* host build or GPU build time */
    uint32_t num_spheres; /* analytic primitives (rt_scene_set_analytic) */
    uint32_t num_cubes;
    uint32_t num_planes;
    uint32_t pad_;
} rt_scene_info;

/* Launch parameters.  They change speed only, never results. */
typedef struct rt_trace_opts {
    uint32_t block_size;  /* threads per block: 256, or 128 (LDS ring 8-32) */
    uint32_t lds_stack;   /* LDS ring entries per lane: 1, 4, 8, 16, 32, or 0
                             for the query's default (closest 16, occlusion 8) */
    uint32_t waves_per_cu;/* persistent grid size in resident waves per CU,
                             capped by occupancy; 0 = the query's default
                             (dynamic schedule 12, static 24) */
    uint32_t schedule;    /* 0 = dynamic (idle lanes take new rays as they
                             finish), 1 = static chunks of 64 rays, 2 = the
                             query's default (closest dynamic, occlusion
                             static); wave64 masks always use static chunks */
    uint32_t refill;      /* dynamic: idle lanes before a wave hands out new
                             rays, 1..64, or 0 for the query's default
                             (closest 8, occlusion 16; batches 16) */
    uint32_t chunk;       /* dynamic: rays a wave reserves at a time,
                             16..65536 (default 64) */
    uint32_t pool_shift;  /* dynamic: n >> pool_shift rays form a pool shared
                             by all waves, the rest is split evenly over the
                             waves; 1..32 (32 = no pool), 0 = the query's
                             default (single queries 8, batches 3) */
    uint32_t reserved;
} rt_trace_opts;

void rt_trace_opts_default(rt_trace_opts *opts);

/* ---- context ----------------------------------------------------------- */
int rt_ctx_create(int device, rt_ctx *out);
int rt_ctx_destroy(rt_ctx ctx);
/* Message of the last failing call on this context (never NULL). */
const char *rt_last_error(rt_ctx ctx);
/* Message of the last failing call made without a context (host builder). */
const char *rt_last_error_global(void);
int rt_sync(rt_ctx ctx, void *stream);
int rt_device_count(int *count);

/* ---- host BVH (no device needed) --------------------------------------- */
/* Builds the BVH on the host from input-order triangles tri_xyz[n][9]
 * (p0 p1 p2), the raw positions the reference's Triangle carries
 * (primitive.glsl:28-38, without the packed normals). */
int rt_bvh_build(const float *tri_xyz, uint32_t n, const rt_build_opts *opts,
                 rt_bvh *out);
int rt_bvh_info(rt_bvh bvh, rt_scene_info *info);
/* Host views of the built arrays: nodes (info.num_nodes * info.node_bytes
 * bytes) and triangles (num_tris * 48 bytes: p0.xyz id | e1.xyz 0 | e2.xyz 0,
 * in leaf order).  Valid until rt_bvh_free. */
int rt_bvh_arrays(rt_bvh bvh, const void **nodes, const void **tris);
int rt_bvh_free(rt_bvh bvh);

/* ---- scenes ------------------------------------------------------------ */
/* Every builder refuses a scene with an infinite coordinate (RT_ERR_INVALID,
 * also from rt_bvh_build): a box with an infinite side has no quantised
 * form.  NaN coordinates are accepted; box arithmetic ignores them, and the
 * traversal rejects the NaN t they produce. */
/* tri_xyz is a HOST pointer (n x 9 floats). */
int rt_scene_create(rt_ctx ctx, const float *tri_xyz, uint32_t n,
                    const rt_build_opts *opts, rt_scene *out);
/* The reference's own triangle buffer: n records of its std430 Triangle
 * (primitive.glsl:28-38; the `triangles[]` SSBO, scene.glsl:26-28), 48 bytes
 * each: p0.xyz n0 p1.xyz n1 p2.xyz n2, with n* the half2-packed spheremap
 * vertex normals (decodeSpheremap, primitive.glsl:95-104).  HOST pointer.
 * Builds like rt_scene_create from the positions and keeps the normals on
 * the device, so rt_raygen_closest interpolates them (trace.glsl:48-57)
 * without a normals argument.  Object ids are record indices. */
int rt_scene_create_packed(rt_ctx ctx, const void *triangles, uint32_t n, const rt_build_opts *opts,
                           rt_scene *out);
/* d_tri_xyz is a DEVICE pointer; builder must be RT_BUILD_LBVH_GPU. */
int rt_scene_create_device(rt_ctx ctx, const float *d_tri_xyz, uint32_t n,
                           const rt_build_opts *opts, void *stream,
                           rt_scene *out);
int rt_scene_from_bvh(rt_ctx ctx, rt_bvh bvh, const float *tri_xyz,
                      rt_scene *out);
/* The binary tree every GPU LBVH build starts from (30-bit Morton codes of
 * the triangle box centres in the centroid box, stable radix sort, Karras
 * radix tree with equal codes ordered by index, bottom-up boxes), for
 * inspection and export.  All pointers are DEVICE pointers the caller
 * allocates: morton[n] sorted codes, order[n] triangle index at each sorted
 * position, left[n-1] / right[n-1] children of internal node i (node 0 is
 * the root; bit 31 set = sorted leaf position), boxes[(2n-1)*6] unpadded
 * lo.xyz hi.xyz of internal nodes 0..n-2 then leaves 0..n-1. */
int rt_lbvh_binary(rt_ctx ctx, const float *d_tri_xyz, uint32_t n,
                   uint32_t *morton, uint32_t *order, uint32_t *left,
                   uint32_t *right, float *boxes, void *stream);
int rt_scene_info_get(rt_scene scene, rt_scene_info *info);
/* Analytic primitives of the scene, replacing any set before; host arrays.
 * They take the object ids after the n triangles, in the reference's order
 * (trace.glsl:22-46): spheres n.., then cubes, then planes.
 *   spheres[ns][4]: centre.xyz, radius (the shader squares w, so w is the
 *                   radius; primitive.glsl:173-210)
 *   cubes[nc][6]:   start.xyz, end.xyz (Cube{xy0, z0_x1, yz1},
 *                   primitive.glsl:40-44,286-333)
 *   planes[np][4]:  normal.xyz (normalised by the test), offset w: the plane
 *                   dot(p, normalize(normal)) = w (primitive.glsl:212-237)
 * Every query, linear scan, raygen and shadow-tile call on the scene tests
 * them after the triangles with the reference's per-type rules (spheres and
 * planes: 0 <= t < best; cubes: tmin <= best, a negative tmin from inside
 * included; so a cube wins a tie and a sphere or plane loses it).  The cost
 * is linear in their number per ray (the reference's scenes hold a handful).
 * Replaces the objects the reference re-uploads per frame
 * (NielsScene::update, test/scene/niels_scene.cpp:61-71): waits for the
 * device to be idle first, so no query in flight reads the old set.  Batches
 * and the server test them too (a pass after their traversal launch); they
 * hold the set a scene had when it joined them, so a batch or server whose
 * scene replaced its set since refuses to trace that scene (RT_ERR_INVALID;
 * make a new batch).  The path tracer takes triangle-only scenes. */
int rt_scene_set_analytic(rt_ctx ctx, rt_scene scene, const float *spheres, uint32_t ns, const float *cubes,
                          uint32_t nc, const float *planes, uint32_t np);
/* Copies the device node and triangle arrays to host memory (sizes from
 * rt_scene_info_get). */
int rt_scene_download(rt_scene scene, void *nodes, void *tris);
int rt_scene_destroy(rt_scene scene);

/* ---- queries (device pointers) ---------------------------------------- */
/* Closest hit; replaces traceGeometry (trace.glsl:9-66) as called by
 * raygen.comp:31. */
int rt_trace_closest(rt_ctx ctx, rt_scene scene, const rt_ray *rays,
                     uint32_t n, rt_hit *hits, void *stream);
/* Any hit; replaces traceOcclusion (trace.glsl:70-98) as called by
 * nv_all.shadow.comp:119,125.  flags[i] = 1 if ray i is occluded. */
int rt_trace_occluded(rt_ctx ctx, rt_scene scene, const rt_ray *rays,
                      uint32_t n, uint8_t *flags, void *stream);
/* Same query, wave64 ballot output as nv_all.shadow.comp:128-140:
 * bit (i & 63) of masks[i >> 6] = ray i occluded; ceil(n/64) words. */
int rt_trace_occluded_mask(rt_ctx ctx, rt_scene scene, const rt_ray *rays,
                           uint32_t n, uint64_t *masks, void *stream);
/* Extended forms: launch options and optional per-ray counters
 * counts[2*i+0] = BVH nodes fetched, counts[2*i+1] = triangles tested. */
int rt_trace_closest_ex(rt_ctx ctx, rt_scene scene, const rt_ray *rays,
                        uint32_t n, rt_hit *hits, const rt_trace_opts *opts,
                        uint32_t *counts, void *stream);
int rt_trace_occluded_ex(rt_ctx ctx, rt_scene scene, const rt_ray *rays,
                         uint32_t n, uint8_t *flags, uint64_t *masks,
                         const rt_trace_opts *opts, uint32_t *counts,
                         void *stream);
/* No counterpart in the reference.  rt_trace_closest for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same query and the same
 * results, but it calls no hipMalloc, hipFree or other allocating routine.
 * The traversal's spill stack and work counter live in a fixed scratch
 * array of RT_NOALLOC_SCRATCH_BYTES inside the library, one per device,
 * resolved by rt_ctx_create and held until the process ends.  The
 * persistent grid is sized to fit it; a scene whose traversal stack does
 * not fit even one block is refused (RT_ERR_INVALID).  Every call on a
 * device shares that device's scratch: calls that overlap in time on one
 * device (other streams, other threads, other contexts) race and must not
 * be issued; calls ordered on one stream are fine.  The default launch
 * shape of rt_trace_closest is used. */
#define RT_NOALLOC_SCRATCH_BYTES (64u << 20)
int rt_trace_closest_noalloc(rt_ctx ctx, rt_scene scene, const rt_ray *rays,
                             uint32_t n, rt_hit *hits, void *stream);
/* The reference's own algorithm on the GPU: every ray scans every triangle
 * in input order (trace.glsl:25-27, :78-79), including the NaN behaviour of
 * the strict `t >= hit.hitT` reject.  Needs keep_source. */
int rt_trace_closest_brute(rt_ctx ctx, rt_scene scene, const rt_ray *rays,
                           uint32_t n, rt_hit *hits, void *stream);
int rt_trace_occluded_brute(rt_ctx ctx, rt_scene scene, const rt_ray *rays,
                            uint32_t n, uint8_t *flags, void *stream);

/* ---- callers ------------------------------------------------------------ */
/* The screen camera of camera.glsl:72-87: a pixel's ray leaves `eye` toward
 * p0 + cx (p1 - p0) + cy (p2 - p0), cx = (x + jitter.x) / width,
 * cy = 1 - (y + jitter.y) / height (camera.glsl:115-116; jitter 0.5 = pixel
 * centres). */
typedef struct rt_camera {
    float eye[3];
    float p0[3], p1[3], p2[3];
    float jitter[2];
} rt_camera;

/* raygen.comp as one call: a primary ray per pixel, closest hit, and the
 * reference's packing (raygen.comp:39-51) into two float4 images [h][w]:
 * dir_object = (dir * t, bits(id)) on a hit, (dir, bits(RT_MISS)) on a miss;
 * uv_normal = (u, v, bits of encodeNormal(n)) written on hits only, where n
 * is the interpolated shading normal from `normals` (device, 3 spheremap
 * half2-packed vertex normals per input triangle, trace.glsl:50-57); when
 * normals is NULL, from the scene's own normals if it was created by
 * rt_scene_create_packed, else the geometric normal (primitive.glsl:281).
 * An analytic object's hit packs its geometry normal (trace.glsl:58-59). */
int rt_raygen_closest(rt_ctx ctx, rt_scene scene, const rt_camera *cam,
                      uint32_t width, uint32_t height, const uint32_t *normals,
                      float *dir_object, float *uv_normal, void *stream);
/* calculatePrimary (camera.glsl:113-138) in full: the projection types of
 * camera.glsl:5-11 and the per-pixel offset rand(loc + randLoc)
 * (camera.glsl:115; rand, rand_util.glsl:119-133).  The fields are the
 * reference's Camera (camera.glsl:14-44) that it reads, and Seed.randomX/Y
 * (rand_util.glsl:5-10). */
enum {
    RT_PROJ_DEFAULT = 0,        /* calculateScreen over p0..p2 */
    RT_PROJ_OMNI = 1,           /* calculateOmni (equirectangular) */
    RT_PROJ_OMNI_STEREO_TB = 2, /* two omni eyes, top (right) / bottom (left) */
    RT_PROJ_STEREO_TB = 3,      /* bottom half p0..p2, top half p3..p5 */
    RT_PROJ_OMNI_STEREO_LR = 4, /* two omni eyes, left / right halves */
    RT_PROJ_STEREO_LR = 5       /* left half p0..p2, right half p3..p5 */
};
typedef struct rt_primary_camera {
    float eye[3];
    float p0[3], p1[3], p2[3]; /* left (or only) screen: corner, +x, +y */
    float p3[3], p4[3], p5[3]; /* right screen of the stereo projections */
    float ipd;                 /* omni eye offset: ipd * 5e-4 each side */
    uint32_t projection;       /* RT_PROJ_* */
    uint32_t jitter_mode;      /* 0: rand(loc + seed) per pixel, as the
                                  reference; 1: the fixed offset below */
    float seed[2];             /* randLoc: Seed.randomX, Seed.randomY */
    float jitter[2];           /* fixed offset in pixels (jitter_mode 1) */
} rt_primary_camera;

/* raygen.comp with calculatePrimary in full: as rt_raygen_closest, with the
 * ray of each pixel from `cam`.  Omnidirectional rays leave eye + the ipd
 * offset; dir_object packs dir * t as raygen.comp:39 does.  The reference's
 * sin/cos are driver-defined; here they are correctly rounded FP32 sine and
 * cosine of the FP32 argument (evaluated through FP64), so offsets and
 * directions agree with the reference's to the accuracy of its sin. */
int rt_raygen_primary(rt_ctx ctx, rt_scene scene, const rt_primary_camera *cam,
                      uint32_t width, uint32_t height, const uint32_t *normals,
                      float *dir_object, float *uv_normal, void *stream);

/* nv_all.shadow.comp's output layout: occlusion of rays[s][y][x] as one
 * 64-bit ballot per 16x4 pixel tile per sample, word tile.x + tile.y *
 * tilesX + s * tilesX * tilesY, bit (x & 15) | (y & 3) << 4
 * (nv_all.shadow.comp:128-141, light.glsl:221-229).  Pixels outside the
 * image and rays with tmax 0 or NaN contribute 0 bits.  Test a bit with a
 * 64-bit mask: the reference's reader (light_rt.glsl:39) shifts a 32-bit 1,
 * undefined for the tile's rows 2 and 3. */
int rt_shadow_tiles(rt_ctx ctx, rt_scene scene, const rt_ray *rays,
                    uint32_t width, uint32_t height, uint32_t samples,
                    uint64_t *tiles, void *stream);

/* The reference's Light (primitive.glsl:46-54), 32 bytes: position; radius
 * and origin as half2 (x = radius, y = origin, unpackHalf2x16); direction
 * packed by encodeNormal (decodeNormal, primitive.glsl:89-92); colour and
 * type, the type in color_type[1] >> 16 (light.glsl:9-12: 0 directional,
 * 2 point). */
typedef struct rt_light {
    float pos[3];
    uint32_t rad_origin;
    uint32_t dir[2];
    uint32_t color_type[2];
} rt_light;

/* nv_all.shadow.comp (:50-143) as one call, after raygen: for every pixel
 * and sample, the hit position eye + dir_object.xyz (:84) of the pixel's
 * primary hit (dir_object: DEVICE float4 [height][width], as
 * rt_raygen_closest / rt_raygen_primary write it), a light sample from
 * rand(loc + seed) over 128x128 tiles plus hammersley(sample, samples)
 * (:88-94) through getDirToLight (light.glsl:98-133: a point on a point
 * light's sphere, or a direction in a directional light's cone), then
 * traceOcclusion toward it with prevHit = the pixel's object: a point light
 * only while its distance d is in [origin, radius), with maxDist = d -
 * origin (:104-117), a directional light with maxDist = noHit (:121-125).
 * Output as rt_shadow_tiles: one 64-bit ballot per 16x4 tile per sample.
 * A tile none of whose pixels hit is written as 0 (the reference leaves it
 * unwritten outside DEBUG builds, :69-82).  sin/cos/acos are correctly
 * rounded FP32, as in rt_raygen_primary. */
int rt_shadow_pass(rt_ctx ctx, rt_scene scene, const float *dir_object,
                   const float eye[3], const rt_light *light,
                   const float seed[2], uint32_t width, uint32_t height,
                   uint32_t samples, uint64_t *tiles, void *stream);

/* ---- wavefront path tracing ---------------------------------------------- */
/* Diffuse path tracing over the closest-hit query: per frame, one path per
 * (pixel, sample) with PCG32 seeded per pixel (stream = sample), a jittered
 * screen-camera ray, then max_depth cosine-weighted bounces off geometric
 * normals (albedo, Russian roulette from the second bounce, constant sky
 * 0.5 + 0.5 dir.y on a miss).  Each bounce is one persistent launch over a
 * device-side ray queue, compacted with one atomic per wave; the frame is
 * captured in a hipGraph on first use (use_graph) and replayed.  image is
 * a device float[height][width] (mean radiance over the samples). */
typedef struct rt_pathtracer_s *rt_pathtracer;
typedef struct rt_pt_opts {
    uint32_t width, height, spp, max_depth;
    float rr_survive;  /* survival probability (default 0.8) */
    float albedo;      /* default 0.7 */
    uint32_t chunk;    /* bounce feed: rays a wave reserves at a time,
                          16..4096 (0 = 64) */
    uint32_t refill;   /* bounce feed: idle lanes before a wave hands out
                          new rays, 1..64 (0 = 8) */
} rt_pt_opts;
void rt_pt_opts_default(rt_pt_opts *opts);
int rt_pt_create(rt_ctx ctx, rt_scene scene, const rt_camera *cam,
                 const rt_pt_opts *opts, rt_pathtracer *out);
int rt_pt_render(rt_pathtracer pt, uint64_t frame_seed, float *image,
                 int use_graph, void *stream);
/* Rays traced at each depth 0..max_depth by the last frame (synchronizes). */
int rt_pt_counts(rt_pathtracer pt, uint32_t *counts, uint32_t n);
/* The first n rays that the last frame traced at `depth`, in queue order
 * (which varies from frame to frame), copied to out (host or device memory;
 * synchronizes).  Only the last two depths are kept: depth must be
 * max_depth or max_depth - 1, and n at most width * height * spp; take the
 * count from rt_pt_counts.  No counterpart in the reference, whose passes
 * keep no ray queue: it lets tests and benchmarks trace the same rays
 * through the plain queries (node counts, hits). */
int rt_pt_queue_rays(rt_pathtracer pt, uint32_t depth, rt_ray *out, uint32_t n);
int rt_pt_destroy(rt_pathtracer pt);

/* ---- autotuner ---------------------------------------------------------- */
/* Times candidate launch shapes (block 128/256, LDS ring 8/16/32, static or
 * dynamic schedule with chunk 32/64/128 and refill 4/16/32) of one query
 * kind (RT_QUERY_*) on a device ray sample, `iters` timed launches each, and
 * writes the fastest to *best (its mean time to *best_ms if not NULL).
 * `out` receives the query's results.  Shapes never change results; node
 * width is chosen by building the scene in each width and tuning each. */
int rt_autotune(rt_ctx ctx, rt_scene scene, uint32_t kind, const rt_ray *rays,
                uint32_t n, void *out, uint32_t iters, rt_trace_opts *best,
                float *best_ms, void *stream);

/* ---- multi-scene batches ------------------------------------------------ */
/* Many scenes and many ray batches of uneven size in ONE persistent launch:
 * the batch holds a device table of its scenes (all of one node width); a
 * query names a scene of the batch, a kind and its device buffers.  Results
 * equal rt_trace_closest / rt_trace_occluded of each query alone, analytic
 * primitives included.  A scene must keep the analytic set it had when the
 * batch was made (rt_batch_trace fails otherwise). */
typedef struct rt_batch_s *rt_batch;
enum { RT_QUERY_CLOSEST = 0, RT_QUERY_OCCLUDED = 1 };
typedef struct rt_query {
    uint32_t scene;      /* index into the batch's scene list */
    uint32_t kind;       /* RT_QUERY_CLOSEST or RT_QUERY_OCCLUDED */
    uint32_t n;          /* rays */
    uint32_t reserved;
    const rt_ray *rays;  /* device */
    void *out;           /* device: rt_hit[n] (closest) or uint8_t[n] */
} rt_query;

int rt_batch_create(rt_ctx ctx, const rt_scene *scenes, uint32_t nscenes,
                    rt_batch *out);
/* Launch shape from opts (lds_stack 8, 16 or 32; dynamic schedule). */
int rt_batch_trace(rt_ctx ctx, rt_batch batch, const rt_query *queries,
                   uint32_t nq, const rt_trace_opts *opts, void *stream);
int rt_batch_destroy(rt_batch batch);
/* The autotuner for batch launches: block 128/256, LDS ring 8/16/32, chunk
 * 64/128/256, refill 16/24/32, timed on these queries (whose outputs are
 * written).  Node width is chosen by building the batch in each width. */
int rt_batch_autotune(rt_ctx ctx, rt_batch batch, const rt_query *queries,
                      uint32_t nq, uint32_t iters, rt_trace_opts *best,
                      float *best_ms, void *stream);

/* ---- serving ----------------------------------------------------------- */
/* A request loop over resident scenes with continuous batching: requests are
 * admitted FIFO while the server's HBM budget for request buffers has room
 * (rays + results), the admitted ones form one multi-scene launch ("tick")
 * on a free stream, and requests that arrive while ticks run join the next
 * tick on another stream, so their copies and kernels overlap on the device.
 * A plain server (plain = 1) serves one request at a time with one launch.
 * Results of a request are the same alone and in any mix. */
typedef struct rt_server_s *rt_server;

typedef struct rt_server_opts {
    uint64_t mem_budget;     /* bytes of HBM for request buffers; 0 = half of
                                the free HBM at creation, at most 64 GiB */
    uint32_t streams;        /* ticks in flight (default 4, at most 16) */
    uint32_t max_tick_rays;  /* rays per tick (default 8,388,608) */
    uint32_t plain;          /* 1 = one request per launch, one at a time */
    uint32_t node_width;     /* width of every scene served (default 4) */
    uint32_t max_scenes;     /* capacity of the scene table (default 1024) */
    uint32_t simulate;       /* 1 = no device: ticks complete after a modelled
                                time (for host tests of the scheduler);
                                2 = the same, but every tick completes with
                                a HIP error (RT_ERR_HIP, failure-path tests) */
    rt_trace_opts launch;    /* tick launch shape (zero = defaults); the
                                default chunk and refill are 256 and 24 for
                                a batching server, 64 and 16 for a plain one */
} rt_server_opts;

void rt_server_opts_default(rt_server_opts *opts);

enum { RT_REQ_HOST = 0, RT_REQ_DEVICE = 1 };
typedef struct rt_request {
    uint32_t scene;        /* server scene id */
    uint32_t kind;         /* RT_QUERY_CLOSEST or RT_QUERY_OCCLUDED */
    uint32_t n;            /* rays */
    uint32_t memory;       /* RT_REQ_HOST: rays/out are host memory (pinned
                              for asynchronous copies), staged through the
                              server's HBM budget; RT_REQ_DEVICE: device
                              buffers, nothing staged */
    const rt_ray *rays;
    void *out;             /* rt_hit[n] or uint8_t[n] */
    double deadline_ms;    /* cancel if not started this long after submit;
                              0 = no deadline */
    void *stream;          /* RT_REQ_DEVICE: the caller's hipStream_t that
                              produced the rays and last used the output
                              buffer (NULL = the null stream).  Submit
                              records an event on it and the tick's launch
                              waits for that event, so work still queued
                              there (a pending kernel writing the memory a
                              caching allocator just handed out as `out`)
                              finishes before the tick writes.  Ignored for
                              RT_REQ_HOST. */
} rt_request;

enum { RT_TICKET_PENDING = 0, RT_TICKET_RUNNING = 1, RT_TICKET_DONE = 2, RT_TICKET_FAILED = 3 };

typedef struct rt_server_stats {
    uint64_t submitted, completed, failed, ticks, rays;
    uint64_t mem_budget, mem_peak;
    double busy_ms;        /* wall time with at least one tick in flight */
    double mean_requests_per_tick;
} rt_server_stats;

int rt_server_create(rt_ctx ctx, const rt_server_opts *opts, rt_server *out);
/* Registers a built scene (the caller keeps ownership; width must match). */
int rt_server_add_scene(rt_server s, rt_scene scene, uint32_t *id);
/* Builds a scene from host triangles with the GPU LBVH builder on the
 * server's build stream while ticks keep running; the server owns it.
 * Requests for the id wait until the build has finished. */
int rt_server_upload_scene(rt_server s, const float *tri_xyz, uint32_t n, uint32_t *id);
/* Non-blocking.  RT_ERR_ADMISSION if the request can never fit the budget.
 * A device request records an event on req->stream (see rt_request).  Once
 * its ticket is DONE, its results are in the output buffer for any stream
 * or the host; the caller frees or reuses the buffers only after that. */
int rt_server_submit(rt_server s, const rt_request *req, uint64_t *ticket);
int rt_server_cancel(rt_server s, uint64_t ticket);
/* state = RT_TICKET_*; status = the request's RT_* result once finished;
 * latency_ms = submit to completion. */
int rt_server_poll(rt_server s, uint64_t ticket, int *state, int *status, double *latency_ms);
/* Blocks until the ticket finishes (or timeout_ms passes, RT_ERR_CANCELLED
 * is then NOT implied: state stays pending/running). */
int rt_server_wait(rt_server s, uint64_t ticket, double timeout_ms, int *state, int *status, double *latency_ms);
/* Drops finished tickets' records (latencies already read). */
int rt_server_forget(rt_server s, uint64_t ticket);
int rt_server_stats_get(rt_server s, rt_server_stats *stats);
int rt_server_destroy(rt_server s);

/* Device error word of the context: non-zero after a kernel hit a bounded
 * structure's limit (reads and clears it; synchronizes the stream). */
int rt_check_device_error(rt_ctx ctx, void *stream, uint32_t *code);

#ifdef __cplusplus
}
#endif
#endif /* IGX_RT_C_H */
