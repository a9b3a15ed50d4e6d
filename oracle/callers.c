/*
 * callers.c - CPU restatement of the two callers of the query path.
 * TEST INFRASTRUCTURE ONLY.
 *
 * raygen: calculatePrimary with its projections (res/shaders/camera.glsl:
 * 49-138) and its rand(loc + seed) offset or a fixed one, traceGeometry by
 * the linear scan
 * (trace.glsl:9-66), and the packing of raygen.comp:39-51 with the normal of
 * trace.glsl:48-63 (decodeSpheremap primitive.glsl:95-104, interpolate
 * :146-151) or primitive.glsl:281, encoded by encodeNormal (:84-87).
 * Vector helpers: dot(x,y) = (x.x*y.x + x.y*y.y) + x.z*y.z; normalize(v) =
 * v / sqrt(dot(v,v)).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NO_HIT 3.4028235e38f
#define MISS 0xFFFFFFFFu

static inline float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}

static void normalize3(float *v) {
    const float l = sqrtf(dot3(v[0], v[1], v[2], v[0], v[1], v[2]));
    v[0] = v[0] / l;
    v[1] = v[1] / l;
    v[2] = v[2] / l;
}

/* IEEE binary16 -> binary32, exact. */
static float half_to_float(uint32_t h) {
    const uint32_t sign = (h >> 15) & 1u, ex = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
    float v;
    if (ex == 0)
        v = ldexpf((float)man, -24);
    else if (ex == 31)
        v = man ? NAN : INFINITY;
    else
        v = ldexpf((float)(man | 0x400u), (int)ex - 25);
    return sign ? -v : v;
}

static void decode_spheremap(uint32_t n, float out[3]) {
    float nx = half_to_float(n & 0xFFFFu), ny = half_to_float(n >> 16);
    const float l = dot3(nx, ny, 1.f, -nx, -ny, 1.f);
    const float r = sqrtf(l);
    nx = nx * r;
    ny = ny * r;
    out[0] = nx * 2.f + 0.f;
    out[1] = ny * 2.f + 0.f;
    out[2] = l * 2.f + -1.f;
}

/* The geometry normal of analytic object k (0-based after the triangles) hit
 * at t: the primitive's own test re-run with nothing to beat, which sets the
 * normal from the same operations as when it won the scan. */
static void analytic_normal(const orc_analytic *a, uint32_t k, const float o[3], const float d[3], float n[3]) {
    float t = INFINITY, uv[2];
    if (k < a->num_spheres) {
        orc_sphere(o, d, a->spheres + 4 * (size_t)k, 0, MISS, &t, uv, n);
        return;
    }
    k -= a->num_spheres;
    if (k < a->num_cubes) {
        orc_cube(o, d, a->cubes + 6 * (size_t)k, 0, MISS, &t, uv, n);
        return;
    }
    k -= a->num_cubes;
    orc_plane(o, d, a->planes + 4 * (size_t)k, 0, MISS, &t, uv, n);
}

/* sin and cos as the FP32 rounding of the FP64 function of the FP32
 * argument (GLSL leaves them to the driver). */
static float sin_r(float x) { return (float)sin((double)x); }
static float cos_r(float x) { return (float)cos((double)x); }
static float fract_f(float x) { return x - floorf(x); }

/* rand1(vec2), res/shaders/rand_util.glsl:123-125. */
static float rand1(float px, float py) { return fract_f(sin_r(px * 12.9898f + py * 78.233f) * 43758.5453f); }

/* calculateScreen, res/shaders/camera.glsl:72-87. */
static void screen_ray(const orc_camera *c, float cx, float cy, int right, float org[3], float dir[3]) {
    const float *p0 = right ? c->p3 : c->p0, *p1 = right ? c->p4 : c->p1, *p2 = right ? c->p5 : c->p2;
    for (int k = 0; k < 3; ++k) {
        const float r = p1[k] - p0[k], u = p2[k] - p0[k];
        dir[k] = (p0[k] + cx * r) + cy * u - c->eye[k];
        org[k] = c->eye[k];
    }
    normalize3(dir);
}

/* calculateOmni, res/shaders/camera.glsl:49-70. */
static void omni_ray(const orc_camera *c, float cx, float cy, int left, float org[3], float dir[3]) {
    const float pi = 3.1415927410125732421875f; /* rand_util.glsl:13 */
    const float sx = (cx - 0.5f) * (2.f * pi), sy = (0.5f - cy) * pi;
    const float sinx = sin_r(sx), siny = sin_r(sy), cosx = cos_r(sx), cosy = cos_r(sy);
    const float e = c->ipd * 5e-4f, sg = left ? -1.f : 1.f;
    org[0] = c->eye[0] + (cosx * e) * sg;
    org[1] = c->eye[1] + (0.f * e) * sg;
    org[2] = c->eye[2] + (sinx * e) * sg;
    dir[0] = sinx * cosy;
    dir[1] = siny;
    dir[2] = -cosx * cosy;
}

/* calculatePrimary, res/shaders/camera.glsl:113-138, with randLoc = seed
 * (camera.glsl:115) or a fixed offset. */
static void primary_ray(const orc_camera *c, uint32_t x, uint32_t y, float invx, float invy, float org[3],
                        float dir[3]) {
    float jx = c->jitter[0], jy = c->jitter[1];
    if (c->rand_jitter) {
        /* rand(vec2), rand_util.glsl:131-133 */
        const float px = (float)x + c->seed[0], py = (float)y + c->seed[1];
        jx = rand1(px, py);
        jy = rand1(px * 1103515245.f + 12345.f, py * 1103515245.f + 12345.f);
    }
    const float cx = ((float)x + jx) * invx;
    float cy = ((float)y + jy) * invy;
    cy = 1.f - cy;
    switch (c->projection) {
        case 1: omni_ray(c, cx, cy, 0, org, dir); break;
        case 2: omni_ray(c, cx, fract_f(cy * 2.f), cy < 0.5f, org, dir); break;
        case 4: omni_ray(c, fract_f(cx * 2.f), cy, cx < 0.5f, org, dir); break;
        case 3:
            if (cy < 0.5f) screen_ray(c, cx, cy * 2.f, 0, org, dir);
            else screen_ray(c, cx, cy * 2.f - 1.f, 1, org, dir);
            break;
        case 5:
            if (cx < 0.5f) screen_ray(c, cx * 2.f, cy, 0, org, dir);
            else screen_ray(c, cx * 2.f - 1.f, cy, 1, org, dir);
            break;
        default: screen_ray(c, cx, cy, 0, org, dir); break;
    }
}

void orc_primary_rays(const orc_camera *cam, uint32_t width, uint32_t height, float *rays) {
    const float invx = 1.0f / (float)width, invy = 1.0f / (float)height;
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            float *r = rays + 8 * ((size_t)y * width + x);
            float org[3], d[3];
            primary_ray(cam, x, y, invx, invy, org, d);
            r[0] = org[0];
            r[1] = org[1];
            r[2] = org[2];
            r[3] = NO_HIT;
            r[4] = d[0];
            r[5] = d[1];
            r[6] = d[2];
            const uint32_t miss = MISS;
            memcpy(r + 7, &miss, 4);
        }
}

void orc_raygen(const float *tri, uint32_t ntri, const uint32_t *normals, const float cam[14], uint32_t width,
                uint32_t height, float *dir_object, float *uv_normal, int nthreads) {
    orc_raygen_ana(tri, ntri, normals, NULL, cam, width, height, dir_object, uv_normal, nthreads);
}

void orc_raygen_ana(const float *tri, uint32_t ntri, const uint32_t *normals, const orc_analytic *ana,
                    const float cam[14], uint32_t width, uint32_t height, float *dir_object, float *uv_normal,
                    int nthreads) {
    orc_camera c;
    memset(&c, 0, sizeof c);
    memcpy(c.eye, cam, 3 * sizeof(float));
    memcpy(c.p0, cam + 3, 3 * sizeof(float));
    memcpy(c.p1, cam + 6, 3 * sizeof(float));
    memcpy(c.p2, cam + 9, 3 * sizeof(float));
    c.jitter[0] = cam[12];
    c.jitter[1] = cam[13];
    orc_raygen_primary(tri, ntri, normals, ana, &c, width, height, dir_object, uv_normal, nthreads);
}

void orc_raygen_primary(const float *tri, uint32_t ntri, const uint32_t *normals, const orc_analytic *ana,
                        const orc_camera *cam, uint32_t width, uint32_t height, float *dir_object, float *uv_normal,
                        int nthreads) {
    const uint32_t npix = width * height;
    if (npix == 0) return;
    float *rays = (float *)calloc(8 * (size_t)npix, sizeof(float));
    float *dirs = (float *)calloc(3 * (size_t)npix, sizeof(float));
    uint32_t *hits = (uint32_t *)calloc(4 * (size_t)npix, sizeof(uint32_t));
    if (!rays || !dirs || !hits) {
        free(rays);
        free(dirs);
        free(hits);
        return;
    }
    orc_primary_rays(cam, width, height, rays);
    for (uint32_t p = 0; p < npix; ++p) memcpy(dirs + 3 * (size_t)p, rays + 8 * (size_t)p + 4, 3 * sizeof(float));
    orc_closest_brute_ana(tri, ntri, ana, rays, npix, hits, NULL, nthreads);
    for (uint32_t p = 0; p < npix; ++p) {
        const float *d = dirs + 3 * (size_t)p;
        const uint32_t *h = hits + 4 * (size_t)p;
        float *o = dir_object + 4 * (size_t)p;
        if (h[0] != MISS) {
            const float t = u2f(h[1]);
            o[0] = d[0] * t;
            o[1] = d[1] * t;
            o[2] = d[2] * t;
            o[3] = u2f(h[0]);
        } else {
            o[0] = d[0];
            o[1] = d[1];
            o[2] = d[2];
            o[3] = u2f(MISS);
            continue;
        }
        const float u = u2f(h[2]), v = u2f(h[3]);
        float n[3];
        const float *t3 = tri + 9 * (size_t)h[0];
        if (h[0] >= ntri) {
            /* trace.glsl:58-59: objectNormal = geometryNormal. */
            analytic_normal(ana, h[0] - ntri, rays + 8 * (size_t)p, d, n);
        } else if (normals) {
            float a[3], b[3], c[3];
            decode_spheremap(normals[3 * (size_t)h[0]], a);
            decode_spheremap(normals[3 * (size_t)h[0] + 1], b);
            decode_spheremap(normals[3 * (size_t)h[0] + 2], c);
            const float bz = (1.f - u) - v;
            for (int k = 0; k < 3; ++k) n[k] = (u * b[k] + v * c[k]) + bz * a[k];
        } else {
            float e1[3], e2[3];
            for (int k = 0; k < 3; ++k) {
                e1[k] = t3[3 + k] - t3[k];
                e2[k] = t3[6 + k] - t3[k];
            }
            const float hh[3] = {d[1] * e2[2] - e2[1] * d[2], d[2] * e2[0] - e2[2] * d[0], d[0] * e2[1] - e2[0] * d[1]};
            const float av = dot3(e1[0], e1[1], e1[2], hh[0], hh[1], hh[2]);
            const float sg = av > 0.f ? -1.f : (av < 0.f ? 1.f : -0.f);
            normalize3(e1);
            normalize3(e2);
            n[0] = (e1[1] * e2[2] - e2[1] * e1[2]) * sg;
            n[1] = (e1[2] * e2[0] - e2[2] * e1[0]) * sg;
            n[2] = (e1[0] * e2[1] - e2[0] * e1[1]) * sg;
        }
        normalize3(n);
        const uint32_t hx = (uint32_t)((n[0] * 0.5f + 0.5f) * 65535.f);
        const uint32_t hy = (uint32_t)((n[1] * 0.5f + 0.5f) * 65535.f);
        const uint32_t hz = (uint32_t)((n[2] * 0.5f + 0.5f) * 65535.f);
        const uint32_t enc0 = (hx << 16) | hy;
        float *q = uv_normal + 4 * (size_t)p;
        q[0] = u;
        q[1] = v;
        q[2] = u2f(enc0);
        q[3] = u2f(hz);
    }
    free(rays);
    free(dirs);
    free(hits);
}
