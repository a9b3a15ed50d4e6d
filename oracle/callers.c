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

/* ---- nv_all.shadow.comp:50-143 ------------------------------------------- */

static float acos_r(float x) { return (float)acos((double)x); }
static float glsl_max(float x, float y) { return x < y ? y : x; }
static float glsl_min(float x, float y) { return y < x ? y : x; }

static void rand2(float px, float py, float *rx, float *ry) {
    *rx = rand1(px, py);
    *ry = rand1(px * 1103515245.f + 12345.f, py * 1103515245.f + 12345.f);
}

/* hammersley, res/shaders/rand_util.glsl:96-111 */
static void hammersley(uint32_t i, uint32_t n, float *hx, float *hy) {
    uint32_t v = (i << 16u) | (i >> 16u);
    v = ((v & 0x55555555u) << 1u) | ((v & 0xAAAAAAAAu) >> 1u);
    v = ((v & 0x33333333u) << 2u) | ((v & 0xCCCCCCCCu) >> 2u);
    v = ((v & 0x0F0F0F0Fu) << 4u) | ((v & 0xF0F0F0F0u) >> 4u);
    v = ((v & 0x00FF00FFu) << 8u) | ((v & 0xFF00FF00u) >> 8u);
    *hx = (float)i / (float)n;
    *hy = (float)v * 2.3283064365386963e-10f;
}

static void cross3(const float *x, const float *y, float *o) {
    o[0] = x[1] * y[2] - y[1] * x[2];
    o[1] = x[2] * y[0] - y[2] * x[0];
    o[2] = x[0] * y[1] - y[0] * x[1];
}

/* getDirToLight, res/shaders/light.glsl:98-133 (brightness not needed). */
void orc_dir_to_light(const orc_light *L, const float pos[3], float rx, float ry, float l[3], float *dist) {
    const float pi = 3.1415927410125732421875f;
    *dist = -1.f;
    float ro_x = glsl_max(half_to_float(L->rad_origin & 0xFFFFu), 0.f);
    float ro_y = glsl_max(half_to_float(L->rad_origin >> 16), 0.f);
    ro_y = glsl_min(ro_y, ro_x);
    if ((L->color_type[1] >> 16) == 2u) {
        for (int k = 0; k < 3; ++k) l[k] = pos[k] - L->pos[k];
        *dist = sqrtf(dot3(l[0], l[1], l[2], l[0], l[1], l[2]));
        /* randomPointOnHemisphere / randomPointOnUnitSphere, rand_util.glsl:32-51 */
        const float px = (2.f * pi) * rx, py = acos_r(1.f - 2.f * ry);
        const float spx = sin_r(px), spy = sin_r(py), cpx = cos_r(px), cpy = cos_r(py);
        const float n[3] = {spx * cpy, spx * spy, cpx};
        for (int k = 0; k < 3; ++k) l[k] = pos[k] - (L->pos[k] + n[k] * ro_y);
    } else {
        /* getSunDirection(random, normalize(decodeNormal(dir)), radius), rand_util.glsl:55-86 */
        float d[3] = {(((float)(L->dir[0] >> 16) / 65535.f) * 2.f) - 1.f,
                      (((float)(L->dir[0] & 65535u) / 65535.f) * 2.f) - 1.f,
                      (((float)L->dir[1] / 65535.f) * 2.f) - 1.f};
        normalize3(d);
        const float h = cos_r(ro_x);
        const float phi = (2.f * pi) * rx;
        const float z = h + (1.f - h) * ry;
        const float sin_t = sqrtf(1.f - z * z);
        const float x = cos_r(phi) * sin_t, y = sin_r(phi) * sin_t;
        const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
        const uint32_t xm = (ax - ay < 0.f && ax - az < 0.f) ? 1u : 0u;
        const uint32_t ym = ay - az < 0.f ? (1u ^ xm) : 0u;
        const uint32_t zm = 1u ^ (xm | ym);
        const float m[3] = {(float)xm, (float)ym, (float)zm};
        float b[3], t[3];
        cross3(d, m, b);
        cross3(b, d, t);
        for (int k = 0; k < 3; ++k) l[k] = (b[k] * x + t[k] * y) + d[k] * z;
    }
    normalize3(l);
}

/* The shadow rays of nv_all.shadow.comp:50-125 for every (sample, pixel):
 * rays[s][y][x] (8 floats) and trace[s][y][x] = 1 where the reference calls
 * traceOcclusion. */
void orc_shadow_rays(const float *dir_object, const float eye[3], const orc_light *L, const float seed[2],
                     uint32_t width, uint32_t height, uint32_t samples, float *rays, uint8_t *trace) {
    for (uint32_t s = 0; s < samples; ++s)
        for (uint32_t y = 0; y < height; ++y)
            for (uint32_t x = 0; x < width; ++x) {
                const size_t i = ((size_t)s * height + y) * width + x;
                const float *dob = dir_object + 4 * ((size_t)y * width + x);
                uint32_t object;
                memcpy(&object, dob + 3, 4);
                float *r = rays + 8 * i;
                memset(r, 0, 8 * sizeof(float));
                trace[i] = 0;
                if (object == MISS) continue;
                const float pos[3] = {eye[0] + dob[0], eye[1] + dob[1], eye[2] + dob[2]};
                float jx, jy, hx, hy, rx, ry;
                rand2((float)x + seed[0], (float)y + seed[1], &jx, &jy);
                float ux = ((float)x + jx) / 128.f, uy = ((float)y + jy) / 128.f;
                hammersley(s, samples, &hx, &hy);
                ux = ux + hx;
                uy = uy + hy;
                rand2(ux, uy, &rx, &ry);
                float l[3], dist;
                orc_dir_to_light(L, pos, rx, ry, l, &dist);
                float tmax = -1.f;
                if (dist >= 0.f) {
                    const float rad = half_to_float(L->rad_origin & 0xFFFFu), org = half_to_float(L->rad_origin >> 16);
                    if (dist >= org && dist < rad) tmax = dist - org;
                } else {
                    tmax = NO_HIT;
                }
                if (!(tmax >= 0.f)) continue;
                r[0] = pos[0];
                r[1] = pos[1];
                r[2] = pos[2];
                r[3] = tmax;
                r[4] = -l[0];
                r[5] = -l[1];
                r[6] = -l[2];
                memcpy(r + 7, &object, 4);
                trace[i] = 1;
            }
}

void orc_shadow_pass(const float *tri, uint32_t ntri, const orc_analytic *ana, const float *dir_object,
                     const float eye[3], const orc_light *L, const float seed[2], uint32_t width, uint32_t height,
                     uint32_t samples, uint64_t *tiles, int nthreads) {
    const size_t n = (size_t)width * height * samples;
    const uint32_t tiles_x = (width + 15u) / 16u, tiles_y = (height + 3u) / 4u;
    memset(tiles, 0, sizeof(uint64_t) * (size_t)tiles_x * tiles_y * samples);
    if (n == 0) return;
    float *rays = (float *)calloc(8 * n, sizeof(float));
    uint8_t *trace = (uint8_t *)calloc(n, 1), *occ = (uint8_t *)calloc(n, 1);
    if (!rays || !trace || !occ) {
        free(rays);
        free(trace);
        free(occ);
        return;
    }
    orc_shadow_rays(dir_object, eye, L, seed, width, height, samples, rays, trace);
    orc_occluded_brute_ana(tri, ntri, ana, rays, (uint32_t)n, occ, nthreads);
    for (uint32_t s = 0; s < samples; ++s)
        for (uint32_t y = 0; y < height; ++y)
            for (uint32_t x = 0; x < width; ++x) {
                const size_t i = ((size_t)s * height + y) * width + x;
                if (!trace[i] || !occ[i]) continue;
                /* indexToLight, light.glsl:221-229; bit (x & 15) | (y & 3) << 4 */
                const size_t w = (x >> 4) + (size_t)(y >> 2) * tiles_x + (size_t)s * tiles_x * tiles_y;
                tiles[w] |= 1ull << ((x & 15u) | ((y & 3u) << 4));
            }
    free(rays);
    free(trace);
    free(occ);
}
