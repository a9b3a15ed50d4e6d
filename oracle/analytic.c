/*
 * analytic.c - CPU restatement of the reference's analytic primitives.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): never linked into the product.
 *
 * Restates rayIntersectSphere (res/shaders/primitive.glsl:173-210),
 * rayIntersectPlane (:212-237) and rayIntersectCube (:286-333) one statement
 * at a time, in FP32 with no contraction.  Vector helpers as everywhere in
 * this oracle: dot(x,y) = (x.x*y.x + x.y*y.y) + x.z*y.z, normalize(v) =
 * v / sqrt(dot(v,v)), GLSL cross(x,y) = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x,
 * x.x*y.y - y.x*x.y), GLSL min(x,y) = y < x ? y : x and max(x,y) =
 * x < y ? y : x (the GLSL 4.60 definitions, which also fix NaN and signed
 * zero).
 *
 * asin and atan (sphere uv, :198-199) are the Cephes single-precision
 * polynomials (S. L. Moshier, Cephes Math Library asinf.c / atanf.c)
 * evaluated in FP32, one operation at a time; the kernels evaluate the same
 * operations, so sphere uv is bit-exact between them.  A |x| > 1 argument of
 * asin returns NaN (GLSL leaves it undefined).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static inline float dot3(const float *a, const float *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

static inline void normalize3(float *v) {
    const float l = sqrtf(dot3(v, v));
    v[0] = v[0] / l;
    v[1] = v[1] / l;
    v[2] = v[2] / l;
}

static inline void cross3(const float *x, const float *y, float *o) {
    o[0] = x[1] * y[2] - y[1] * x[2];
    o[1] = x[2] * y[0] - y[2] * x[0];
    o[2] = x[0] * y[1] - y[0] * x[1];
}

static inline float gmin(float x, float y) { return y < x ? y : x; }
static inline float gmax(float x, float y) { return x < y ? y : x; }

#define PIO2F 1.5707963267948966192f
#define PIO4F 0.7853981633974483096f

float orc_asinf(float xx) {
    float a, sign;
    if (xx < 0.f) {
        sign = -1.f;
        a = -xx;
    } else {
        sign = 1.f;
        a = xx;
    }
    if (a > 1.f) return NAN;
    float z;
    if (a < 1.0e-4f) {
        z = a;
    } else {
        float x;
        int flag;
        if (a > 0.5f) {
            z = 0.5f * (1.f - a);
            x = sqrtf(z);
            flag = 1;
        } else {
            x = a;
            z = x * x;
            flag = 0;
        }
        float p = 4.2163199048e-2f * z + 2.4181311049e-2f;
        p = p * z + 4.5470025998e-2f;
        p = p * z + 7.4953002686e-2f;
        p = p * z + 1.6666752422e-1f;
        z = ((p * z) * x) + x;
        if (flag) {
            z = z + z;
            z = PIO2F - z;
        }
    }
    return sign < 0.f ? -z : z;
}

float orc_atanf(float xx) {
    float x, y, sign;
    if (xx < 0.f) {
        sign = -1.f;
        x = -xx;
    } else {
        sign = 1.f;
        x = xx;
    }
    if (x > 2.414213562373095f) {
        y = PIO2F;
        x = -(1.f / x);
    } else if (x > 0.4142135623730950f) {
        y = PIO4F;
        x = (x - 1.f) / (x + 1.f);
    } else {
        y = 0.f;
    }
    const float z = x * x;
    float p = 8.05374449538e-2f * z - 1.38776856032e-1f;
    p = p * z + 1.99777106478e-1f;
    p = p * z - 3.33329491539e-1f;
    y = y + (((p * z) * x) + x);
    return sign < 0.f ? -y : y;
}

/* primitive.glsl:173-210. */
int orc_sphere(const float o[3], const float d[3], const float s[4], uint32_t obj, uint32_t prev, float *best_t,
               float uv[2], float n[3]) {
    const float dif[3] = {s[0] - o[0], s[1] - o[1], s[2] - o[2]};              /* :175 */
    const float t = dot3(dif, d);                                               /* :176 */
    const float q[3] = {dif[0] - t * d[0], dif[1] - t * d[1], dif[2] - t * d[2]}; /* :178 */
    const float q2 = dot3(q, q);                                                /* :179 */
    const float r2 = s[3] * s[3];                                               /* :180 */
    const int out = q2 > r2;                                                    /* :182 */
    const float hit_t = t - sqrtf(r2 - q2);                                     /* :183 */
    if (!(!out && obj != prev && hit_t >= 0.f && hit_t < *best_t)) return 0;    /* :185 */
    *best_t = hit_t;                                                            /* :187 */
    const float p[3] = {hit_t * d[0] + o[0], hit_t * d[1] + o[1], hit_t * d[2] + o[2]}; /* :189 */
    float nn[3] = {s[0] - p[0], s[1] - p[1], s[2] - p[2]};                      /* :190 */
    normalize3(nn);
    const float lat = orc_asinf(nn[2]);                                         /* :198 */
    float lon = orc_atanf(nn[1] / nn[0]);                                       /* :199 */
    if (lon != lon) lon = 0.f;                                                  /* :201-202 */
    const float c = 0.636619746685f * 0.5f;                                     /* :204 */
    uv[0] = lat * c + 0.5f;
    uv[1] = lon * c + 0.5f;
    if (n) memcpy(n, nn, sizeof nn);                                            /* :192 */
    return 1;
}

/* primitive.glsl:212-237. */
int orc_plane(const float o[3], const float d[3], const float pl[4], uint32_t obj, uint32_t prev, float *best_t,
              float uv[2], float n[3]) {
    float dir[3] = {pl[0], pl[1], pl[2]};
    normalize3(dir);                                                            /* :214 */
    const float nd[3] = {-dir[0], -dir[1], -dir[2]};
    const float dif = dot3(d, nd);                                              /* :215 */
    const float hit_t = -(dot3(o, nd) + pl[3]) / dif;                           /* :219 */
    if (!(hit_t >= 0.f && obj != prev && hit_t < *best_t)) return 0;            /* :221 */
    *best_t = hit_t;                                                            /* :223 */
    if (n) {                                                                    /* :224 */
        n[0] = dif > 0.f ? nd[0] : dir[0];
        n[1] = dif > 0.f ? nd[1] : dir[1];
        n[2] = dif > 0.f ? nd[2] : dir[2];
    }
    const float p[3] = {hit_t * d[0] + o[0], hit_t * d[1] + o[1], hit_t * d[2] + o[2]}; /* :226 */
    const float z1[3] = {0.f, 0.f, 1.f}, x1[3] = {1.f, 0.f, 0.f};
    float px[3], pz[3];
    cross3(pl, z1, px);                                                         /* :228 */
    cross3(pl, x1, pz);                                                         /* :229 */
    uv[0] = dot3(p, px);                                                        /* :231 */
    uv[1] = dot3(p, pz);
    return 1;
}

/* primitive.glsl:286-333; c = start.xyz, end.xyz (Cube{xy0, z0_x1, yz1}). */
int orc_cube(const float o[3], const float d[3], const float c[6], uint32_t obj, uint32_t prev, float *best_t,
             float uv[2], float n[3]) {
    const float rev[3] = {1.f / d[0], 1.f / d[1], 1.f / d[2]};                /* :288 */
    const float *start = c, *end = c + 3;                                       /* :290-291 */
    float sd[3], ed[3], mi[3], ma[3];
    for (int k = 0; k < 3; ++k) {
        sd[k] = (start[k] - o[k]) * rev[k];                                     /* :293 */
        ed[k] = (end[k] - o[k]) * rev[k];                                       /* :294 */
        mi[k] = gmin(sd[k], ed[k]);                                             /* :296 */
        ma[k] = gmax(sd[k], ed[k]);                                             /* :297 */
    }
    const float tmin = gmax(gmax(mi[0], mi[1]), mi[2]);                         /* :299 */
    const float tmax = gmin(gmin(ma[0], ma[1]), ma[2]);                         /* :300 */
    if (tmax < 0.f || tmin > tmax || tmin > *best_t || obj == prev) return 0;   /* :302 */
    float pos[3];
    for (int k = 0; k < 3; ++k) {
        pos[k] = (d[k] * tmin + o[k]) - start[k];                               /* :308 */
        pos[k] = pos[k] / end[k];                                               /* :309 */
    }
    float nn[3] = {0.f, 0.f, 0.f};
    if (tmin == mi[0]) {                                                        /* :311-314 */
        nn[0] = (float)((mi[0] == sd[0]) * 2 - 1);
        uv[0] = pos[1];
        uv[1] = pos[2];
    } else if (tmin == mi[1]) {                                                 /* :317-320 */
        nn[1] = (float)((mi[1] == sd[1]) * 2 - 1);
        uv[0] = pos[0];
        uv[1] = pos[2];
    } else {                                                                    /* :323-326 */
        nn[2] = (float)((mi[2] == sd[2]) * 2 - 1);
        uv[0] = pos[0];
        uv[1] = pos[1];
    }
    if (n) memcpy(n, nn, sizeof nn);
    *best_t = tmin;                                                             /* :331 */
    return 1;
}

/* trace.glsl:30-46: the analytic loops of traceGeometry / traceOcclusion,
 * continuing from the state the triangle loop left (object ids continue
 * after the triangles: spheres, then cubes, then planes). */
void orc_analytic_scan(const orc_analytic *a, uint32_t base, const float o[3], const float d[3], uint32_t prev,
                       float *hit_t, uint32_t *obj, float *hu, float *hv, float n[3], int *nan_seen) {
    if (!a) return;
    uint32_t j = base;
    float uv[2], nn[3];
    for (uint32_t i = 0; i < a->num_spheres; ++i, ++j)
        if (orc_sphere(o, d, a->spheres + 4 * (size_t)i, j, prev, hit_t, uv, nn)) {
            *obj = j;
            *hu = uv[0];
            *hv = uv[1];
            if (n) memcpy(n, nn, sizeof nn);
            if (nan_seen && *hit_t != *hit_t) *nan_seen = 1;
        }
    for (uint32_t i = 0; i < a->num_cubes; ++i, ++j)
        if (orc_cube(o, d, a->cubes + 6 * (size_t)i, j, prev, hit_t, uv, nn)) {
            *obj = j;
            *hu = uv[0];
            *hv = uv[1];
            if (n) memcpy(n, nn, sizeof nn);
            if (nan_seen && *hit_t != *hit_t) *nan_seen = 1;
        }
    for (uint32_t i = 0; i < a->num_planes; ++i, ++j)
        if (orc_plane(o, d, a->planes + 4 * (size_t)i, j, prev, hit_t, uv, nn)) {
            *obj = j;
            *hu = uv[0];
            *hv = uv[1];
            if (n) memcpy(n, nn, sizeof nn);
            if (nan_seen && *hit_t != *hit_t) *nan_seen = 1;
        }
}
