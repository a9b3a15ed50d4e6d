/*
 * oracle.c - CPU restatement of the reference's ray-query algorithm.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): never linked into the product.
 *
 * Compiled with -ffp-contract=off and without fast-math so that every FP32
 * product and sum is rounded as written, the way the reference states them
 * in res/shaders/primitive.glsl:239-284.  x86-64 evaluates float in SSE, so
 * there is no excess precision.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define NO_HIT 3.4028235e38f /* primitive.glsl:6 */
#define MISS 0xFFFFFFFFu     /* primitive.glsl:7 */

static inline float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* primitive.glsl:239-284.  GLSL cross(x, y) = (x.y*y.z - y.y*x.z,
 * x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y); dot(x, y) = (x.x*y.x + x.y*y.y) +
 * x.z*y.z. */
int orc_mt_f32(const float o[3], const float d[3], const float p0[3], const float p1[3], const float p2[3],
               float *t, float *u, float *v) {
    const float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]}; /* :245 */
    const float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]}; /* :246 */
    const float h[3] = {d[1] * e2[2] - e2[1] * d[2], d[2] * e2[0] - e2[2] * d[0],
                        d[0] * e2[1] - e2[0] * d[1]};                /* :255 */
    const float a = (e1[0] * h[0] + e1[1] * h[1]) + e1[2] * h[2];    /* :256 */
    /* :258 `if (abs(a) < 0)` can never be true: no parallel cull. */
    const float f = 1.0f / a;                                        /* :261 */
    const float s[3] = {o[0] - p0[0], o[1] - p0[1], o[2] - p0[2]};   /* :262 */
    const float uu = f * ((s[0] * h[0] + s[1] * h[1]) + s[2] * h[2]); /* :263 */
    if (uu < 0.f || uu > 1.f) return 0;                              /* :265 */
    const float q[3] = {s[1] * e1[2] - e1[1] * s[2], s[2] * e1[0] - e1[2] * s[0],
                        s[0] * e1[1] - e1[0] * s[1]};                /* :268 */
    const float vv = f * ((d[0] * q[0] + d[1] * q[1]) + d[2] * q[2]); /* :269 */
    if (vv < 0.f || uu + vv > 1.f) return 0;                         /* :271 */
    *t = f * ((e2[0] * q[0] + e2[1] * q[1]) + e2[2] * q[2]);         /* :274 */
    *u = uu;
    *v = vv;
    return 1;
}

int orc_mt_f64(const double o[3], const double d[3], const double p0[3], const double p1[3],
               const double p2[3], double *t, double *u, double *v) {
    const double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double h[3] = {d[1] * e2[2] - e2[1] * d[2], d[2] * e2[0] - e2[2] * d[0], d[0] * e2[1] - e2[0] * d[1]};
    const double a = (e1[0] * h[0] + e1[1] * h[1]) + e1[2] * h[2];
    const double f = 1.0 / a;
    const double s[3] = {o[0] - p0[0], o[1] - p0[1], o[2] - p0[2]};
    const double uu = f * ((s[0] * h[0] + s[1] * h[1]) + s[2] * h[2]);
    if (uu < 0.0 || uu > 1.0) return 0;
    const double q[3] = {s[1] * e1[2] - e1[1] * s[2], s[2] * e1[0] - e1[2] * s[0], s[0] * e1[1] - e1[0] * s[1]};
    const double vv = f * ((d[0] * q[0] + d[1] * q[1]) + d[2] * q[2]);
    if (vv < 0.0 || uu + vv > 1.0) return 0;
    *t = f * ((e2[0] * q[0] + e2[1] * q[1]) + e2[2] * q[2]);
    *u = uu;
    *v = vv;
    return 1;
}

/* ---- a small thread pool over ray ranges ----------------------------- */

typedef void (*range_fn)(void *ctx, uint32_t lo, uint32_t hi);
typedef struct {
    range_fn fn;
    void *ctx;
    uint32_t n;
    uint32_t grain;
    volatile uint32_t next;
} pool_job;

static void *pool_worker(void *arg) {
    pool_job *j = (pool_job *)arg;
    for (;;) {
        uint32_t lo = __atomic_fetch_add(&j->next, j->grain, __ATOMIC_RELAXED);
        if (lo >= j->n) break;
        uint32_t hi = lo + j->grain < j->n ? lo + j->grain : j->n;
        j->fn(j->ctx, lo, hi);
    }
    return NULL;
}

static void run_pool(range_fn fn, void *ctx, uint32_t n, int nthreads) {
    pool_job j = {fn, ctx, n, 64, 0};
    if (nthreads <= 1 || n <= 64) {
        fn(ctx, 0, n);
        return;
    }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    int started = 0;
    for (int i = 0; i < nthreads; ++i)
        if (pthread_create(&th[i], NULL, pool_worker, &j) == 0) started++;
        else break;
    if (started == 0) {
        fn(ctx, 0, n);
        return;
    }
    for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
}

/* ---- brute force: trace.glsl:9-66 and :70-98 --------------------------- */

typedef struct {
    const float *tri;
    uint32_t ntri;
    const float *rays;
    uint32_t *hits;
    uint8_t *nan_flag;
    uint8_t *flags;
    const orc_analytic *ana;
    double guard; /* < 0: off; else skip FP32 artifacts (see orc_closest_brute_guard) */
} brute_ctx;

/* Whether an FP32-accepted pair is an artifact of FP32 rounding rather than
 * a point of the triangle: a NaN t, or a hit point o + t d (evaluated in
 * FP64 from the FP32 t) lying more than `guard` outside the triangle's
 * box. */
static int fp32_artifact(const float o[3], const float d[3], const float *p, float t, double guard) {
    if (t != t) return 1;
    double out = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double x = (double)o[k] + (double)t * (double)d[k];
        const double lo = fmin(fmin((double)p[k], (double)p[3 + k]), (double)p[6 + k]);
        const double hi = fmax(fmax((double)p[k], (double)p[3 + k]), (double)p[6 + k]);
        out = fmax(out, fmax(lo - x, x - hi));
    }
    return out > guard;
}

/* One ray scanned against every triangle in index order; returns hitT and
 * sets object/u/v exactly as traceGeometry's loop does (trace.glsl:22-28). */
static float scan_ray(const brute_ctx *c, const float *r, uint32_t *obj, float *hu, float *hv, int *nan_seen) {
    const float o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
    const uint32_t prev = f2u(r[7]);
    float hit_t = NO_HIT; /* trace.glsl:14 */
    *hu = 0.f;             /* :16 */
    *hv = 0.f;
    *obj = 0; /* :17 */
    *nan_seen = 0;
    for (uint32_t j = 0; j < c->ntri; ++j) {
        const float *p = c->tri + (size_t)j * 9;
        float t, u, v;
        if (!orc_mt_f32(o, d, p, p + 3, p + 6, &t, &u, &v)) continue;
        if (t <= 0.f || j == prev || t >= hit_t) continue; /* primitive.glsl:276 */
        if (c->guard >= 0.0 && fp32_artifact(o, d, p, t, c->guard)) continue;
        if (t != t) *nan_seen = 1;
        hit_t = t; /* :279-280 */
        *hu = u;
        *hv = v;
        *obj = j; /* trace.glsl:27 */
    }
    /* trace.glsl:30-46: spheres, cubes, planes after the triangles. */
    orc_analytic_scan(c->ana, c->ntri, o, d, prev, &hit_t, obj, hu, hv, NULL, nan_seen);
    return hit_t;
}

static void closest_range(void *ctx, uint32_t lo, uint32_t hi) {
    const brute_ctx *c = (const brute_ctx *)ctx;
    for (uint32_t i = lo; i < hi; ++i) {
        const float *r = c->rays + (size_t)i * 8;
        uint32_t obj;
        float hu, hv;
        int nan_seen;
        float hit_t = scan_ray(c, r, &obj, &hu, &hv, &nan_seen);
        /* raygen.comp:41: a miss is hitT == noHit.  A finite tmax filters
         * the closest hit afterwards (the reference has no tmax here). */
        int hit = hit_t != NO_HIT;
        const float tmax = r[3];
        if (!(tmax >= NO_HIT) && !(hit_t < tmax)) hit = 0;
        uint32_t *h = c->hits + (size_t)i * 4;
        if (hit) {
            h[0] = obj;
            h[1] = f2u(hit_t);
            h[2] = f2u(hu);
            h[3] = f2u(hv);
        } else {
            h[0] = MISS;
            h[1] = f2u(NO_HIT);
            h[2] = 0;
            h[3] = 0;
        }
        if (c->nan_flag) c->nan_flag[i] = (uint8_t)nan_seen;
    }
}

static void occluded_range(void *ctx, uint32_t lo, uint32_t hi) {
    const brute_ctx *c = (const brute_ctx *)ctx;
    for (uint32_t i = lo; i < hi; ++i) {
        const float *r = c->rays + (size_t)i * 8;
        uint32_t obj;
        float hu, hv;
        int nan_seen;
        /* traceOcclusion runs the same scan with no early exit (:78-79). */
        float hit_t = scan_ray(c, r, &obj, &hu, &hv, &nan_seen);
        c->flags[i] = hit_t < r[3] ? 1 : 0; /* :97 */
    }
}

void orc_closest_brute_ana(const float *tri, uint32_t ntri, const orc_analytic *ana, const float *rays,
                           uint32_t nrays, uint32_t *hits, uint8_t *nan_flag, int nthreads) {
    brute_ctx c = {tri, ntri, rays, hits, nan_flag, NULL, ana, -1.0};
    run_pool(closest_range, &c, nrays, nthreads);
}

void orc_occluded_brute_ana(const float *tri, uint32_t ntri, const orc_analytic *ana, const float *rays,
                            uint32_t nrays, uint8_t *flags, int nthreads) {
    brute_ctx c = {tri, ntri, rays, NULL, NULL, flags, ana, -1.0};
    run_pool(occluded_range, &c, nrays, nthreads);
}

void orc_closest_brute_guard(const float *tri, uint32_t ntri, const float *rays, uint32_t nrays, uint32_t *hits,
                             double guard, int nthreads) {
    brute_ctx c = {tri, ntri, rays, hits, NULL, NULL, NULL, guard};
    run_pool(closest_range, &c, nrays, nthreads);
}

void orc_occluded_brute_guard(const float *tri, uint32_t ntri, const float *rays, uint32_t nrays, uint8_t *flags,
                              double guard, int nthreads) {
    brute_ctx c = {tri, ntri, rays, NULL, NULL, flags, NULL, guard};
    run_pool(occluded_range, &c, nrays, nthreads);
}

void orc_closest_brute(const float *tri, uint32_t ntri, const float *rays, uint32_t nrays, uint32_t *hits,
                       uint8_t *nan_flag, int nthreads) {
    orc_closest_brute_ana(tri, ntri, NULL, rays, nrays, hits, nan_flag, nthreads);
}

void orc_occluded_brute(const float *tri, uint32_t ntri, const float *rays, uint32_t nrays, uint8_t *flags,
                        int nthreads) {
    orc_occluded_brute_ana(tri, ntri, NULL, rays, nrays, flags, nthreads);
}

/* ---- traversal of a built scene ---------------------------------------- */
/* Node formats (little-endian dwords):
 *   W=2 (64 B):  lo0.xyz hi0.xyz lo1.xyz hi1.xyz child0 child1 pad pad
 *   W=4 (64 B):  origin.xyz, ex (3 biased exponent bytes), qlo_x qlo_y qlo_z
 *                (one byte per child), qhi_x qhi_y qhi_z, child[4], pad[2]
 *   W=8 (128 B): origin.xyz, ex, qlo_x[2] qlo_y[2] qlo_z[2], qhi_*[2],
 *                child[8], pad[8]
 * Child reference: 0xFFFFFFFF empty; bit 31 clear = node index; bit 31 set =
 * leaf with count-1 in bits 28..30 and first record in bits 0..27. */

#define REF_NONE 0xFFFFFFFFu
#define KEY_MISS 0xFFFFFFFFu

typedef struct {
    float ix, iy, iz, nox, noy, noz;
    int negx, negy, negz;
} box_ray;

static inline float safe_inv(float d) { return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d)); }

static inline uint32_t child_key(float tn, float tf, int c, uint32_t slot_mask) {
    return tn <= tf ? ((f2u(tn) & ~slot_mask) | (uint32_t)c) : KEY_MISS;
}

static inline float qbyte(uint32_t w, int c) { return (float)((w >> (8 * c)) & 0xFFu); }

static void node_test(int W, const uint32_t *nd, const box_ray *b, float tmax, uint32_t *key, uint32_t *ref) {
    const uint32_t m = (uint32_t)W - 1u;
    if (W == 2) {
        for (int c = 0; c < 2; ++c) {
            const float *lo = (const float *)(nd + 6 * c);
            const float *hi = lo + 3;
            const float tnx = fmaf(b->negx ? hi[0] : lo[0], b->ix, b->nox);
            const float tfx = fmaf(b->negx ? lo[0] : hi[0], b->ix, b->nox);
            const float tny = fmaf(b->negy ? hi[1] : lo[1], b->iy, b->noy);
            const float tfy = fmaf(b->negy ? lo[1] : hi[1], b->iy, b->noy);
            const float tnz = fmaf(b->negz ? hi[2] : lo[2], b->iz, b->noz);
            const float tfz = fmaf(b->negz ? lo[2] : hi[2], b->iz, b->noz);
            const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.f));
            const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
            key[c] = child_key(tn, tf, c, m);
            ref[c] = nd[12 + c];
        }
        return;
    }
    const int nw = W / 4; /* dwords per axis per bound */
    const float ax = fmaf(u2f(nd[0]), b->ix, b->nox);
    const float ay = fmaf(u2f(nd[1]), b->iy, b->noy);
    const float az = fmaf(u2f(nd[2]), b->iz, b->noz);
    const float bx = u2f(((nd[3] >> 0) & 0xFFu) << 23) * b->ix;
    const float by = u2f(((nd[3] >> 8) & 0xFFu) << 23) * b->iy;
    const float bz = u2f(((nd[3] >> 16) & 0xFFu) << 23) * b->iz;
    const uint32_t *qlo = nd + 4, *qhi = nd + 4 + 3 * nw, *child = nd + 4 + 6 * nw;
    for (int c = 0; c < W; ++c) {
        const int w = c >> 2, s = c & 3;
        const uint32_t lx = qlo[0 * nw + w], ly = qlo[1 * nw + w], lz = qlo[2 * nw + w];
        const uint32_t hx = qhi[0 * nw + w], hy = qhi[1 * nw + w], hz = qhi[2 * nw + w];
        const float tnx = fmaf(qbyte(b->negx ? hx : lx, s), bx, ax), tfx = fmaf(qbyte(b->negx ? lx : hx, s), bx, ax);
        const float tny = fmaf(qbyte(b->negy ? hy : ly, s), by, ay), tfy = fmaf(qbyte(b->negy ? ly : hy, s), by, ay);
        const float tnz = fmaf(qbyte(b->negz ? hz : lz, s), bz, az), tfz = fmaf(qbyte(b->negz ? lz : hz, s), bz, az);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.f));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        key[c] = child_key(tn, tf, c, m);
        ref[c] = child[c];
    }
}

typedef struct {
    int width;
    const uint8_t *nodes;
    const float *tris;
    uint32_t root;
    const float *rays;
    int occlusion;
    uint32_t *hits;
    uint8_t *flags;
    uint32_t *counts;
    uint32_t node_words;
} walk_ctx;

typedef struct {
    uint32_t ref, tn;
} entry;

static void walk_range(void *ctx, uint32_t lo, uint32_t hi) {
    const walk_ctx *c = (const walk_ctx *)ctx;
    const int W = c->width;
    const uint32_t m = (uint32_t)W - 1u;
    size_t cap = 256;
    entry *stack = (entry *)malloc(cap * sizeof(entry));
    for (uint32_t i = lo; i < hi; ++i) {
        const float *r = c->rays + (size_t)i * 8;
        const float o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
        const float tmax = r[3];
        const uint32_t prev = f2u(r[7]);
        int skip = !(tmax == tmax);
        int occluded = 0;
        float tlim;
        if (c->occlusion) {
            if (!skip && tmax > NO_HIT) { /* hitT (<= FLT_MAX) < maxDist always */
                occluded = 1;
                skip = 1;
            }
            tlim = tmax;
        } else {
            tlim = tmax < NO_HIT ? tmax : NO_HIT;
        }
        box_ray b;
        b.ix = safe_inv(d[0]);
        b.iy = safe_inv(d[1]);
        b.iz = safe_inv(d[2]);
        b.nox = -(o[0] * b.ix);
        b.noy = -(o[1] * b.iy);
        b.noz = -(o[2] * b.iz);
        b.negx = b.ix < 0.f;
        b.negy = b.iy < 0.f;
        b.negz = b.iz < 0.f;
        float best_t = tlim, bu = 0.f, bv = 0.f;
        uint32_t best_id = 0;
        uint32_t nn = 0, nt = 0;
        size_t sp = 0;
        uint32_t cur = skip ? REF_NONE : c->root;
        while (cur != REF_NONE) {
            if (!(cur & 0x80000000u)) {
                ++nn;
                uint32_t key[8], ref[8];
                node_test(W, (const uint32_t *)(c->nodes + (size_t)cur * c->node_words * 4), &b, best_t, key, ref);
                /* ascending insertion sort of (key, ref); keys of hit children
                 * are distinct (their low bits are the slot) */
                for (int a = 1; a < W; ++a) {
                    uint32_t k = key[a], rr = ref[a];
                    int j = a - 1;
                    while (j >= 0 && key[j] > k) {
                        key[j + 1] = key[j];
                        ref[j + 1] = ref[j];
                        --j;
                    }
                    key[j + 1] = k;
                    ref[j + 1] = rr;
                }
                if (key[0] != KEY_MISS) {
                    cur = ref[0];
                    for (int j = W - 1; j >= 1; --j)
                        if (key[j] != KEY_MISS) {
                            if (sp == cap) {
                                cap *= 2;
                                stack = (entry *)realloc(stack, cap * sizeof(entry));
                            }
                            stack[sp].ref = ref[j];
                            stack[sp].tn = key[j] & ~m;
                            ++sp;
                        }
                    continue;
                }
            } else {
                const uint32_t first = cur & 0x0FFFFFFFu, cnt = ((cur >> 28) & 7u) + 1u;
                for (uint32_t k = 0; k < cnt; ++k) {
                    ++nt;
                    const float *rec = c->tris + (size_t)(first + k) * 12;
                    /* The record stores e1 = p1 - p0 and e2 = p2 - p0 rounded
                     * as primitive.glsl:245-246 computes them; the rest is
                     * orc_mt_f32 line by line. */
                    const float *p0 = rec, *e1 = rec + 4, *e2 = rec + 8;
                    const uint32_t id = f2u(rec[3]);
                    const float h[3] = {d[1] * e2[2] - e2[1] * d[2], d[2] * e2[0] - e2[2] * d[0],
                                        d[0] * e2[1] - e2[0] * d[1]};
                    const float a = (e1[0] * h[0] + e1[1] * h[1]) + e1[2] * h[2];
                    const float f = 1.0f / a;
                    const float s[3] = {o[0] - p0[0], o[1] - p0[1], o[2] - p0[2]};
                    const float u = f * ((s[0] * h[0] + s[1] * h[1]) + s[2] * h[2]);
                    if (u < 0.f || u > 1.f) continue;
                    const float q[3] = {s[1] * e1[2] - e1[1] * s[2], s[2] * e1[0] - e1[2] * s[0],
                                        s[0] * e1[1] - e1[0] * s[1]};
                    const float v = f * ((d[0] * q[0] + d[1] * q[1]) + d[2] * q[2]);
                    if (v < 0.f || u + v > 1.f) continue;
                    const float t = f * ((e2[0] * q[0] + e2[1] * q[1]) + e2[2] * q[2]);
                    if (c->occlusion) {
                        if (t > 0.f && id != prev && t < tlim) {
                            occluded = 1;
                            break;
                        }
                    } else if (t > 0.f && id != prev && (t < best_t || (t == best_t && id < best_id))) {
                        best_t = t;
                        best_id = id;
                        bu = u;
                        bv = v;
                    }
                }
                if (occluded) break;
            }
            cur = REF_NONE;
            while (sp > 0) {
                entry e = stack[--sp];
                if (c->occlusion || u2f(e.tn) <= best_t) {
                    cur = e.ref;
                    break;
                }
            }
        }
        if (c->occlusion) {
            c->flags[i] = (uint8_t)occluded;
        } else {
            uint32_t *h = c->hits + (size_t)i * 4;
            if (best_t < tlim) {
                h[0] = best_id;
                h[1] = f2u(best_t);
                h[2] = f2u(bu);
                h[3] = f2u(bv);
            } else {
                h[0] = MISS;
                h[1] = f2u(NO_HIT);
                h[2] = 0;
                h[3] = 0;
            }
        }
        if (c->counts) {
            c->counts[2 * (size_t)i] = nn;
            c->counts[2 * (size_t)i + 1] = nt;
        }
    }
    free(stack);
}

void orc_walk(int width, const void *nodes, const void *tris, uint32_t root, const float *rays, uint32_t nrays,
              int occlusion, uint32_t *hits, uint8_t *flags, uint32_t *counts, int nthreads) {
    walk_ctx c;
    c.width = width;
    c.nodes = (const uint8_t *)nodes;
    c.tris = (const float *)tris;
    c.root = root;
    c.rays = rays;
    c.occlusion = occlusion;
    c.hits = hits;
    c.flags = flags;
    c.counts = counts;
    c.node_words = width == 8 ? 32u : 16u;
    run_pool(walk_range, &c, nrays, nthreads);
}
