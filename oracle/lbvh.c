/*
 * lbvh.c - CPU restatement of the GPU LBVH builder's binary tree (steps 1-5
 * of igx_raytracing_amd/csrc/lbvh.hip).  TEST INFRASTRUCTURE ONLY.
 *
 * The arithmetic is the GPU's: FP32 min/max for boxes, centre = 0.5*lo +
 * 0.5*hi, quantisation (c - cmin) * (1024 / extent) clamped to [0, 1023],
 * 10-bit spreads interleaved x, y, z (x highest), a sort by (code, index),
 * the Karras (2012) split search with equal codes ordered by index, and
 * exact min/max unions bottom-up.  The sort is written as a plain comparison
 * sort, not a radix sort, so the two implementations share no code.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int32_t ord_f(float f) {
    int32_t i;
    memcpy(&i, &f, 4);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
static float unord_f(int32_t i) {
    const int32_t b = i >= 0 ? i : i ^ 0x7FFFFFFF;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static void tri_box(const float *t, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(fminf(t[k], t[3 + k]), t[6 + k]);
        hi[k] = fmaxf(fmaxf(t[k], t[3 + k]), t[6 + k]);
    }
}

static uint32_t spread10(uint32_t v) {
    /* bit b of v goes to bit 3b */
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

static uint32_t quant10(float c, float cmin, float cmax) {
    const float ext = cmax - cmin;
    const float scale = ext > 0.f ? 1024.0f / ext : 0.f;
    const float q = fminf(fmaxf((c - cmin) * scale, 0.f), 1023.f);
    return (uint32_t)q;
}

typedef struct {
    uint32_t code, idx;
} pair_t;

static int cmp_pair(const void *a, const void *b) {
    const pair_t *x = (const pair_t *)a, *y = (const pair_t *)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0);
}

static int delta(const uint32_t *codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    if (codes[i] == codes[j]) return 32 + __builtin_clz((uint32_t)(i ^ j));
    return __builtin_clz(codes[i] ^ codes[j]);
}

int orc_lbvh_binary(const float *tri, uint32_t n, uint32_t *morton, uint32_t *order, uint32_t *left,
                    uint32_t *right, float *boxes) {
    if (n == 0) return 0;
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i) {
        float lo[3], hi[3];
        tri_box(tri + (size_t)i * 9, lo, hi);
        for (int k = 0; k < 3; ++k) {
            const float c = 0.5f * lo[k] + 0.5f * hi[k];
            clo[k] = fminf(clo[k], c);
            chi[k] = fmaxf(chi[k], c);
        }
    }
    pair_t *p = (pair_t *)malloc(sizeof(pair_t) * n);
    if (!p) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        float lo[3], hi[3];
        tri_box(tri + (size_t)i * 9, lo, hi);
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) q[k] = quant10(0.5f * lo[k] + 0.5f * hi[k], clo[k], chi[k]);
        p[i].code = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
        p[i].idx = i;
    }
    qsort(p, n, sizeof(pair_t), cmp_pair);
    for (uint32_t i = 0; i < n; ++i) {
        morton[i] = p[i].code;
        order[i] = p[i].idx;
    }
    free(p);
    const int N = (int)n;
    for (int i = 0; i < N - 1; ++i) {
        const int d = delta(morton, N, i, i + 1) - delta(morton, N, i, i - 1) > 0 ? 1 : -1;
        const int dmin = delta(morton, N, i, i - d);
        int lmax = 2;
        while (delta(morton, N, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(morton, N, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(morton, N, i, j);
        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (delta(morton, N, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int g = i + s * d + (d < 0 ? -1 : 0);
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        left[i] = lo == g ? (0x80000000u | (uint32_t)g) : (uint32_t)g;
        right[i] = hi == g + 1 ? (0x80000000u | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
    }
    /* leaf boxes, then internal boxes in post-order from the root */
    for (uint32_t j = 0; j < n; ++j) {
        float lo[3], hi[3];
        tri_box(tri + (size_t)order[j] * 9, lo, hi);
        float *b = boxes + ((size_t)(n - 1) + j) * 6;
        memcpy(b, lo, sizeof lo);
        memcpy(b + 3, hi, sizeof hi);
    }
    if (n == 1) return 0;
    uint32_t *st = (uint32_t *)malloc(sizeof(uint32_t) * 2 * n);
    uint8_t *seen = (uint8_t *)calloc(n, 1);
    if (!st || !seen) {
        free(st);
        free(seen);
        return -1;
    }
    size_t sp = 0;
    st[sp++] = 0;
    while (sp) {
        const uint32_t node = st[sp - 1];
        const uint32_t c[2] = {left[node], right[node]};
        if (!seen[node]) {
            seen[node] = 1;
            for (int s = 0; s < 2; ++s)
                if (!(c[s] & 0x80000000u)) st[sp++] = c[s];
            continue;
        }
        --sp;
        /* Unions in the order-preserving integer image of the floats (-0 < +0,
         * NaNs placed by their bits), the order the device's atomic min/max
         * folds use. */
        int32_t u[6] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
        for (int s = 0; s < 2; ++s) {
            const size_t b = (c[s] & 0x80000000u) ? (size_t)(n - 1) + (c[s] & 0x7FFFFFFFu) : c[s];
            for (int k = 0; k < 3; ++k) {
                const int32_t l = ord_f(boxes[b * 6 + k]), h = ord_f(boxes[b * 6 + 3 + k]);
                u[k] = l < u[k] ? l : u[k];
                u[3 + k] = h > u[3 + k] ? h : u[3 + k];
            }
        }
        for (int k = 0; k < 6; ++k) boxes[(size_t)node * 6 + k] = unord_f(u[k]);
    }
    free(st);
    free(seen);
    return 0;
}
