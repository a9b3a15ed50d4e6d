// Device-side traversal shared by the query kernels (trace.hip), the caller
// adapters (callers.hip) and the wavefront path tracer (wavefront.hip):
// the Moller-Trumbore test in the reference's operation order, slab tests of
// BVH2 / quantised BVH4 / BVH8 nodes, the LDS-ring + global-spill stack, and
// the per-lane tracer.  Every translation unit that includes it is built
// with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "analytic.hpp"
#include "kernels.hpp"

namespace rtx {
namespace dev {

constexpr float kNoHit = 3.4028235e38f;  // primitive.glsl:6
constexpr uint32_t kKeyMiss = 0xFFFFFFFFu;

struct RayD {
    float ox, oy, oz, dx, dy, dz;
};

// Moller-Trumbore in the reference's operation order, primitive.glsl:239-284.
// GLSL cross(x,y) = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y);
// dot(x,y) = (x.x*y.x + x.y*y.y) + x.z*y.z.  Returns false when the u/v tests
// reject (:265, :271); t is then unset.  A = (p0, id), B = e1, C = e2.
__device__ __forceinline__ bool mt_test(const RayD &r, const float4 &A, const float4 &B,
                                        const float4 &C, float &t, float &u, float &v) {
    const float hx = r.dy * C.z - C.y * r.dz;  // h = cross(dir, p2 - p0)   :255
    const float hy = r.dz * C.x - C.z * r.dx;
    const float hz = r.dx * C.y - C.x * r.dy;
    const float a = (B.x * hx + B.y * hy) + B.z * hz;  // a = dot(p1 - p0, h)  :256
    const float f = 1.0f / a;                          // :261 (no epsilon, :258 is dead)
    const float sx = r.ox - A.x, sy = r.oy - A.y, sz = r.oz - A.z;  // s = pos - p0  :262
    u = f * ((sx * hx + sy * hy) + sz * hz);           // :263
    if (u < 0.f || u > 1.f) return false;              // :265 (NaN passes)
    const float qx = sy * B.z - B.y * sz;              // q = cross(s, p1 - p0)  :268
    const float qy = sz * B.x - B.z * sx;
    const float qz = sx * B.y - B.x * sy;
    v = f * ((r.dx * qx + r.dy * qy) + r.dz * qz);     // :269
    if (v < 0.f || u + v > 1.f) return false;          // :271
    t = f * ((C.x * qx + C.y * qy) + C.z * qz);        // :274
    return true;
}

// Per-ray constants of the slab test.
struct BoxRay {
    float ix, iy, iz;     // 1/dir, zero components replaced by +-1e-20
    float nox, noy, noz;  // -org * inv
    bool negx, negy, negz;
};

__device__ __forceinline__ float safe_inv(float d) {
    return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d));
}

__device__ __forceinline__ BoxRay make_box_ray(const RayD &r) {
    BoxRay b;
    b.ix = safe_inv(r.dx);
    b.iy = safe_inv(r.dy);
    b.iz = safe_inv(r.dz);
    b.nox = -(r.ox * b.ix);
    b.noy = -(r.oy * b.iy);
    b.noz = -(r.oz * b.iz);
    b.negx = b.ix < 0.f;
    b.negy = b.iy < 0.f;
    b.negz = b.iz < 0.f;
    return b;
}

__device__ __forceinline__ float4 as_float4(uint4 v) {
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}

// Scene, ray and hit memory is accessed in the global address space.  A
// pointer read from memory (the batch kernel's per-query scene and ray
// pointers) is generic to the compiler, which then emits flat loads; those
// count against the same wait counter as LDS operations, so every stack
// access would also wait for the node fetches in flight.
#define RTX_GLOBAL __attribute__((address_space(1)))
typedef uint32_t rtx_u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t rtx_u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint4 ldg4(const void *p) {
    const rtx_u32x4 v = *(const RTX_GLOBAL rtx_u32x4 *)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 ldg4f(const void *p) { return as_float4(ldg4(p)); }
__device__ __forceinline__ void stg4(void *p, float4 v) {
    *(RTX_GLOBAL rtx_u32x4 *)p =
        rtx_u32x4{__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w)};
}
__device__ __forceinline__ void stg1(uint8_t *p, uint8_t v) { *(RTX_GLOBAL uint8_t *)p = v; }

// An empty asm statement that reads the registers: the values count as used
// at this point on every path through the surrounding code.
__device__ __forceinline__ void keep(const uint4 &v) { asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w)); }
__device__ __forceinline__ void keep(const float4 &v) { asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w)); }

__device__ __forceinline__ float qbyte(uint32_t w, int c) {
    return static_cast<float>((w >> (8 * c)) & 0xFFu);
}

__device__ __forceinline__ float exp_scale(uint32_t w, int axis) {
    return __uint_as_float(((w >> (8 * axis)) & 0xFFu) << 23);
}

// Sort key of a child: tnear's bits with the low log2(W) bits replaced by the
// slot, so equal distances order by slot; kKeyMiss for a culled child.
template <int W>
__device__ __forceinline__ uint32_t child_key(float tn, float tf, int c) {
    constexpr uint32_t m = W - 1;
    return tn <= tf ? ((__float_as_uint(tn) & ~m) | static_cast<uint32_t>(c)) : kKeyMiss;
}

template <int W>
struct NodeTest;

// Node words arrive already loaded (w[k] = the node's k-th 16-byte word).

// NodeB2: [lo0 hi0 lo1 hi1] floats, child[2].
template <>
struct NodeTest<2> {
    __device__ static __forceinline__ void run(const uint4 *w, const BoxRay &b, float tmax, uint32_t key[2],
                                               uint32_t ref[2]) {
        const uint4 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
        const float lo[2][3] = {{__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)},
                                {__uint_as_float(w1.z), __uint_as_float(w1.w), __uint_as_float(w2.x)}};
        const float hi[2][3] = {{__uint_as_float(w0.w), __uint_as_float(w1.x), __uint_as_float(w1.y)},
                                {__uint_as_float(w2.y), __uint_as_float(w2.z), __uint_as_float(w2.w)}};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const float tnx = fmaf(b.negx ? hi[c][0] : lo[c][0], b.ix, b.nox);
            const float tfx = fmaf(b.negx ? lo[c][0] : hi[c][0], b.ix, b.nox);
            const float tny = fmaf(b.negy ? hi[c][1] : lo[c][1], b.iy, b.noy);
            const float tfy = fmaf(b.negy ? lo[c][1] : hi[c][1], b.iy, b.noy);
            const float tnz = fmaf(b.negz ? hi[c][2] : lo[c][2], b.iz, b.noz);
            const float tfz = fmaf(b.negz ? lo[c][2] : hi[c][2], b.iz, b.noz);
            const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.f));
            const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
            key[c] = child_key<2>(tn, tf, c);
        }
        ref[0] = w3.x;
        ref[1] = w3.y;
    }
};

// Quantised child boxes: t = fma(q, scale*inv, fma(origin, inv, -org*inv)).
template <int W>
__device__ __forceinline__ void quant_test(const BoxRay &b, float ox, float oy, float oz, uint32_t ex,
                                           const uint32_t *qlo_x, const uint32_t *qlo_y,
                                           const uint32_t *qlo_z, const uint32_t *qhi_x,
                                           const uint32_t *qhi_y, const uint32_t *qhi_z, float tmax,
                                           uint32_t key[W]) {
    const float ax = fmaf(ox, b.ix, b.nox), ay = fmaf(oy, b.iy, b.noy), az = fmaf(oz, b.iz, b.noz);
    const float bx = exp_scale(ex, 0) * b.ix, by = exp_scale(ex, 1) * b.iy, bz = exp_scale(ex, 2) * b.iz;
#pragma unroll
    for (int c = 0; c < W; ++c) {
        const int w = c >> 2, s = c & 3;
        const uint32_t nqx = b.negx ? qhi_x[w] : qlo_x[w], fqx = b.negx ? qlo_x[w] : qhi_x[w];
        const uint32_t nqy = b.negy ? qhi_y[w] : qlo_y[w], fqy = b.negy ? qlo_y[w] : qhi_y[w];
        const uint32_t nqz = b.negz ? qhi_z[w] : qlo_z[w], fqz = b.negz ? qlo_z[w] : qhi_z[w];
        const float tnx = fmaf(qbyte(nqx, s), bx, ax), tfx = fmaf(qbyte(fqx, s), bx, ax);
        const float tny = fmaf(qbyte(nqy, s), by, ay), tfy = fmaf(qbyte(fqy, s), by, ay);
        const float tnz = fmaf(qbyte(nqz, s), bz, az), tfz = fmaf(qbyte(fqz, s), bz, az);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.f));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        key[c] = child_key<W>(tn, tf, c);
    }
}

// NodeB4 dwords: 0-2 origin, 3 ex, 4-6 qlo[x|y|z], 7-9 qhi[x|y|z], 10-13 child.
template <>
struct NodeTest<4> {
    __device__ static __forceinline__ void run(const uint4 *w, const BoxRay &b, float tmax, uint32_t key[4],
                                               uint32_t ref[4]) {
        const uint4 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
        const uint32_t qlx[1] = {w1.x}, qly[1] = {w1.y}, qlz[1] = {w1.z};
        const uint32_t qhx[1] = {w1.w}, qhy[1] = {w2.x}, qhz[1] = {w2.y};
        quant_test<4>(b, __uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z), w0.w, qlx,
                      qly, qlz, qhx, qhy, qhz, tmax, key);
        ref[0] = w2.z;
        ref[1] = w2.w;
        ref[2] = w3.x;
        ref[3] = w3.y;
    }
};

// NodeB8 dwords: 0-2 origin, 3 ex, 4-9 qlo[x0 x1 y0 y1 z0 z1], 10-15 qhi, 16-23 child
// (the second 64-byte half carries nothing else the test reads).
template <>
struct NodeTest<8> {
    __device__ static __forceinline__ void run(const uint4 *w, const BoxRay &b, float tmax, uint32_t key[8],
                                               uint32_t ref[8]) {
        const uint4 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4], w5 = w[5];
        const uint32_t qlx[2] = {w1.x, w1.y}, qly[2] = {w1.z, w1.w}, qlz[2] = {w2.x, w2.y};
        const uint32_t qhx[2] = {w2.z, w2.w}, qhy[2] = {w3.x, w3.y}, qhz[2] = {w3.z, w3.w};
        quant_test<8>(b, __uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z), w0.w, qlx,
                      qly, qlz, qhx, qhy, qhz, tmax, key);
        ref[0] = w4.x;
        ref[1] = w4.y;
        ref[2] = w4.z;
        ref[3] = w4.w;
        ref[4] = w5.x;
        ref[5] = w5.y;
        ref[6] = w5.z;
        ref[7] = w5.w;
    }
};

// 16-byte words a node fetch reads, and the node stride in 16-byte words.
template <int W>
struct NodeWords {
    static constexpr int kRead = W == 8 ? 6 : 4;
    static constexpr int kStride = W == 8 ? 8 : 4;
};

// Word k of a node.  BVH2 and BVH4 nodes use only the first 8 bytes of
// their last word (the last two child references), so that word is read
// with an 8-byte load: a 16-byte load whose upper half nothing reads lets
// the compiler reuse those registers while the load is still in flight,
// and the wait it then has to insert holds back the rest of the step's
// fetches.
template <int W>
__device__ __forceinline__ uint4 node_word(const uint4 *np, int k) {
    if (W <= 4 && k == 3) {
        const rtx_u32x2 v = *(const RTX_GLOBAL rtx_u32x2 *)(np + k);
        return make_uint4(v.x, v.y, 0u, 0u);
    }
    return ldg4(np + k);
}

// The root node's words, loaded once per wave into wave-uniform registers.
// Every ray of a single-scene launch begins at the root, so a ray can test
// it from these when it begins and spend its first step on the next level:
// one dependent node fetch fewer per ray.
template <int W>
struct RootWords {
    uint4 w[NodeWords<W>::kRead];
    bool ok;  // the root is an interior node

    __device__ __forceinline__ void load(const void *nodes, uint32_t root) {
        ok = root != kRefNone && !ref_is_leaf(root);
        if (!ok) return;
        const uint4 *np = reinterpret_cast<const uint4 *>(nodes) + static_cast<size_t>(root) * NodeWords<W>::kStride;
#pragma unroll
        for (int k = 0; k < NodeWords<W>::kRead; ++k) {
            const uint4 v = node_word<W>(np, k);
            w[k] = make_uint4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                              __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
        }
    }
};

// Child references of a node from its words (the layouts of NodeTest).
template <int W>
__device__ __forceinline__ uint32_t node_child(const uint4 *w, int c) {
    if (W == 2) return c == 0 ? w[3].x : w[3].y;
    if (W == 4) {
        const uint32_t r[4] = {w[2].z, w[2].w, w[3].x, w[3].y};
        return r[c];
    }
    const uint32_t r[8] = {w[4].x, w[4].y, w[4].z, w[4].w, w[5].x, w[5].y, w[5].z, w[5].w};
    return r[c];
}

// The root's interior children, copied once per block into LDS: with the
// root in registers, a ray that begins tests two levels without a fetch.
template <int W>
struct TopNodes {
    uint4 w[W][NodeWords<W>::kRead];
    uint32_t ref[W];  // kRefNone where the child is a leaf or empty

    // Call with the whole block (contains a barrier).
    __device__ __forceinline__ void load(const void *nodes, const RootWords<W> &rw) {
        constexpr int NR = NodeWords<W>::kRead;
        for (int k = threadIdx.x; k < W * NR; k += blockDim.x) {
            const int c = k / NR, j = k % NR;
            const uint32_t r = rw.ok ? node_child<W>(rw.w, c) : kRefNone;
            const bool in = r != kRefNone && !ref_is_leaf(r);
            if (j == 0) ref[c] = in ? r : kRefNone;
            if (in) w[c][j] = ldg4(reinterpret_cast<const uint4 *>(nodes) + static_cast<size_t>(r) * NodeWords<W>::kStride + j);
        }
        __syncthreads();
    }
};

__device__ __forceinline__ void cswap(uint32_t *k, uint32_t *r, int i, int j) {
    const bool s = k[j] < k[i];
    const uint32_t ki = k[i], kj = k[j], ri = r[i], rj = r[j];
    k[i] = s ? kj : ki;
    k[j] = s ? ki : kj;
    r[i] = s ? rj : ri;
    r[j] = s ? ri : rj;
}

// Ascending sort of (key, ref) pairs by key; fixed networks.
template <int W>
__device__ __forceinline__ void sort_children(uint32_t *k, uint32_t *r);
template <>
__device__ __forceinline__ void sort_children<2>(uint32_t *k, uint32_t *r) {
    cswap(k, r, 0, 1);
}
template <>
__device__ __forceinline__ void sort_children<4>(uint32_t *k, uint32_t *r) {
    cswap(k, r, 0, 1);
    cswap(k, r, 2, 3);
    cswap(k, r, 0, 2);
    cswap(k, r, 1, 3);
    cswap(k, r, 1, 2);
}
template <>
__device__ __forceinline__ void sort_children<8>(uint32_t *k, uint32_t *r) {
    // Batcher odd-even merge sort, 19 comparators.
    cswap(k, r, 0, 1); cswap(k, r, 2, 3); cswap(k, r, 4, 5); cswap(k, r, 6, 7);
    cswap(k, r, 0, 2); cswap(k, r, 1, 3); cswap(k, r, 4, 6); cswap(k, r, 5, 7);
    cswap(k, r, 1, 2); cswap(k, r, 5, 6);
    cswap(k, r, 0, 4); cswap(k, r, 1, 5); cswap(k, r, 2, 6); cswap(k, r, 3, 7);
    cswap(k, r, 2, 4); cswap(k, r, 3, 5);
    cswap(k, r, 1, 2); cswap(k, r, 3, 4); cswap(k, r, 5, 6);
}

// Stack entry: (child ref, tnear bits) when pops are culled by distance
// (closest hit), the child ref alone when they are not (any hit).
template <bool TN>
struct StackEntry;
template <>
struct StackEntry<true> {
    using type = uint2;
    __device__ static __forceinline__ type make(uint32_t ref, uint32_t tn) { return make_uint2(ref, tn); }
    __device__ static __forceinline__ uint32_t ref(type e) { return e.x; }
    __device__ static __forceinline__ uint32_t tn(type e) { return e.y; }
};
template <>
struct StackEntry<false> {
    using type = uint32_t;
    __device__ static __forceinline__ type make(uint32_t ref, uint32_t) { return ref; }
    __device__ static __forceinline__ uint32_t ref(type e) { return e; }
    __device__ static __forceinline__ uint32_t tn(type) { return 0u; }
};

// Per-lane traversal stack: a ring of S entries in LDS holding the newest
// entries, spilling the oldest to a per-thread global stack when full, so no
// entry is ever dropped.
template <int S, int BLOCK, bool TN>
struct LaneStack {
    using E = StackEntry<TN>;
    using T = typename E::type;
    static constexpr int kS = S;
    static constexpr int kBlock = BLOCK;
    T *lds;          // &lds_base[tid], stride BLOCK
    T *glob;         // &gstack[gtid], stride gstride
    uint32_t gstride, gdepth;
    int top, lc, gc;

    __device__ __forceinline__ void reset() {
        top = S - 1;
        lc = 0;
        gc = 0;
    }
    __device__ __forceinline__ bool push(uint32_t ref, uint32_t tn) {
        if (lc == S) {
            if (static_cast<uint32_t>(gc) >= gdepth) return false;
            const int oldest = (top + 1) & (S - 1);
            glob[static_cast<size_t>(gc) * gstride] = lds[oldest * BLOCK];
            ++gc;
            --lc;
        }
        top = (top + 1) & (S - 1);
        lds[top * BLOCK] = E::make(ref, tn);
        ++lc;
        return true;
    }
    // Pushes the hit children key[F..W-1] of a node (sorted ascending, culled
    // ones last with key kKeyMiss) far-to-near, so the nearest ends on top:
    // the same ring contents as pushing them one by one, without a branch
    // per child.  Needs lc + W - F <= S: every slot above top is then free,
    // so the culled children are written there too and simply left above
    // the new top.
    template <int W, int F>
    __device__ __forceinline__ void push_sorted(const uint32_t *key, const uint32_t *ref, uint32_t slot_mask) {
        int np = 0;
#pragma unroll
        for (int j = F; j < W; ++j) np += key[j] != 0xFFFFFFFFu ? 1 : 0;
#pragma unroll
        for (int j = F; j < W; ++j) {
            const int k = j - F + 1;  // 1-based rank among the pushed
            const int pos = (top + (k <= np ? np + 1 - k : k)) & (S - 1);
            lds[pos * BLOCK] = E::make(ref[j], key[j] & ~slot_mask);
        }
        top = (top + np) & (S - 1);
        lc += np;
    }

    // The newest entry when it is in the LDS ring, without removing it.  An
    // LDS read only: a peek that could also read the global spill compiles
    // to one generic (flat) load, which counts as a vector memory access, so
    // the wave would wait for every global load in flight (the ray records
    // just requested) before issuing its next fetch.
    __device__ __forceinline__ bool peek_ring(T &e) const {
        if (lc > 0) {
            e = lds[top * BLOCK];
            return true;
        }
        return false;
    }

    // Removes the newest entry, which peek_ring returned.
    __device__ __forceinline__ void drop_ring() {
        top = (top - 1) & (S - 1);
        --lc;
    }

    __device__ __forceinline__ bool pop(T &e) {
        if (lc > 0) {
            e = lds[top * BLOCK];
            top = (top - 1) & (S - 1);
            --lc;
            return true;
        }
        if (gc > 0) {
            --gc;
            e = glob[static_cast<size_t>(gc) * gstride];
            return true;
        }
        return false;
    }
};

// A ray record as two 16-byte words: (org, tmax) and (dir, prev_id).
struct RayRec {
    float4 a, b;
};

__device__ __forceinline__ RayRec load_ray(const rt_ray *ray) {
    const float4 *p = reinterpret_cast<const float4 *>(ray);
    return RayRec{ldg4f(p), ldg4f(p + 1)};
}

// Rank of the calling lane among the set lanes of m below it.
__device__ __forceinline__ uint32_t lane_rank(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Ray indices of a persistent dynamic-schedule launch.  The first `local`
// of the n indices are dealt to the waves in chunks, round-robin: wave w
// takes chunks w, w + W, w + 2W, ... (W waves in the grid) without any
// shared counter, so every wave samples the whole index space, and a batch
// whose rays differ in cost by query (scene size, query kind) still gives
// each wave a similar mix.  The remaining indices form a pool that waves
// draw from, one global atomic per chunk, once their own chunks are used up.
// Single-address global atomics serialise across the whole device, so only
// the pool, which evens out what the dealing did not, uses one.
struct WaveFeed {
    uint32_t next = 0, end = 0;  // the wave's current range [next, end)
    uint32_t local = 0, n = 0;
    uint32_t deal = 0, stride = 0;  // the wave's next local chunk, chunk stride
    bool exhausted = false;

    __device__ __forceinline__ void init(uint32_t n_, uint32_t pool_shift, uint32_t wave, uint32_t total_waves) {
        n = n_;
        local = n - (pool_shift < 32u ? (n >> pool_shift) : 0u);
        deal = wave;
        stride = total_waves;
        next = end = 0;
    }

    // Refills the range when it is used up: the wave's next local chunk,
    // else a chunk of the pool.  Wave-uniform.
    __device__ __forceinline__ void reserve(uint32_t *work, uint32_t chunk) {
        if (next < end || exhausted) return;
        const uint64_t lo = static_cast<uint64_t>(deal) * chunk;
        if (lo < local) {
            next = static_cast<uint32_t>(lo);
            end = local - next > chunk ? next + chunk : local;
            deal += stride;
            return;
        }
        uint32_t b = 0;
        if ((threadIdx.x & 63u) == 0u) b = atomicAdd(work, chunk);
        b = __builtin_amdgcn_readfirstlane(b);
        const uint64_t g = static_cast<uint64_t>(local) + b;
        if (g >= n) {
            exhausted = true;
        } else {
            next = static_cast<uint32_t>(g);
            end = n - next > chunk ? next + chunk : n;
        }
    }
    // Index for the idle lane of rank `rank` among the `nneed` idle lanes,
    // or ~0u; then advances past the indices handed out.  Wave-uniform call.
    __device__ __forceinline__ uint32_t take(uint32_t rank, uint32_t nneed) {
        const uint32_t i = rank < end - next ? next + rank : ~0u;
        next = end - next > nneed ? next + nneed : end;
        return i;
    }
};

// The same feed with each wave's dealt rays handed out longest first.
//
// The launch ends when the last wave's last ray does, and a wave's last
// rays are the ones it started last.  So each wave takes its dealt rays in
// windows of up to kWin, estimates every ray's cost from the length of its
// path through the scene's box (Spearman 0.77 with its node count on 1M
// incoherent rays), and hands the window out in four classes of that
// length, longest first, through a list in LDS.  Each ray record is read
// once more for this.  Which rays a wave gets is unchanged; only their order
// within the wave is, so results are identical.  Measured on 1M incoherent
// rays over 1M triangles: the same order made on the host ran 4.6% faster.
// The pool is handed out as before.
struct RankedFeed {
    static constexpr uint32_t kWin = 512;  // slots per window: 8 per lane
    static constexpr uint32_t kPer = kWin / 64;
    uint32_t *list = nullptr;              // the wave's kWin LDS entries
    uint32_t wpos = 0, wlen = 0;           // [wpos, wlen) of the list not yet handed out
    uint64_t slot = 0;                     // dealt rays of the wave already windowed
    uint32_t next = 0, end = 0;            // pool range
    uint32_t local = 0, n = 0, deal = 0, stride = 0;
    bool windows_done = false, exhausted = false;
    const rt_ray *rays = nullptr;
    float lo[3], hi[3], scale = 0.f;

    __device__ __forceinline__ void init(uint32_t n_, uint32_t pool_shift, uint32_t wave, uint32_t total_waves,
                                         uint32_t *lds_list, const rt_ray *rays_, const float *lo_, const float *hi_) {
        n = n_;
        local = n - (pool_shift < 32u ? (n >> pool_shift) : 0u);
        deal = wave;
        stride = total_waves;
        list = lds_list;
        rays = rays_;
        float d2 = 0.f;
        for (int k = 0; k < 3; ++k) {
            lo[k] = lo_[k];
            hi[k] = hi_[k];
            d2 += (hi_[k] - lo_[k]) * (hi_[k] - lo_[k]);
        }
        // Classes at 1/6, 2/6 and 3/6 of the box diagonal.
        scale = d2 > 0.f ? 6.f / sqrtf(d2) : 0.f;
    }

    // Path length class (0-3) of a ray record through the scene's box.
    __device__ __forceinline__ uint32_t length_class(const RayRec &r) const {
        const float o[3] = {r.a.x, r.a.y, r.a.z}, d[3] = {r.b.x, r.b.y, r.b.z};
        float tn = 0.f, tf = r.a.w < kNoHit ? r.a.w : kNoHit;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float inv = __builtin_amdgcn_rcpf(d[k]);
            const float t0 = (lo[k] - o[k]) * inv, t1 = (hi[k] - o[k]) * inv;
            tn = fmaxf(tn, fminf(t0, t1));
            tf = fminf(tf, fmaxf(t0, t1));
        }
        const float len = (tf - tn) * sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * scale;
        return len >= 3.f ? 3u : (len >= 2.f ? 2u : (len >= 1.f ? 1u : 0u));  // NaN: class 0
    }

    // Fills the list with the wave's next kWin dealt rays, longest class
    // first.  Wave-uniform call.
    __device__ __forceinline__ void fill(uint32_t chunk) {
        const uint32_t lane = __lane_id();
        uint32_t idx[kPer], cls[kPer];
        bool ok[kPer];
#pragma unroll
        for (int k = 0; k < static_cast<int>(kPer); ++k) {
            const uint64_t sl = slot + static_cast<uint64_t>(k) * 64u + lane;
            const uint64_t c = deal + (sl / chunk) * stride;
            const uint64_t i = c * chunk + sl % chunk;
            ok[k] = i < local;
            idx[k] = ok[k] ? static_cast<uint32_t>(i) : 0u;
        }
#pragma unroll
        for (int k = 0; k < static_cast<int>(kPer); ++k) cls[k] = ok[k] ? length_class(load_ray(rays + idx[k])) : 0u;
        uint32_t base = 0;
#pragma unroll
        for (int c = 3; c >= 0; --c) {
#pragma unroll
            for (int k = 0; k < static_cast<int>(kPer); ++k) {
                const bool in = ok[k] && cls[k] == static_cast<uint32_t>(c);
                const unsigned long long m = __ballot(in);
                if (in) list[base + lane_rank(m)] = idx[k];
                base += static_cast<uint32_t>(__popcll(m));
            }
        }
        __builtin_amdgcn_wave_barrier();
        slot += kWin;
        wpos = 0;
        wlen = base;
        // Dealt slots are in increasing index order, so a window that is not
        // full was the last.
        if (base < kWin) windows_done = true;
    }

    __device__ __forceinline__ void reserve(uint32_t *work, uint32_t chunk) {
        if (wpos < wlen || next < end || exhausted) return;
        if (!windows_done) {
            fill(chunk);
            if (wpos < wlen) return;
        }
        uint32_t b = 0;
        if ((threadIdx.x & 63u) == 0u) b = atomicAdd(work, chunk);
        b = __builtin_amdgcn_readfirstlane(b);
        const uint64_t g = static_cast<uint64_t>(local) + b;
        if (g >= n) {
            exhausted = true;
        } else {
            next = static_cast<uint32_t>(g);
            end = n - next > chunk ? next + chunk : n;
        }
    }

    __device__ __forceinline__ uint32_t take(uint32_t rank, uint32_t nneed) {
        if (wpos < wlen) {
            const uint32_t i = rank < wlen - wpos ? list[wpos + rank] : ~0u;
            wpos = wlen - wpos > nneed ? wpos + nneed : wlen;
            return i;
        }
        const uint32_t i = rank < end - next ? next + rank : ~0u;
        next = end - next > nneed ? next + nneed : end;
        return i;
    }
};

// Forces a callback lambda to be inlined: a lambda left out of line makes
// the tracer state it captures live in scratch memory.
#define RTX_INLINE __attribute__((always_inline))

// The last wave of a launch zeroes the pool counter (work[0]) and the wave
// counter (work[1]) for the next launch on the stream.
__device__ __forceinline__ void release_work(uint32_t *work, uint32_t total_waves) {
    if ((threadIdx.x & 63u) == 0u) {
        __threadfence();
        const uint32_t done = atomicAdd(work + 1, 1u);
        if (done == total_waves - 1u) {
            atomicExch(work, 0u);
            atomicExch(work + 1, 0u);
        }
    }
}

// Per-block LDS slots through which work-stealing helpers hand their
// results to the lane that owns the ray (one slot per lane): the minimum
// (t bits << 32 | id) key over the helpers' closest hits, or 0 once one of
// them found occlusion; the u, v and triangle record of that key; and the
// count of helpers still at work, with bit 31 set when one overflowed its
// stack.  LDS atomics make a merge cost the same however many helpers finish
// in one step.
template <int BLOCK>
struct MergeSlots {
    unsigned long long key[BLOCK];
    uint4 uvr[BLOCK];
    uint32_t pend[BLOCK];
};

// Per-lane traversal state of one ray.  Visiting order: children of a node
// by ascending (tnear, slot); the nearest is taken next, the others pushed
// far-to-near; a popped entry whose tnear exceeds the current closest t is
// skipped.  Closest hit keeps the (t, id)-lexicographic minimum, which equals
// the scan's first-index-wins result for any visiting order.
// KIND: 0 closest hit, 1 any hit, 2 chosen per lane at begin (batches that
// mix both query kinds in one launch).
template <int W, int KIND, bool COUNT, int S, int BLOCK, bool ONE_FETCH = (KIND == 0)>
struct Tracer {
    static constexpr bool kTN = KIND != 1;
    using Stack = LaneStack<S, BLOCK, kTN>;
    using Entry = typename Stack::T;

    RayD rd;
    BoxRay br;
    float tlim, best_t, bu, bv;
    uint32_t best_id, best_rec, prev, cur, nn, nt;
    uint32_t owner;  // helper: the lane (of this wave) whose ray it traverses
    bool occluded, overflow, any_lane, helping;
    Stack st;
    // Stack size (LDS ring + spill) up to which a node step may take a
    // second node; -1 never (see dual_stack in kernels.hpp).
    int dual_max;
    // The wave's merge slots in LDS, indexed by lane (see MergeSlots); null
    // in kernels without work stealing.
    unsigned long long *mkey;
    uint4 *muvr;
    uint32_t *mpend;

    __device__ __forceinline__ bool any_hit() const { return KIND == 2 ? any_lane : KIND == 1; }

    __device__ __forceinline__ void begin(const RayRec &r, uint32_t root, bool lane_any = false) {
        begin_ray(RayD{r.a.x, r.a.y, r.a.z, r.b.x, r.b.y, r.b.z}, r.a.w, __float_as_uint(r.b.w), root, lane_any);
    }

    __device__ __forceinline__ void begin_ray(const RayD &ray, float tmax, uint32_t prev_id, uint32_t root,
                                              bool lane_any = false) {
        any_lane = lane_any;
        const bool ANY = any_hit();
        rd = ray;
        prev = prev_id;
        bool skip = !(tmax == tmax);  // NaN tmax: no candidate
        occluded = false;
        overflow = false;
        if (ANY) {
            // trace.glsl:97 returns hit.hitT < maxDist with hitT <= FLT_MAX,
            // so maxDist > FLT_MAX is always occluded.
            if (!skip && tmax > kNoHit) {
                occluded = true;
                skip = true;
            }
            tlim = tmax;
        } else {
            // A NaN tmax keeps tlim NaN, so no candidate (analytic ones
            // included) can pass `t < tlim` and the ray reports a miss.
            tlim = skip ? tmax : (tmax < kNoHit ? tmax : kNoHit);
        }
        br = make_box_ray(rd);
        best_t = tlim;
        bu = 0.f;
        bv = 0.f;
        best_id = 0u;  // with best_t = tlim no tie can be accepted before a hit
        best_rec = 0u;
        nn = 0;
        nt = 0;
        helping = false;
        st.reset();
        cur = skip ? kRefNone : root;
    }

    // Stealing kernels: opens this lane's merge slot for the ray it begins
    // (no helper of an earlier ray is left: a ray finishes only when its
    // helpers have).
    __device__ __forceinline__ void claim_slot() {
        const uint32_t l = __lane_id();
        mkey[l] = ~0ull;
        mpend[l] = 0u;
    }

    // Helpers still at work on this lane's ray.
    __device__ __forceinline__ uint32_t helpers_left() const { return mpend[__lane_id()] & 0x7FFFFFFFu; }

    // Before the owner finishes: takes in what its helpers left in the slot.
    // The slot holds the minimum (t bits << 32 | id) key over the helpers'
    // results (t > 0, so the bits order like the values, and equal t orders
    // by id: the (t, id)-lexicographic minimum), u, v and the record of that
    // key, the overflow bit, and 0 as the key once a helper found occlusion.
    __device__ __forceinline__ void collect() {
        const uint32_t l = __lane_id();
        const unsigned long long m = mkey[l];
        if (mpend[l] >> 31) overflow = true;
        if (any_hit()) {
            if (m == 0ull) occluded = true;
        } else {
            const unsigned long long own =
                (static_cast<unsigned long long>(__float_as_uint(best_t)) << 32) | best_id;
            if (m < own) {
                const uint4 q = muvr[l];
                best_t = __uint_as_float(static_cast<uint32_t>(m >> 32));
                best_id = static_cast<uint32_t>(m);
                bu = __uint_as_float(q.x);
                bv = __uint_as_float(q.y);
                best_rec = q.z;
            }
        }
    }

    __device__ __forceinline__ uint32_t pop() {
        Entry e;
        const bool ANY = any_hit();
        while (st.pop(e))
            if (ANY || __uint_as_float(Stack::E::tn(e)) <= best_t) return Stack::E::ref(e);
        return kRefNone;
    }

    // Pushes every hit child of a node far-to-near.
    __device__ __forceinline__ void push_all(const uint32_t *key, const uint32_t *ref) {
        constexpr uint32_t kSlotMask = W - 1;
        if (W <= 4 && st.lc + W <= S) {
            st.template push_sorted<W, 0>(key, ref, kSlotMask);
        } else {
#pragma unroll
            for (int j = W - 1; j >= 0; --j)
                if (key[j] != kKeyMiss && !st.push(ref[j], key[j] & ~kSlotMask)) overflow = true;
        }
    }

    // cur is an interior node.  A closest-hit walk (BVH2/BVH4, not the
    // counting one) also takes up to kExtra entries from the top of the
    // stack, while they are interior nodes the ray still has to visit, and
    // tests them in the same step: a lane then has several node fetches of
    // one ray in flight instead of a chain of single dependent ones.  That
    // shortens the critical path of long rays, which set the end of a launch
    // (the last ~half of a 1M-ray launch drains a few long rays).  Their hit
    // children go on the stack below those of cur.  The visiting order
    // changes, the result does not: the closest hit is the
    // (t, id)-lexicographic minimum over every accepted triangle, in any
    // order.  Any-hit queries stop at the first hit and lose more to the
    // extra fetches than they gain (-11% measured with one extra node, -26%
    // re-measured later), so they test one node a step; the any-hit lanes of
    // a mixed batch take the extra node (+1.5% on the 64-scene batch), whose
    // closest-hit lanes set the pace anyway.
    // One extra node: two or three measured slower (more VALU and divergence
    // than the latency they hide).
    static constexpr int kExtra = (!COUNT && W <= 4 && KIND != 1) ? 1 : 0;
    static constexpr int kKX = kExtra > 0 ? kExtra : 1;
    static constexpr int kNR = NodeWords<W>::kRead;

    // Takes the extra nodes of a two-node step off the stack top: interior
    // entries of the LDS ring the ray still has to visit, while the stack is
    // small enough (dual_max) for the step's pushes to stay within it.  An
    // entry in the global spill is left for a later pop.
    __device__ __forceinline__ void take_extra(uint32_t (&curx)[kKX]) {
#pragma unroll
        for (int x = 0; x < kKX; ++x) curx[x] = kRefNone;
        if (kExtra > 0 && (KIND == 2 || !any_hit()) && st.lc + st.gc <= dual_max) {
            bool more = true;
#pragma unroll
            for (int x = 0; x < kKX; ++x) {
                Entry e;
                if (more && st.peek_ring(e)) {
                    const uint32_t r = Stack::E::ref(e);
                    if (!ref_is_leaf(r) && __uint_as_float(Stack::E::tn(e)) <= best_t) {
                        st.drop_ring();
                        curx[x] = r;
                    } else {
                        more = false;
                    }
                } else {
                    more = false;
                }
            }
        }
    }

    // The extra nodes' words.  A missing extra node re-reads cur's node (the
    // same cache lines), so every fetch issues back to back with no branch.
    __device__ __forceinline__ void fetch_extra(const void *nodes, const uint32_t (&curx)[kKX],
                                                uint4 (&wx)[kKX][kNR]) const {
        if (kExtra > 0) {
#pragma unroll
            for (int x = 0; x < kKX; ++x) {
                const uint4 *npx = reinterpret_cast<const uint4 *>(nodes) +
                                   static_cast<size_t>(curx[x] != kRefNone ? curx[x] : cur) * NodeWords<W>::kStride;
#pragma unroll
                for (int k = 0; k < kNR; ++k) wx[x][k] = node_word<W>(npx, k);
            }
        }
    }

    // cur is an interior node whose words (w) and extra nodes (curx, wx) are
    // loaded: tests them, pushes the hit children, and moves cur to the
    // nearest child of cur's node or to the next stack entry.
    __device__ __forceinline__ void node_visit(const uint4 (&w)[kNR], const uint32_t (&curx)[kKX],
                                               const uint4 (&wx)[kKX][kNR]) {
        constexpr uint32_t kSlotMask = W - 1;
        if (COUNT) ++nn;
        uint32_t key[W], ref[W];
        if (kExtra > 0) {
#pragma unroll
            for (int x = kKX - 1; x >= 0; --x) {
                if (curx[x] != kRefNone) {
                    NodeTest<W>::run(wx[x], br, best_t, key, ref);
                    sort_children<W>(key, ref);
                    push_all(key, ref);
                }
            }
        }
        NodeTest<W>::run(w, br, best_t, key, ref);
        sort_children<W>(key, ref);
        if (key[0] != kKeyMiss) {
            cur = ref[0];
            if (W <= 4 && st.lc + (W - 1) <= S) {
                st.template push_sorted<W, 1>(key, ref, kSlotMask);
            } else {
#pragma unroll
                for (int j = W - 1; j >= 1; --j)
                    if (key[j] != kKeyMiss && !st.push(ref[j], key[j] & ~kSlotMask)) overflow = true;
            }
        } else {
            cur = pop();
        }
        if (overflow) cur = kRefNone;
    }

    // cur is an interior node.  A closest-hit walk (BVH2/BVH4, not the
    // counting one) also takes up to kExtra entries from the top of the
    // stack, while they are interior nodes the ray still has to visit, and
    // tests them in the same step: a lane then has several node fetches of
    // one ray in flight instead of a chain of single dependent ones.  That
    // shortens the critical path of long rays, which set the end of a launch
    // (the last ~half of a 1M-ray launch drains a few long rays).  Their hit
    // children go on the stack below those of cur.  The visiting order
    // changes, the result does not: the closest hit is the
    // (t, id)-lexicographic minimum over every accepted triangle, in any
    // order.  Any-hit walks stop at the first hit and lose more to the extra
    // fetches than they gain (-11% measured with one extra node), so they
    // test one node a step.
    // One extra node: two or three measured slower (more VALU and divergence
    // than the latency they hide).
    __device__ __forceinline__ void node_step(const void *nodes) {
        uint32_t curx[kKX];
        take_extra(curx);
        const uint4 *np = reinterpret_cast<const uint4 *>(nodes) + static_cast<size_t>(cur) * NodeWords<W>::kStride;
        uint4 w[kNR], wx[kKX][kNR];
#pragma unroll
        for (int k = 0; k < kNR; ++k) w[k] = node_word<W>(np, k);
        fetch_extra(nodes, curx, wx);
        node_visit(w, curx, wx);
    }

    // begin, then the root's visit from the wave's copy of its words: the
    // ray's first step without its fetch.  The visit sequence is unchanged.
    __device__ __forceinline__ void begin_at_root(const RayRec &r, uint32_t root, const RootWords<W> &rw,
                                                  bool lane_any = false) {
        begin(r, root, lane_any);
        if (rw.ok && cur == root) {
            uint32_t curx[kKX];
            uint4 wx[kKX][kNR];
#pragma unroll
            for (int x = 0; x < kKX; ++x) curx[x] = kRefNone;
            node_visit(rw.w, curx, wx);
        }
    }

    // begin_at_root, then the next level from the block's LDS copy when the
    // ray's nearest child of the root is one of the root's interior
    // children.
    __device__ __forceinline__ void begin_at_top(const RayRec &r, uint32_t root, const RootWords<W> &rw,
                                                 const TopNodes<W> &top, bool lane_any = false) {
        begin_at_root(r, root, rw, lane_any);
        if (!rw.ok || cur == kRefNone || ref_is_leaf(cur)) return;
        int c = -1;
#pragma unroll
        for (int k = 0; k < W; ++k) c = top.ref[k] == cur ? k : c;
        if (c < 0) return;
        uint4 w[kNR], wx[kKX][kNR];
        uint32_t curx[kKX];
#pragma unroll
        for (int x = 0; x < kKX; ++x) curx[x] = kRefNone;
#pragma unroll
        for (int j = 0; j < kNR; ++j) w[j] = top.w[c][j];
        node_visit(w, curx, wx);
    }

    // cur is a leaf whose first triangle record (A, B, C) is loaded: tests
    // its triangles in order and moves cur to the next stack entry.
    __device__ __forceinline__ void leaf_visit(const float4 *tris, float4 A, float4 B, float4 C) {
        const bool ANY = any_hit();
        const uint32_t first = leaf_first(cur), cnt = leaf_count(cur);
        for (uint32_t k = 0; k < cnt; ++k) {
            if (COUNT) ++nt;
            if (k > 0) {
                const float4 *tp = tris + static_cast<size_t>(first + k) * 3;
                A = ldg4f(tp);
                B = ldg4f(tp + 1);
                C = ldg4f(tp + 2);
            }
            float t, u, v;
            if (!mt_test(rd, A, B, C, t, u, v)) continue;
            const uint32_t id = __float_as_uint(A.w);
            if (ANY) {
                // trace.glsl:97 with the primitive.glsl:276 rejects.
                if (t > 0.f && id != prev && t < tlim) {
                    occluded = true;
                    break;
                }
            } else if (t > 0.f && id != prev && (t < best_t || (t == best_t && id < best_id))) {
                best_t = t;
                best_id = id;
                best_rec = first + k;
                bu = u;
                bv = v;
            }
        }
        cur = (ANY && occluded) ? kRefNone : pop();
    }

    // cur is a leaf.
    __device__ __forceinline__ void leaf_step(const float4 *tris) {
        const float4 *tp = tris + static_cast<size_t>(leaf_first(cur)) * 3;
        leaf_visit(tris, ldg4f(tp), ldg4f(tp + 1), ldg4f(tp + 2));
    }

    // Steps.
    //
    // ONE_FETCH (closest-hit query kernels): every lane issues its one fetch
    // first, a node's words or a leaf's first triangle record, and only then
    // tests what it fetched, so a wave waits on one memory latency per step
    // even when some lanes are at leaves and others at interior nodes.  A
    // leaf that a node visit makes cur is tested in the next step.  The
    // sequence of visits of each ray is the same as visiting the leaf at
    // once; only its timing moves.  Measured on 1M incoherent rays, which
    // mostly miss: +3.3% (BVH4), +4.3% (BVH8).
    //
    // Otherwise such a leaf is visited in the same step, after the node (a
    // second fetch the wave waits on).  Rays that mostly end soon after their
    // first leaf (any-hit walks, diffuse path-tracing bounces) lose 2-8% to
    // the extra step; mixed-kind batches gain 3% from one fetch per step.
    //
    // A step is split into fetch() and visit() so that the dynamic loop can
    // issue the fetch before it requests new ray records: loads return in
    // order, so the records then arrive behind the node words instead of
    // delaying them.
    static constexpr bool kOneFetch = ONE_FETCH;

    // The loads of a one-fetch step, between fetch() and visit().
    struct Fetched {
        uint4 w[kNR];
        uint4 wx[kKX][kNR];
        uint32_t curx[kKX];
    };

    // First half of a step: issues its loads (one-fetch walks only; a
    // two-fetch walk does its whole step in visit()).
    __device__ __forceinline__ void fetch(const void *nodes, const float4 *tris, Fetched &f) {
        if (!kOneFetch || cur == kRefNone) return;
        const bool leaf = ref_is_leaf(cur);
        if (leaf) {
#pragma unroll
            for (int x = 0; x < kKX; ++x) f.curx[x] = kRefNone;
        } else {
            take_extra(f.curx);
        }
        // A leaf's record is 3 words, a node's 4 or 6: the first three words
        // of either come from one address.
        const uint4 *p = leaf ? reinterpret_cast<const uint4 *>(tris + static_cast<size_t>(leaf_first(cur)) * 3)
                              : reinterpret_cast<const uint4 *>(nodes) +
                                    static_cast<size_t>(cur) * NodeWords<W>::kStride;
#pragma unroll
        for (int k = 0; k < 3; ++k) f.w[k] = ldg4(p + k);
        if (!leaf) {
#pragma unroll
            for (int k = 3; k < kNR; ++k) f.w[k] = node_word<W>(p, k);
            fetch_extra(nodes, f.curx, f.wx);
        }
    }

    // Second half: tests what fetch() loaded.  False when the ray is done.
    __device__ __forceinline__ bool visit(const void *nodes, const float4 *tris, const Fetched &f) {
        if (!kOneFetch) {
            if (cur != kRefNone && !ref_is_leaf(cur)) node_step(nodes);
            if (cur != kRefNone && ref_is_leaf(cur)) leaf_step(tris);
            return cur != kRefNone;
        }
        if (cur == kRefNone) return false;
        if (ref_is_leaf(cur))
            leaf_visit(tris, as_float4(f.w[0]), as_float4(f.w[1]), as_float4(f.w[2]));
        else
            node_visit(f.w, f.curx, f.wx);
        return cur != kRefNone;
    }

    // Marks a one-fetch step's words as used, on every path, once visit()
    // has run.  visit() reads them under per-lane conditions that the
    // compiler cannot relate to the fetch, so it would otherwise treat those
    // loads as possibly still in flight at the next step, and make that
    // step's fetch first wait for every older memory operation: the ray
    // records the refill just requested and the hit records just stored.
    __device__ static __forceinline__ void settle(const Fetched &f) {
        if (!kOneFetch) return;
#pragma unroll
        for (int k = 0; k < kNR; ++k) keep_word(f.w[k], k);
#pragma unroll
        for (int x = 0; x < kKX; ++x)
#pragma unroll
            for (int k = 0; k < kNR; ++k) keep_word(f.wx[x][k], k);
    }
    // keep() of the part of node word k that node_word() loads.
    __device__ static __forceinline__ void keep_word(const uint4 &v, int k) {
        if (W <= 4 && k == 3)
            asm volatile("" ::"v"(v.x), "v"(v.y));
        else
            keep(v);
    }

    // One step, false when the ray is done.  The tail loop calls it without
    // settle(): the step's fetch there has no newer record loads to wait
    // for, and settling measured 1% slower.
    __device__ __forceinline__ bool advance(const void *nodes, const float4 *tris) {
        Fetched f;
        fetch(nodes, tris, f);
        return visit(nodes, tris, f);
    }

    // Traverses until the ray is finished.
    __device__ __forceinline__ void run(const void *nodes, const float4 *tris) {
        while (cur != kRefNone) {
            while (cur != kRefNone && !ref_is_leaf(cur)) node_step(nodes);
            while (cur != kRefNone && ref_is_leaf(cur)) leaf_step(tris);
        }
    }

    // The scene's analytic primitives after the triangle walk: closest hit
    // continues the reference's scan from the walk's result; any hit runs
    // the scan when no triangle occludes (trace.glsl:97 then holds iff an
    // analytic primitive's t is below tmax).
    __device__ __forceinline__ void analytic(const AnaDev &ana, uint32_t base) {
        const AnaRay r{rd.ox, rd.oy, rd.oz, rd.dx, rd.dy, rd.dz};
        if (any_hit()) {
            if (occluded) return;
            float t = kNoHit, u, v;
            uint32_t id = 0;
            analytic_scan(ana, base, r, prev, t, id, u, v);
            occluded = t < tlim;
        } else {
            analytic_scan(ana, base, r, prev, best_t, best_id, bu, bv);
        }
    }

    __device__ __forceinline__ void store(rt_hit *hits, uint8_t *flags, uint32_t i, uint32_t *err) const {
        if (overflow) atomicOr(err, 1u);
        if (!any_hit()) {
            const bool hit = best_t < tlim;
            stg4(hits + i, hit ? make_float4(__uint_as_float(best_id), best_t, bu, bv)
                               : make_float4(__uint_as_float(RT_MISS), kNoHit, 0.f, 0.f));
        } else if (flags) {
            stg1(flags + i, occluded ? 1 : 0);
        }
    }

    __device__ __forceinline__ void finish(const TraceLaunch &a, uint32_t i) const {
        store(a.hits, a.flags, i, a.err);
        if (COUNT) {
            a.counts[2 * static_cast<size_t>(i)] = nn;
            a.counts[2 * static_cast<size_t>(i) + 1] = nt;
        }
    }
};

template <int W, int KIND, bool COUNT, int S, int BLOCK, bool OF>
__device__ __forceinline__ void init_stack(Tracer<W, KIND, COUNT, S, BLOCK, OF> &t,
                                           typename Tracer<W, KIND, COUNT, S, BLOCK, OF>::Entry *lds_stack,
                                           uint2 *gstack, uint32_t gdepth, MergeSlots<BLOCK> *slots = nullptr,
                                           int dual_max = -1) {
    using Entry = typename Tracer<W, KIND, COUNT, S, BLOCK, OF>::Entry;
    t.st.lds = lds_stack + threadIdx.x;
    t.st.gstride = gridDim.x * BLOCK;
    t.st.glob = reinterpret_cast<Entry *>(gstack) + (blockIdx.x * BLOCK + threadIdx.x);
    t.st.gdepth = gdepth;
    t.dual_max = dual_max;
    t.cur = kRefNone;
    t.helping = false;
    t.occluded = false;
    const uint32_t wave_base = threadIdx.x & ~63u;
    t.mkey = slots ? slots->key + wave_base : nullptr;
    t.muvr = slots ? slots->uvr + wave_base : nullptr;
    t.mpend = slots ? slots->pend + wave_base : nullptr;
}

// ---- work stealing within a wave ---------------------------------------------
//
// Once a wave has no rays left to hand out, its idle lanes help the lanes
// still traversing: an idle lane (thief) takes the oldest entry of a busy
// lane's LDS ring, which is the largest subtree that lane still has to
// visit, and traverses it with a copy of that lane's ray.  When the thief's
// subtree is done it hands its best hit back to the owner, which keeps the
// (t, id)-lexicographic minimum, the same as if it had visited the subtree
// itself, so results do not change.  An owner finishes only when its own
// walk and all its helpers are done.  Everything happens in lockstep within
// the wave: pairing by ballots and ranks, ray state by cross-lane reads,
// results by a scalar loop over the finishing helpers.  Helpers donate too,
// so a large subtree handed to one helper is split again as lanes free up.

// Position of the r-th (0-based) set bit of m; r < popcount(m).
__device__ __forceinline__ uint32_t nth_set_bit(unsigned long long m, uint32_t r) {
    uint32_t pos = 0;
    uint32_t c = __popc(static_cast<uint32_t>(m));
    if (r >= c) { r -= c; m >>= 32; pos += 32; }
    c = __popc(static_cast<uint32_t>(m) & 0xFFFFu);
    if (r >= c) { r -= c; m >>= 16; pos += 16; }
    c = __popc(static_cast<uint32_t>(m) & 0xFFu);
    if (r >= c) { r -= c; m >>= 8; pos += 8; }
    c = __popc(static_cast<uint32_t>(m) & 0xFu);
    if (r >= c) { r -= c; m >>= 4; pos += 4; }
    c = __popc(static_cast<uint32_t>(m) & 0x3u);
    if (r >= c) { r -= c; m >>= 2; pos += 2; }
    if (r >= (static_cast<uint32_t>(m) & 1u)) pos += 1;
    return pos;
}

__device__ __forceinline__ float readlane_f(float v, uint32_t j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), static_cast<int>(j)));
}
__device__ __forceinline__ uint32_t readlane_u(uint32_t v, uint32_t j) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(j)));
}

__device__ __forceinline__ float shfl_f(float v, uint32_t src) { return __shfl(v, static_cast<int>(src)); }
__device__ __forceinline__ uint32_t shfl_u(uint32_t v, uint32_t src) { return __shfl(v, static_cast<int>(src)); }

// Pairs thieves with donors (lanes traversing a ray, their own or as a
// helper, with at least one LDS ring entry).  Returns true in a thief that
// took work; the lane that owns the ray is in t.owner, and that lane's
// helper count has been raised.  `extra` lets a kernel copy per-lane context
// of the donor (scene pointers of a batch) for the thief:
// extra(donor_lane, is_thief).
template <typename TR, typename Extra>
__device__ __forceinline__ bool steal(TR &t, bool thief, bool donor, Extra &&extra) {
    const unsigned long long tm = __ballot(thief);
    if (tm == 0ull) return false;
    const unsigned long long dm = __ballot(donor);
    if (dm == 0ull) return false;
    const uint32_t np = min(static_cast<uint32_t>(__popcll(tm)), static_cast<uint32_t>(__popcll(dm)));
    const uint32_t lane = __lane_id();
    const bool take = thief && lane_rank(tm) < np;
    const bool give = donor && lane_rank(dm) < np;
    const uint32_t d = take ? nth_set_bit(dm, lane_rank(tm)) : lane;
    // Donor state, read by every lane (only thieves use it).
    const int dtop = __shfl(t.st.top, static_cast<int>(d));
    const int dlc = __shfl(t.st.lc, static_cast<int>(d));
    const float ox = shfl_f(t.rd.ox, d), oy = shfl_f(t.rd.oy, d), oz = shfl_f(t.rd.oz, d);
    const float dx = shfl_f(t.rd.dx, d), dy = shfl_f(t.rd.dy, d), dz = shfl_f(t.rd.dz, d);
    const float ix = shfl_f(t.br.ix, d), iy = shfl_f(t.br.iy, d), iz = shfl_f(t.br.iz, d);
    const float nox = shfl_f(t.br.nox, d), noy = shfl_f(t.br.noy, d), noz = shfl_f(t.br.noz, d);
    const float tlim = shfl_f(t.tlim, d), best_t = shfl_f(t.best_t, d), bu = shfl_f(t.bu, d), bv = shfl_f(t.bv, d);
    const uint32_t prev = shfl_u(t.prev, d), best_id = shfl_u(t.best_id, d), best_rec = shfl_u(t.best_rec, d);
    const uint32_t dany = shfl_u(t.any_lane ? 1u : 0u, d);
    const uint32_t downer = shfl_u(t.helping ? t.owner : lane, d);
    extra(d, take);
    typename TR::Entry e{};
    if (take) {
        // The donor's oldest LDS entry, read before the donor gives it up.
        const int oldest = (dtop - dlc + 1) & (TR::Stack::kS - 1);
        e = t.st.lds[oldest * TR::Stack::kBlock + (static_cast<int>(d) - static_cast<int>(lane))];
    }
    __builtin_amdgcn_wave_barrier();
    if (give) --t.st.lc;
    if (take) {
        t.helping = true;
        t.owner = downer;
        t.rd = RayD{ox, oy, oz, dx, dy, dz};
        t.br.ix = ix;
        t.br.iy = iy;
        t.br.iz = iz;
        t.br.nox = nox;
        t.br.noy = noy;
        t.br.noz = noz;
        t.br.negx = ix < 0.f;
        t.br.negy = iy < 0.f;
        t.br.negz = iz < 0.f;
        t.tlim = tlim;
        t.prev = prev;
        t.best_t = best_t;
        t.best_id = best_id;
        t.best_rec = best_rec;
        t.bu = bu;
        t.bv = bv;
        t.any_lane = dany != 0u;
        t.occluded = false;
        t.overflow = false;
        t.st.reset();
        const bool keep = t.any_hit() || __uint_as_float(TR::Stack::E::tn(e)) <= best_t;
        t.cur = keep ? TR::Stack::E::ref(e) : kRefNone;
    }
    // Each new helper counts in the slot of the lane that owns its ray.
    if (take) atomicAdd(&t.mpend[t.owner], 1u);
    return take;
}

// Helpers whose subtree is done hand their result to their owners' slots and
// become idle; any-hit walks whose ray is already occluded stop.  Call with
// all lanes of the wave.  The wave's LDS operations complete in program
// order, so each read below sees every lane's atomic above it.
template <typename TR>
__device__ __forceinline__ void merge_helpers(TR &t) {
    const bool done = t.helping && t.cur == kRefNone;
    const uint32_t o = t.owner;
    const unsigned long long key =
        t.any_hit() ? (t.occluded ? 0ull : ~0ull)
                    : (static_cast<unsigned long long>(__float_as_uint(t.best_t)) << 32) | t.best_id;
    if (done) atomicMin(&t.mkey[o], key);
    if (done && !t.any_hit() && t.mkey[o] == key)
        t.muvr[o] = make_uint4(__float_as_uint(t.bu), __float_as_uint(t.bv), t.best_rec, 0u);
    if (done) {
        if (t.overflow) atomicOr(&t.mpend[o], 0x80000000u);
        atomicSub(&t.mpend[o], 1u);
        t.helping = false;
    }
    // An occluded ray needs no more of its walk, by its owner or a helper.
    if (t.any_hit() && t.cur != kRefNone && t.mkey[t.helping ? o : __lane_id()] == 0ull) {
        t.occluded = true;
        t.cur = kRefNone;
    }
}


// ---- the dynamic-schedule loop ---------------------------------------------
//
// Shared by the query, batch and path-tracing kernels.  Main phase: while
// the wave's feed has rays, every lane traces its own ray and idle lanes
// take new ones (refill).  Tail phase, once the feed is exhausted: no
// refill; idle lanes steal subtrees of the rays still being traced.  The two
// phases are separate loops so that the main loop carries none of the
// stealing code.
//
// A main-phase iteration issues the step's fetch, visits what it fetched,
// starts the rays whose records the previous iteration requested, and only
// then refills.  The record loads a refill issues are thus the newest loads
// in flight when the wave next waits for a step's fetch, which was issued
// before them, and a record arrives during the next step's fetch instead of
// holding that fetch back (+2.5% closest hit on 1M incoherent rays, +0.6-1%
// on the 64-scene batch, results identical).  A lane that finishes in a
// step is idle at that iteration's refill.
//
// Right after the visit, before any store, the step's words (settle) and
// the record requested for a pending lane (held) are marked as used on
// every path.  The compiler tracks loads in flight per register, not per
// lane: a record or word that a skipped branch did not read would count as
// still in flight when its registers are next loaded, and the wait it then
// inserts (memory operations complete in order) would also wait for the
// newer ones: the hit records just stored or the ray records just
// requested.  A wave would then wait twice per step.
//
// Callbacks:
//   assign(k)       lane takes ray index k: load its record (consumed by
//                   begin() in the next iteration)
//   begin()         start the ray whose record assign() loaded
//   fetch(f)        first half of a step of the lane's ray: issue its loads
//   visit(f)        second half: test what fetch() loaded; false when the
//                   ray is done
//   advance()       a whole step (fetch, then visit), for the tail
//   finish()        the lane's own ray is done: store or shade it
//   extra(d, take)  steal: copy per-lane context of donor lane d if take
//   held()          marks the record assign() loaded as used (keep())
template <typename TR, typename Feed, typename Assign, typename Begin, typename Fetch, typename Visit,
          typename Advance, typename Finish, typename Extra, typename Held>
__device__ __forceinline__ void dynamic_loop(TR &t, Feed &feed, uint32_t *work, uint32_t chunk, uint32_t refill,
                                             Assign &&assign, Begin &&begin, Fetch &&fetch, Visit &&visit,
                                             Advance &&advance, Finish &&finish, Extra &&extra, Held &&held) {
    bool alive = false, pending = false, retired = false;
    for (;;) {
        typename TR::Fetched f;
        if (alive) fetch(f);
        bool done = false;
        if (alive) done = !visit(f);
        TR::settle(f);
        held();
        if (done) {
            finish();
            alive = false;
        }
        if (pending) {
            begin();
            t.claim_slot();
            pending = false;
            alive = true;
        }
        const bool idle = !alive && !retired;
        const unsigned long long need = __ballot(idle);
        const uint32_t nneed = static_cast<uint32_t>(__popcll(need));
        if (nneed != 0u && (nneed >= refill || __ballot(alive) == 0ull)) {
            feed.reserve(work, chunk);
            const uint32_t k = feed.take(lane_rank(need), nneed);
            if (idle) {
                if (k != ~0u) {
                    assign(k);
                    pending = true;
                } else if (feed.exhausted) {
                    retired = true;
                }
            }
        }
        // A reserve that found the feed exhausted handed out no index, so no
        // lane is pending here.
        if (feed.exhausted) break;
    }
    // Tail: the remaining rays, with idle lanes stealing their subtrees,
    // pairing thieves with donors every step.  (Pairing every second step
    // had measured 1.3% faster for closest hit before the load-ordering and
    // priority measures here; with them, every step is 0.7% faster.)
    //
    // A wave in its tail takes the highest issue priority.  The waves of a
    // CU otherwise issue oldest first, so the tail of a younger wave, which
    // the launch's end waits for, yields to main-phase steps of older waves
    // that have work enough to hide the delay (+0.5% closest hit).
    __builtin_amdgcn_s_setprio(3);
    retired = !alive && !pending;
    for (;;) {
        if (__ballot(alive || pending || t.helping) == 0ull) break;
        steal(t, retired && !t.helping, (alive || t.helping) && t.cur != kRefNone && t.st.lc >= 1, extra);
        if ((alive || t.helping) && t.cur != kRefNone) advance();
        merge_helpers(t);
        if (alive && t.cur == kRefNone && t.helpers_left() == 0u) {
            t.collect();
            finish();
            alive = false;
            retired = true;
        }
        if (pending) {
            begin();
            t.claim_slot();
            pending = false;
            alive = true;
        }
    }
}

}  // namespace dev
}  // namespace rtx
