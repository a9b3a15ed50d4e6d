// GPU LBVH builder for scenes that arrive at run time.
//
// The reference has no acceleration structure (its traversal scans every
// triangle, res/shaders/trace.glsl:25-27); query results do not depend on the
// tree, so this builder trades tree quality for build speed:
//   1. scene bounds (vertex box, centroid box): block min/max reduction
//   2. 30-bit Morton code of each triangle's box centre in the centroid box
//   3. stable LSD radix sort of (code, index): 8-bit digits, per-tile LDS
//      histograms, one digit-major exclusive scan, per-tile stable ranking
//      with wave64 ballots
//   4. Karras (2012) radix tree over the sorted codes, equal codes ordered
//      by index
//   5. bottom-up boxes (exact min/max, so arrival order does not matter)
//   6. a subtree of <= max_leaf triangles becomes a leaf when the host SAH
//      builder's leaf test accepts it, else it is split further; boxes
//      padded outward as the host builder pads them; nodes of width 2/4/8
//      emitted level by level (children of a node contiguous, numbering by
//      prefix sums, so the layout is deterministic)
// Steps 1-5 are restated on the CPU by the oracle; tests compare the binary
// tree bit for bit, and the emitted leaves against the leaf test applied to
// the oracle's tree.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "alloc.hpp"
#include "lbvh.hpp"
#include "rt_format.h"

namespace rtx {

namespace {

constexpr uint32_t kRadixTile = 2048;  // keys per sort tile (256 threads x 8)
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;

// ---- 1. bounds -------------------------------------------------------------

// bounds[0..5]: vertex lo.xyz hi.xyz; bounds[6..11]: centroid lo.xyz hi.xyz.
__device__ __forceinline__ void tri_box(const float *t, float lo[3], float hi[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(fminf(t[k], t[3 + k]), t[6 + k]);
        hi[k] = fmaxf(fmaxf(t[k], t[3 + k]), t[6 + k]);
    }
}

__device__ __forceinline__ float centre(float lo, float hi) { return 0.5f * lo + 0.5f * hi; }

__global__ void __launch_bounds__(256) k_bounds_partial(const float *__restrict__ tri, uint32_t n,
                                                        float *__restrict__ partial) {
    __shared__ float red[12][256];
    float acc[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        acc[k] = INFINITY;
        acc[3 + k] = -INFINITY;
        acc[6 + k] = INFINITY;
        acc[9 + k] = -INFINITY;
    }
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float lo[3], hi[3];
        tri_box(tri + static_cast<size_t>(i) * 9, lo, hi);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            acc[k] = fminf(acc[k], lo[k]);
            acc[3 + k] = fmaxf(acc[3 + k], hi[k]);
            const float c = centre(lo[k], hi[k]);
            acc[6 + k] = fminf(acc[6 + k], c);
            acc[9 + k] = fmaxf(acc[9 + k], c);
        }
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                const float a = red[k][threadIdx.x], b = red[k][threadIdx.x + s];
                red[k][threadIdx.x] = ((k / 3) & 1) ? fmaxf(a, b) : fminf(a, b);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 12) partial[blockIdx.x * 12 + threadIdx.x] = red[threadIdx.x][0];
}

// One block of 256 threads: thread t folds partials t, t + 256, ...; then a
// tree over the threads in LDS.
__global__ void __launch_bounds__(256) k_bounds_final(const float *__restrict__ partial, uint32_t nblocks,
                                                      float *__restrict__ bounds) {
    __shared__ float red[12][256];
    float acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = ((k / 3) & 1) ? -INFINITY : INFINITY;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const float v = partial[b * 12 + k];
            acc[k] = ((k / 3) & 1) ? fmaxf(acc[k], v) : fminf(acc[k], v);
        }
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                const float o = red[k][threadIdx.x + off];
                red[k][threadIdx.x] = ((k / 3) & 1) ? fmaxf(red[k][threadIdx.x], o) : fminf(red[k][threadIdx.x], o);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 12) bounds[threadIdx.x] = red[threadIdx.x][0];
}

// ---- 2. Morton codes ---------------------------------------------------------

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Quantised position of a centre: (c - cmin) * (1024 / extent), clamped to
// [0, 1023]; a flat axis quantises to 0.
__device__ __forceinline__ uint32_t quant10(float c, float cmin, float cmax) {
    const float ext = cmax - cmin;
    const float scale = ext > 0.f ? 1024.0f / ext : 0.f;
    const float q = fminf(fmaxf((c - cmin) * scale, 0.f), 1023.f);
    return static_cast<uint32_t>(q);
}

__global__ void __launch_bounds__(256) k_morton(const float *__restrict__ tri, uint32_t n,
                                                const float *__restrict__ bounds, uint32_t *__restrict__ codes,
                                                uint32_t *__restrict__ idx) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float lo[3], hi[3];
    tri_box(tri + static_cast<size_t>(i) * 9, lo, hi);
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) q[k] = quant10(centre(lo[k], hi[k]), bounds[6 + k], bounds[9 + k]);
    codes[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    idx[i] = i;
}

// ---- 3. stable radix sort ---------------------------------------------------

__global__ void __launch_bounds__(256) k_radix_hist(const uint32_t *__restrict__ keys, uint32_t n, uint32_t shift,
                                                    uint32_t ntiles, uint32_t *__restrict__ ghist) {
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRadixTile;
    for (uint32_t r = 0; r < kRadixTile / 256; ++r) {
        const uint32_t e = base + r * 256 + threadIdx.x;
        if (e < n) atomicAdd(&hist[(keys[e] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    ghist[static_cast<size_t>(threadIdx.x) * ntiles + blockIdx.x] = hist[threadIdx.x];
}

// In-place exclusive scan of m words with one block.  Thread t owns a
// contiguous run of a multiple of 4 words, read and written as 16-byte
// vectors four at a time so that each thread keeps several loads in flight.
__global__ void __launch_bounds__(kScanThreads) k_scan_single(uint32_t *__restrict__ a, uint32_t m,
                                                               uint32_t *__restrict__ total) {
    __shared__ uint32_t part[kScanThreads];
    const uint32_t per = ((m + kScanThreads - 1) / kScanThreads + 3u) & ~3u;
    const uint32_t lo = min(threadIdx.x * per, m);
    const uint32_t hi = min(lo + per, m);
    const uint32_t vhi = lo + ((hi - lo) & ~15u);  // end of the 4-vector groups
    uint32_t s = 0;
    for (uint32_t i = lo; i < vhi; i += 16) {
        const uint4 *p = reinterpret_cast<const uint4 *>(a + i);
        const uint4 v0 = p[0], v1 = p[1], v2 = p[2], v3 = p[3];
        s += (v0.x + v0.y) + (v0.z + v0.w) + (v1.x + v1.y) + (v1.z + v1.w) + (v2.x + v2.y) + (v2.z + v2.w) +
             (v3.x + v3.y) + (v3.z + v3.w);
    }
    for (uint32_t i = vhi; i < hi; ++i) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over the per-thread sums.
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (uint32_t i = lo; i < vhi; i += 16) {
        uint4 *p = reinterpret_cast<uint4 *>(a + i);
        uint4 v[4] = {p[0], p[1], p[2], p[3]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 x = v[q];
            v[q].x = run;
            v[q].y = run + x.x;
            v[q].z = v[q].y + x.y;
            v[q].w = v[q].z + x.z;
            run = v[q].w + x.w;
        }
        p[0] = v[0];
        p[1] = v[1];
        p[2] = v[2];
        p[3] = v[3];
    }
    for (uint32_t i = vhi; i < hi; ++i) {
        const uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
    if (total && threadIdx.x == kScanThreads - 1) *total = part[kScanThreads - 1];
}

// Longer arrays scan in three launches: per-tile sums, a one-block scan of
// those, and a rescan of each tile from its offset.  One block's threads
// each read their words one after another, so a single-block scan of 250K
// words (the largest emission level of a 1M-triangle build) took 55 us.
constexpr uint32_t kScanTile = 4096;         // words per block: 256 threads x 16
constexpr uint32_t kScanSingleMax = 16384;   // longest array scanned by one block

// Thread t's 16 words of its tile, zero past m.
__device__ __forceinline__ void scan_run_load(const uint32_t *a, uint32_t m, uint32_t lo, uint32_t v[16]) {
    if (lo + 16u <= m) {
        const uint4 *p = reinterpret_cast<const uint4 *>(a + lo);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 x = p[q];
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = lo + q < m ? a[lo + q] : 0u;
    }
}

__global__ void __launch_bounds__(256) k_scan_tile_sums(const uint32_t *__restrict__ a, uint32_t m,
                                                        uint32_t *__restrict__ sums) {
    __shared__ uint32_t red[256];
    uint32_t v[16];
    scan_run_load(a, m, blockIdx.x * kScanTile + threadIdx.x * 16u, v);
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) s += v[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

// sums: exclusive offsets of the tiles.
__global__ void __launch_bounds__(256) k_scan_tile_apply(uint32_t *__restrict__ a, uint32_t m,
                                                         const uint32_t *__restrict__ sums) {
    __shared__ uint32_t part[256];
    const uint32_t lo = blockIdx.x * kScanTile + threadIdx.x * 16u;
    uint32_t v[16];
    scan_run_load(a, m, lo, v);
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) s += v[q];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = sums[blockIdx.x] + (threadIdx.x ? part[threadIdx.x - 1] : 0u);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint32_t x = v[q];
        v[q] = run;
        run += x;
    }
    if (lo + 16u <= m) {
        uint4 *p = reinterpret_cast<uint4 *>(a + lo);
#pragma unroll
        for (int q = 0; q < 4; ++q) p[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (lo + q < m) a[lo + q] = v[q];
    }
}

__global__ void __launch_bounds__(256) k_radix_scatter(const uint32_t *__restrict__ keys,
                                                       const uint32_t *__restrict__ vals, uint32_t n,
                                                       uint32_t shift, uint32_t ntiles,
                                                       const uint32_t *__restrict__ ghist,
                                                       uint32_t *__restrict__ okeys, uint32_t *__restrict__ ovals) {
    __shared__ uint32_t base_of[256];
    __shared__ uint32_t running[256];
    __shared__ uint32_t wave_cnt[4][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    base_of[tid] = ghist[static_cast<size_t>(tid) * ntiles + blockIdx.x];
    running[tid] = 0;
    const uint32_t tile = blockIdx.x * kRadixTile;
    for (uint32_t r = 0; r < kRadixTile / 256; ++r) {
        const uint32_t e = tile + r * 256 + tid;
        const bool valid = e < n;
        const uint32_t key = valid ? keys[e] : 0u;
        const uint32_t val = valid ? vals[e] : 0u;
        const uint32_t digit = (key >> shift) & 0xFFu;
        // Lanes of this wave holding the same digit.
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (digit >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(peers >> 32),
                                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(peers), 0u));
        for (uint32_t w = 0; w < 4; ++w) wave_cnt[w][tid] = 0;
        __syncthreads();
        if (valid && below == 0) wave_cnt[wave][digit] = static_cast<uint32_t>(__popcll(peers));
        __syncthreads();
        if (valid) {
            uint32_t pre = running[digit];
            for (uint32_t w = 0; w < wave; ++w) pre += wave_cnt[w][digit];
            const uint32_t dst = base_of[digit] + pre + below;
            okeys[dst] = key;
            ovals[dst] = val;
        }
        __syncthreads();
        running[tid] += wave_cnt[0][tid] + wave_cnt[1][tid] + wave_cnt[2][tid] + wave_cnt[3][tid];
        __syncthreads();
    }
    (void)lane;
}

// ---- 4. Karras radix tree ---------------------------------------------------

__device__ __forceinline__ int delta(const uint32_t *codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz(static_cast<uint32_t>(i ^ j));
    return __clz(a ^ b);
}

// Internal node i (0 = root) covers sorted range [first, last]; children refs
// have kLeafBit set for leaves (sorted position).
__global__ void __launch_bounds__(256) k_karras(const uint32_t *__restrict__ codes, uint32_t nn,
                                                uint32_t *__restrict__ left, uint32_t *__restrict__ right,
                                                uint32_t *__restrict__ first, uint32_t *__restrict__ last,
                                                uint32_t *__restrict__ parent_int, uint32_t *__restrict__ parent_leaf) {
    const int n = static_cast<int>(nn);
    const int i = static_cast<int>(blockIdx.x * 256 + threadIdx.x);
    if (i >= n - 1) return;
    const int d = delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1) > 0 ? 1 : -1;
    const int dmin = delta(codes, n, i, i - d);
    int lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(codes, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int g = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const uint32_t lref = lo == g ? (kLeafBit | static_cast<uint32_t>(g)) : static_cast<uint32_t>(g);
    const uint32_t rref = hi == g + 1 ? (kLeafBit | static_cast<uint32_t>(g + 1)) : static_cast<uint32_t>(g + 1);
    left[i] = lref;
    right[i] = rref;
    first[i] = static_cast<uint32_t>(lo);
    last[i] = static_cast<uint32_t>(hi);
    if (lref & kLeafBit) parent_leaf[g] = static_cast<uint32_t>(i); else parent_int[g] = static_cast<uint32_t>(i);
    if (rref & kLeafBit) parent_leaf[g + 1] = static_cast<uint32_t>(i); else parent_int[g + 1] = static_cast<uint32_t>(i);
}

// ---- 5. bottom-up boxes ------------------------------------------------------

// Order-preserving integer image of a float: ord(x) < ord(y) iff x < y, with
// -0 < +0 and NaNs placed by their bits.  Internal boxes are unions taken
// with integer atomic min/max on these images.
__device__ __forceinline__ int ord_f(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float unord_f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__device__ __forceinline__ int load_coherent(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Leaf j writes its triangle's box (boxes at [(n-1+j)*6], lo.xyz hi.xyz),
// then climbs: at each ancestor it folds its box into the ancestor's ordered
// box with atomic min/max, waits for those to return, and bumps the
// ancestor's arrival count.  The first arrival stops; the second knows both
// children's folds are complete (each was performed before its arrival
// count), reads the union and goes on.
//
// Block b holds leaves [256b, 256b + 255].  An internal node's leaf range
// starts or ends at its own index (Karras 2012), so a node whose whole range
// lies in the block has an index in the block too, and its children come
// from the block's own leaves: such nodes fold in LDS (lbox, lflag at
// index - 256b) and are copied to obox once the block is done.  Only nodes
// whose range crosses a block boundary fold in device memory, with
// device-scope atomics (obox, internal node i at [i*6]; flags), so no
// cache-wide fence is needed between the XCDs' L2s.  Folding every level in
// device memory took 512 us for 1M triangles.  Unions of min/max images do
// not depend on the order of the folds, so the boxes are the same.
__global__ void __launch_bounds__(256) k_refit(const float *__restrict__ tri, uint32_t n,
                                               const uint32_t *__restrict__ order,
                                               const uint32_t *__restrict__ parent_int,
                                               const uint32_t *__restrict__ parent_leaf,
                                               const uint32_t *__restrict__ first, const uint32_t *__restrict__ last,
                                               uint32_t *__restrict__ flags, int *obox, float *boxes) {
    __shared__ int lbox[256][6];
    __shared__ uint32_t lflag[256];
#pragma unroll
    for (int k = 0; k < 6; ++k) lbox[threadIdx.x][k] = k < 3 ? 0x7FFFFFFF : static_cast<int>(0x80000000u);
    lflag[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * 256u;
    const uint32_t bend = min(base + 255u, n - 1u);  // last leaf of the block
    const uint32_t j = base + threadIdx.x;
    int l[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    uint32_t node = kNone;
    if (j < n) {
        float lo[3], hi[3];
        tri_box(tri + static_cast<size_t>(order[j]) * 9, lo, hi);
        float *bl = boxes + static_cast<size_t>(n - 1 + j) * 6;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bl[k] = lo[k];
            bl[3 + k] = hi[k];
            l[k] = ord_f(lo[k]);
            h[k] = ord_f(hi[k]);
        }
        if (n > 1) node = parent_leaf[j];
    }
    // Nodes within the block, in LDS.
    while (node != kNone && first[node] >= base && last[node] <= bend) {
        const uint32_t li = node - base;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            l[k] = min(l[k], atomicMin(&lbox[li][k], l[k]));
            h[k] = max(h[k], atomicMax(&lbox[li][3 + k], h[k]));
        }
        asm volatile("" ::"v"(l[0]), "v"(l[1]), "v"(l[2]), "v"(h[0]), "v"(h[1]), "v"(h[2]) : "memory");
        if (atomicAdd(&lflag[li], 1u) == 0u) {
            node = kNone;  // the sibling subtree is not done yet
            break;
        }
        asm volatile("" ::: "memory");
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            l[k] = lbox[li][k];
            h[k] = lbox[li][3 + k];
        }
        node = parent_int[node];
    }
    __syncthreads();
    // Every node within the block has had both arrivals: publish its box.
    const uint32_t own = base + threadIdx.x;
    if (own + 1u < n && first[own] >= base && last[own] <= bend) {
#pragma unroll
        for (int k = 0; k < 6; ++k) obox[static_cast<size_t>(own) * 6 + k] = lbox[threadIdx.x][k];
    }
    // Nodes across block boundaries, in device memory.
    while (node != kNone) {
        int *ob = obox + static_cast<size_t>(node) * 6;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            l[k] = min(l[k], atomicMin(ob + k, l[k]));
            h[k] = max(h[k], atomicMax(ob + 3 + k, h[k]));
        }
        // The folds' return values are consumed here, so they have been
        // performed before the arrival count below is issued.
        asm volatile("" ::"v"(l[0]), "v"(l[1]), "v"(l[2]), "v"(h[0]), "v"(h[1]), "v"(h[2]) : "memory");
        if (atomicAdd(&flags[node], 1u) == 0u) return;  // the sibling subtree is not done yet
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            l[k] = load_coherent(ob + k);
            h[k] = load_coherent(ob + 3 + k);
        }
        node = parent_int[node];
    }
}

// Ordered images back to floats for the internal boxes.
__global__ void k_unord_boxes(const int *obox, float *boxes, size_t m) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < m;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        boxes[i] = unord_f(obox[i]);
}

__global__ void k_init_obox(int *obox, size_t m) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < m;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        obox[i] = (i % 6) < 3 ? 0x7FFFFFFF : static_cast<int>(0x80000000u);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t m) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < m;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

// ---- 6. emission ---------------------------------------------------------

struct Emit {
    const float *boxes;
    const uint32_t *left, *right, *first, *last;
    uint32_t n, max_leaf;
    double delta;  // padding
    const uint8_t *fin = nullptr;  // per internal node: 1 if it is one leaf of the final tree (k_final_flags)
};

// A child of the binary tree: internal node index, or kLeafBit | sorted leaf.
__device__ __forceinline__ uint32_t sub_count(const Emit &e, uint32_t c) {
    return (c & kLeafBit) ? 1u : e.last[c] - e.first[c] + 1u;
}
__device__ __forceinline__ const float *box_of(const Emit &e, uint32_t c) {
    const uint32_t b = (c & kLeafBit) ? (e.n - 1 + (c & ~kLeafBit)) : c;
    return e.boxes + static_cast<size_t>(b) * 6;
}
// Half surface area of an unpadded box, in the host builder's operation order.
__device__ __forceinline__ double box_half_area(const Emit &e, uint32_t c) {
    const float *b = box_of(e, c);
    const double dx = static_cast<double>(b[3]) - b[0], dy = static_cast<double>(b[4]) - b[1],
                 dz = static_cast<double>(b[5]) - b[2];
    return dx * dy + dy * dz + dz * dx;
}
// A binary subtree is one leaf of the final tree when it holds at most
// max_leaf triangles and the host SAH builder's leaf test accepts it
// (bvh_build.cpp: leaf iff count <= 1 + (A_l n_l + A_r n_r) / A, traversal
// and intersection costs both 1, A the half areas).  Forcing every small
// subtree into a leaf instead made leaves of 2-4 scattered triangles that
// most rays then tested: 1M incoherent rays over 1M clustered triangles
// traced 20% slower (BVH4).
__device__ __forceinline__ bool final_leaf_test(const Emit &e, uint32_t c) {
    const uint32_t n = sub_count(e, c);
    if (n > e.max_leaf) return false;
    const double a = box_half_area(e, c);
    if (!(a > 0.0)) return true;
    const uint32_t l = e.left[c], r = e.right[c];
    const double split = 1.0 + (box_half_area(e, l) * static_cast<double>(sub_count(e, l)) +
                                box_half_area(e, r) * static_cast<double>(sub_count(e, r))) /
                                   a;
    return static_cast<double>(n) <= split;
}
// The test of every internal node, once per build before the emission: the
// emission's collapse asks it of up to W - 1 children per step, and each
// answer is a chain of dependent loads (the node, its children, their
// boxes and counts) that every level would otherwise wait for again.
__global__ void __launch_bounds__(256) k_final_flags(Emit e, uint8_t *__restrict__ fin, uint32_t ni) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < ni) fin[i] = final_leaf_test(e, i) ? 1 : 0;
}
__device__ __forceinline__ bool is_final_leaf(const Emit &e, uint32_t c) {
    return (c & kLeafBit) != 0u || e.fin[c] != 0;
}
__device__ __forceinline__ uint32_t final_leaf_ref(const Emit &e, uint32_t c) {
    return (c & kLeafBit) ? make_leaf(c & ~kLeafBit, 1u) : make_leaf(e.first[c], sub_count(e, c));
}
// Padded box exactly as the host builder pads (bvh_build.cpp pad_and_refit).
__device__ __forceinline__ void padded(const Emit &e, uint32_t c, float lo[3], float hi[3]) {
    const float *b = box_of(e, c);
    for (int k = 0; k < 3; ++k) {
        const double l = static_cast<double>(b[k]) - e.delta;
        const double h = static_cast<double>(b[3 + k]) + e.delta;
        float fl = static_cast<float>(l), fh = static_cast<float>(h);
        if (static_cast<double>(fl) > l) fl = nextafterf(fl, -INFINITY);
        if (static_cast<double>(fh) < h) fh = nextafterf(fh, INFINITY);
        lo[k] = fl;
        hi[k] = fh;
    }
}
__device__ __forceinline__ double half_area(const float lo[3], const float hi[3]) {
    const double dx = static_cast<double>(hi[0]) - lo[0], dy = static_cast<double>(hi[1]) - lo[1],
                 dz = static_cast<double>(hi[2]) - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// Greedy collapse of binary node b into up to W children (largest surface
// area opened first, ties to the lowest slot), as the host collapse does.
template <int W>
__device__ int collapse(const Emit &e, uint32_t b, uint32_t kids[W]) {
    int nk = 2;
    kids[0] = e.left[b];
    kids[1] = e.right[b];
    while (nk < W) {
        int pick = -1;
        double best = -1.0;
        for (int i = 0; i < nk; ++i) {
            if (is_final_leaf(e, kids[i])) continue;
            float lo[3], hi[3];
            padded(e, kids[i], lo, hi);
            const double a = half_area(lo, hi);
            if (a > best) {
                best = a;
                pick = i;
            }
        }
        if (pick < 0) break;
        const uint32_t c = kids[pick];
        for (int i = nk; i > pick + 1; --i) kids[i] = kids[i - 1];
        kids[pick] = e.left[c];
        kids[pick + 1] = e.right[c];
        ++nk;
    }
    return nk;
}

// Pass A: children and interior-child counts of each frontier item.
template <int W>
__global__ void __launch_bounds__(256) k_emit_count(Emit e, const uint32_t *__restrict__ fbin, uint32_t nf,
                                                    uint32_t *__restrict__ kids_out, uint32_t *__restrict__ cnt) {
    const uint32_t f = blockIdx.x * 256 + threadIdx.x;
    if (f >= nf) return;
    uint32_t kids[W];
    const int nk = W == 2 ? (kids[0] = e.left[fbin[f]], kids[1] = e.right[fbin[f]], 2) : collapse<W>(e, fbin[f], kids);
    uint32_t interior = 0;
    for (int i = 0; i < W; ++i) {
        kids_out[static_cast<size_t>(f) * W + i] = i < nk ? kids[i] : kNone;
        if (i < nk && !is_final_leaf(e, kids[i])) ++interior;
    }
    cnt[f] = interior;
}

template <int W>
__device__ void quantise_dev(const float clo[][3], const float chi[][3], int nk, float origin[3], uint8_t ex[3],
                             uint8_t qlo[3][W], uint8_t qhi[3][W]) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < nk; ++c)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], clo[c][k]);
            hi[k] = fmaxf(hi[k], chi[c][k]);
        }
    for (int k = 0; k < 3; ++k) {
        const double ext = static_cast<double>(hi[k]) - lo[k];
        int x = -126;
        if (ext > 0.0) {
            // Boxes of finite coordinates need x <= 122 (the host refuses
            // scenes with infinite ones); the clamps bound the loops anyway.
            const double l = ceil(log2(ext / 255.0));
            x = l > 127.0 ? 127 : l < -126.0 ? -126 : static_cast<int>(l);
            while (x < 127 && ldexp(255.0, x) < ext) ++x;
            while (x > -126 && ldexp(255.0, x - 1) >= ext) --x;
        }
        origin[k] = lo[k];
        ex[k] = static_cast<uint8_t>(x + 127);
        const double s = ldexp(1.0, x);
        for (int c = 0; c < W; ++c) {
            if (c < nk) {
                double ql = floor((static_cast<double>(clo[c][k]) - lo[k]) / s);
                double qh = ceil((static_cast<double>(chi[c][k]) - lo[k]) / s);
                ql = fmin(fmax(ql, 0.0), 255.0);
                qh = fmin(fmax(qh, 0.0), 255.0);
                qlo[k][c] = static_cast<uint8_t>(ql);
                qhi[k][c] = static_cast<uint8_t>(qh);
            } else {
                qlo[k][c] = 255;
                qhi[k][c] = 0;
            }
        }
    }
}

// Pass B: write the nodes of this level and the next frontier.
template <int W>
__global__ void __launch_bounds__(256) k_emit_write(Emit e, const uint32_t *__restrict__ fbin,
                                                    const uint32_t *__restrict__ fidx,
                                                    const uint32_t *__restrict__ fstack, uint32_t nf,
                                                    const uint32_t *__restrict__ kids_in,
                                                    const uint32_t *__restrict__ offs, uint32_t next_base,
                                                    void *nodes, uint32_t *__restrict__ nbin,
                                                    uint32_t *__restrict__ nidx, uint32_t *__restrict__ nstack,
                                                    uint32_t *__restrict__ max_stack) {
    const uint32_t f = blockIdx.x * 256 + threadIdx.x;
    const bool valid = f < nf;
    uint32_t kids[W];
    int nk = 0;
    uint32_t stack_here = 0;
    if (valid) {
        for (int i = 0; i < W; ++i) {
            kids[i] = kids_in[static_cast<size_t>(f) * W + i];
            if (kids[i] != kNone) ++nk;
        }
        stack_here = fstack[f] + static_cast<uint32_t>(nk - 1);
    }
    // The tree's stack bound: one atomic per wave on its one word, not one
    // per node (same-address atomics serialise across the device).
    uint32_t wmax = stack_here;
    for (int off = 32; off > 0; off >>= 1)
        wmax = max(wmax, static_cast<uint32_t>(__shfl_xor(static_cast<int>(wmax), off)));
    if ((threadIdx.x & 63u) == 0u) atomicMax(max_stack, wmax);
    if (!valid) return;
    float clo[W][3], chi[W][3];
    uint32_t ref[W];
    uint32_t slot = 0;
    for (int i = 0; i < W; ++i) {
        if (i >= nk) {
            ref[i] = kRefNone;
            continue;
        }
        padded(e, kids[i], clo[i], chi[i]);
        if (is_final_leaf(e, kids[i])) {
            ref[i] = final_leaf_ref(e, kids[i]);
        } else {
            const uint32_t at = offs[f] + slot++;
            ref[i] = next_base + at;
            nbin[at] = kids[i];
            nidx[at] = next_base + at;
            nstack[at] = stack_here;
        }
    }
    const uint32_t me = fidx[f];
    if constexpr (W == 2) {
        NodeB2 nd;
        for (int k = 0; k < 3; ++k) {
            nd.lo0[k] = clo[0][k];
            nd.hi0[k] = chi[0][k];
            nd.lo1[k] = clo[1][k];
            nd.hi1[k] = chi[1][k];
        }
        nd.child[0] = ref[0];
        nd.child[1] = ref[1];
        nd.pad[0] = nd.pad[1] = 0;
        static_cast<NodeB2 *>(nodes)[me] = nd;
    } else if constexpr (W == 4) {
        NodeB4 nd;
        quantise_dev<4>(clo, chi, nk, nd.origin, nd.ex, nd.qlo, nd.qhi);
        nd.nchild = static_cast<uint8_t>(nk);
        for (int i = 0; i < 4; ++i) nd.child[i] = ref[i];
        nd.pad[0] = nd.pad[1] = 0;
        static_cast<NodeB4 *>(nodes)[me] = nd;
    } else {
        static_assert(W == 8, "node width");
        NodeB8 nd;
        quantise_dev<8>(clo, chi, nk, nd.origin, nd.ex, nd.qlo, nd.qhi);
        nd.nchild = static_cast<uint8_t>(nk);
        for (int i = 0; i < 8; ++i) nd.child[i] = ref[i];
        for (int i = 0; i < 8; ++i) nd.pad[i] = 0;
        static_cast<NodeB8 *>(nodes)[me] = nd;
    }
}

// Triangle records in sorted order and the optional input-order source.
__global__ void __launch_bounds__(256) k_records(const float *__restrict__ tri, uint32_t n,
                                                 const uint32_t *__restrict__ order, TriRec *__restrict__ recs,
                                                 float *__restrict__ src) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    {
        const uint32_t id = order[r];
        const float *t = tri + static_cast<size_t>(id) * 9;
        TriRec rec;
        for (int k = 0; k < 3; ++k) {
            rec.p0[k] = t[k];
            rec.e1[k] = t[3 + k] - t[k];
            rec.e2[k] = t[6 + k] - t[k];
        }
        rec.id = id;
        rec.pad1 = rec.pad2 = 0;
        recs[r] = rec;
    }
    if (src) {
        const float *t = tri + static_cast<size_t>(r) * 9;
        float *o = src + static_cast<size_t>(r) * 9;
        for (int k = 0; k < 3; ++k) {
            o[k] = t[k];
            o[3 + k] = t[3 + k] - t[k];
            o[6 + k] = t[6 + k] - t[k];
        }
    }
}

inline uint32_t blocks_for(size_t n, uint32_t per = 256) {
    return static_cast<uint32_t>(std::max<size_t>(1, (n + per - 1) / per));
}

// In-place exclusive scan of m words; *total (if not null) gets their sum.
// sums holds m / kScanTile + 1 words.
void launch_scan(uint32_t *a, uint32_t m, uint32_t *total, uint32_t *sums, hipStream_t st) {
    if (m <= kScanSingleMax) {
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanThreads), 0, st, a, m, total);
        return;
    }
    const uint32_t tiles = (m + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(tiles), dim3(256), 0, st, a, m, sums);
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanThreads), 0, st, sums, tiles, total);
    hipLaunchKernelGGL(k_scan_tile_apply, dim3(tiles), dim3(256), 0, st, a, m, sums);
}

}  // namespace

// Device scratch of the builder, grown on demand and reused across builds.
struct LbvhWorkspace {
    std::mutex mu;
    void *buf[28] = {};
    size_t cap[28] = {};

    hipError_t get(int slot, size_t bytes, void **out) {
        if (cap[slot] < bytes) {
            if (buf[slot]) (void)hipFree(buf[slot]);
            buf[slot] = nullptr;
            cap[slot] = 0;
            hipError_t e = rtx::dev_malloc(&buf[slot], std::max<size_t>(bytes, 256));
            if (e != hipSuccess) return e;
            cap[slot] = std::max<size_t>(bytes, 256);
        }
        *out = buf[slot];
        return hipSuccess;
    }
    ~LbvhWorkspace() {
        for (auto &b : buf)
            if (b) (void)hipFree(b);
    }
};

LbvhWorkspace *lbvh_workspace_new() { return new LbvhWorkspace; }
void lbvh_workspace_free(LbvhWorkspace *ws) { delete ws; }

namespace {

#define LB_CHECK(expr)                                   \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) {                          \
            why = std::string(#expr) + ": " + hipGetErrorString(e_); \
            return e_ == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP; \
        }                                                \
    } while (0)

struct Binary {
    uint32_t *codes, *order, *left, *right, *first, *last, *pint, *pleaf, *flags;
    int *obox;
    float *boxes, *bounds;
};

// Steps 1-5 into workspace buffers.
int build_binary(LbvhWorkspace *ws, const float *d_tri, uint32_t n, hipStream_t st, Binary &b, std::string &why) {
    const uint32_t ntiles = blocks_for(n, kRadixTile);
    const uint32_t nb_bounds = std::min<uint32_t>(1024, blocks_for(n));
    void *p;
    LB_CHECK(ws->get(0, static_cast<size_t>(n) * 4, &p)); b.codes = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(1, static_cast<size_t>(n) * 4, &p)); b.order = static_cast<uint32_t *>(p);
    uint32_t *kb, *vb;
    LB_CHECK(ws->get(2, static_cast<size_t>(n) * 4, &p)); kb = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(3, static_cast<size_t>(n) * 4, &p)); vb = static_cast<uint32_t *>(p);
    uint32_t *ghist;
    LB_CHECK(ws->get(4, static_cast<size_t>(ntiles) * 256 * 4, &p)); ghist = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(25, (static_cast<size_t>(ntiles) * 256 / kScanTile + 1) * 4, &p));
    uint32_t *rsums = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(5, static_cast<size_t>(nb_bounds) * 12 * 4 + 64, &p));
    float *partial = static_cast<float *>(p);
    LB_CHECK(ws->get(6, 16 * 4, &p)); b.bounds = static_cast<float *>(p);
    const size_t ni = n > 1 ? n - 1 : 1;
    LB_CHECK(ws->get(7, ni * 4, &p)); b.left = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(8, ni * 4, &p)); b.right = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(9, ni * 4, &p)); b.first = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(10, ni * 4, &p)); b.last = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(11, ni * 4, &p)); b.pint = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(12, static_cast<size_t>(n) * 4, &p)); b.pleaf = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(13, ni * 4, &p)); b.flags = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(14, (2 * static_cast<size_t>(n)) * 6 * 4, &p)); b.boxes = static_cast<float *>(p);
    LB_CHECK(ws->get(24, ni * 6 * 4, &p)); b.obox = static_cast<int *>(p);

    hipLaunchKernelGGL(k_bounds_partial, dim3(nb_bounds), dim3(256), 0, st, d_tri, n, partial);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, st, partial, nb_bounds, b.bounds);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(256), 0, st, d_tri, n, b.bounds, b.codes, b.order);
    uint32_t *ks = b.codes, *vs = b.order, *kd = kb, *vd = vb;
    for (uint32_t shift = 0; shift < 32; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(256), 0, st, ks, n, shift, ntiles, ghist);
        launch_scan(ghist, ntiles * 256u, nullptr, rsums, st);
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(256), 0, st, ks, vs, n, shift, ntiles, ghist, kd, vd);
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    // Four passes: the sorted keys are back in b.codes / b.order.
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(ni)), dim3(256), 0, st, b.pint, kNone, ni);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(ni)), dim3(256), 0, st, b.flags, 0u, ni);
    if (n > 1)
        hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1)), dim3(256), 0, st, b.codes, n, b.left, b.right, b.first,
                           b.last, b.pint, b.pleaf);
    else
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, b.pleaf, kNone, static_cast<size_t>(1));
    hipLaunchKernelGGL(k_init_obox, dim3(blocks_for(ni * 6)), dim3(256), 0, st, b.obox, ni * 6);
    hipLaunchKernelGGL(k_refit, dim3(blocks_for(n)), dim3(256), 0, st, d_tri, n, b.order, b.pint, b.pleaf, b.first,
                       b.last, b.flags, b.obox, b.boxes);
    if (n > 1)
        hipLaunchKernelGGL(k_unord_boxes, dim3(blocks_for(ni * 6)), dim3(256), 0, st, b.obox, b.boxes, ni * 6);
    LB_CHECK(hipGetLastError());
    return RT_OK;
}

template <int W>
int emit_levels(LbvhWorkspace *ws, const Emit &e, uint32_t root_bin, uint32_t max_nodes, hipStream_t st,
                void *d_nodes, uint32_t &num_nodes, uint32_t &max_depth, uint32_t &max_stack, std::string &why) {
    void *p;
    uint32_t *fb[2], *fi[2], *fs[2], *kids, *cnt, *tot, *mstack;
    const size_t cap = std::max<uint32_t>(max_nodes, 1);
    LB_CHECK(ws->get(15, cap * 4, &p)); fb[0] = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(16, cap * 4, &p)); fb[1] = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(17, cap * 4, &p)); fi[0] = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(18, cap * 4, &p)); fi[1] = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(19, cap * 4, &p)); fs[0] = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(20, cap * 4, &p)); fs[1] = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(21, cap * W * 4, &p)); kids = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(22, cap * 4, &p)); cnt = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(23, 64, &p)); tot = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(26, (cap / kScanTile + 1) * 4, &p));
    uint32_t *esums = static_cast<uint32_t *>(p);
    LB_CHECK(ws->get(27, e.n, &p));  // n - 1 internal nodes (n >= 2 here)
    Emit ef = e;
    ef.fin = static_cast<const uint8_t *>(p);
    hipLaunchKernelGGL(k_final_flags, dim3(blocks_for(e.n - 1)), dim3(256), 0, st, e, static_cast<uint8_t *>(p),
                       e.n - 1);
    mstack = tot + 4;
    const uint32_t zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    LB_CHECK(hipMemcpyAsync(tot, zero, sizeof zero, hipMemcpyHostToDevice, st));
    LB_CHECK(hipMemcpyAsync(fb[0], &root_bin, 4, hipMemcpyHostToDevice, st));
    LB_CHECK(hipMemcpyAsync(fi[0], zero, 4, hipMemcpyHostToDevice, st));
    LB_CHECK(hipMemcpyAsync(fs[0], zero, 4, hipMemcpyHostToDevice, st));
    uint32_t nf = 1, base = 1, depth = 0;
    int cur = 0;
    while (nf) {
        ++depth;
        hipLaunchKernelGGL(k_emit_count<W>, dim3(blocks_for(nf)), dim3(256), 0, st, ef, fb[cur], nf, kids, cnt);
        launch_scan(cnt, nf, tot, esums, st);
        uint32_t next = 0;
        LB_CHECK(hipMemcpyAsync(&next, tot, 4, hipMemcpyDeviceToHost, st));
        LB_CHECK(hipStreamSynchronize(st));
        if (base + next > max_nodes) {
            why = "LBVH node count exceeded its bound";
            return RT_ERR_OVERFLOW;
        }
        hipLaunchKernelGGL(k_emit_write<W>, dim3(blocks_for(nf)), dim3(256), 0, st, ef, fb[cur], fi[cur], fs[cur], nf,
                           kids, cnt, base, d_nodes, fb[cur ^ 1], fi[cur ^ 1], fs[cur ^ 1], mstack);
        LB_CHECK(hipGetLastError());
        base += next;
        nf = next;
        cur ^= 1;
    }
    uint32_t ms = 0;
    LB_CHECK(hipMemcpyAsync(&ms, mstack, 4, hipMemcpyDeviceToHost, st));
    LB_CHECK(hipStreamSynchronize(st));
    num_nodes = base;
    max_depth = depth;
    max_stack = ms;
    return RT_OK;
}

}  // namespace

int lbvh_build_device(LbvhWorkspace *ws, const float *d_tri, uint32_t n, const rt_build_opts &o, hipStream_t st,
                      LbvhResult &res, std::string &why) {
    std::lock_guard<std::mutex> g(ws->mu);
    const auto t0 = std::chrono::steady_clock::now();
    res = LbvhResult();
    if (n > kMaxLeafOffset) {
        why = "too many triangles for one scene";
        return RT_ERR_INVALID;
    }
    const uint32_t W = o.node_width;
    auto cleanup = [&]() {
        if (res.d_nodes) (void)hipFree(res.d_nodes);
        if (res.d_tris) (void)hipFree(res.d_tris);
        if (res.d_src) (void)hipFree(res.d_src);
        res.d_nodes = res.d_tris = nullptr;
        res.d_src = nullptr;
    };
    hipError_t e;
    if (n) {
        if ((e = rtx::dev_malloc(&res.d_tris, static_cast<size_t>(n) * sizeof(TriRec))) != hipSuccess ||
            (o.keep_source && (e = rtx::dev_malloc(&res.d_src, static_cast<size_t>(n) * 9 * sizeof(float))) != hipSuccess)) {
            cleanup();
            why = std::string("scene memory: ") + hipGetErrorString(e);
            return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
        }
    }
    if (n == 0) {
        res.root_ref = kRefNone;
        return RT_OK;
    }
    Binary b;
    int rc = build_binary(ws, d_tri, n, st, b, why);
    if (rc != RT_OK) {
        cleanup();
        return rc;
    }
    hipLaunchKernelGGL(k_records, dim3(blocks_for(n)), dim3(256), 0, st, d_tri, n, b.order,
                       static_cast<TriRec *>(res.d_tris), res.d_src);
    // Scene bounds and the padding they imply (bvh_build.cpp pad_and_refit).
    float bounds[12];
    if ((e = hipMemcpyAsync(bounds, b.bounds, sizeof bounds, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess) {
        cleanup();
        why = std::string("bounds: ") + hipGetErrorString(e);
        return RT_ERR_HIP;
    }
    // A box with an infinite side has no quantised form.  (The bounds
    // reduction ignores NaN coordinates, as the host builder does.)
    for (int k = 0; k < 3; ++k)
        if (bounds[k] == -INFINITY || bounds[3 + k] == INFINITY) {
            cleanup();
            why = "a triangle has an infinite coordinate";
            return RT_ERR_INVALID;
        }
    double scale = 0.0;
    for (int k = 0; k < 3; ++k) {
        scale = std::max(scale, std::fabs(static_cast<double>(bounds[k])));
        scale = std::max(scale, std::fabs(static_cast<double>(bounds[3 + k])));
        scale = std::max(scale, static_cast<double>(bounds[3 + k]) - bounds[k]);
    }
    Emit em;
    em.boxes = b.boxes;
    em.left = b.left;
    em.right = b.right;
    em.first = b.first;
    em.last = b.last;
    em.n = n;
    em.max_leaf = o.max_leaf;
    em.delta = std::max(scale, 1e-30) * o.pad_rel;
    for (int k = 0; k < 3; ++k) {
        res.lo[k] = bounds[k];
        res.hi[k] = bounds[3 + k];
    }
    if (n <= o.max_leaf) {  // the whole scene is one leaf
        res.root_ref = make_leaf(0, n);
    } else {
        // Interior nodes of the final tree: at most n - 1 (binary) and fewer
        // when wider; allocate the binary bound.
        const uint32_t max_nodes = n - 1;
        const size_t nbytes = static_cast<size_t>(max_nodes) * (W == 8 ? 128 : 64);
        if ((e = rtx::dev_malloc(&res.d_nodes, nbytes)) != hipSuccess) {
            cleanup();
            why = std::string("node memory: ") + hipGetErrorString(e);
            return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
        }
        uint32_t nn = 0, md = 0, ms = 0;
        if (W == 2)
            rc = emit_levels<2>(ws, em, 0, max_nodes, st, res.d_nodes, nn, md, ms, why);
        else if (W == 4)
            rc = emit_levels<4>(ws, em, 0, max_nodes, st, res.d_nodes, nn, md, ms, why);
        else
            rc = emit_levels<8>(ws, em, 0, max_nodes, st, res.d_nodes, nn, md, ms, why);
        if (rc != RT_OK) {
            cleanup();
            return rc;
        }
        res.root_ref = 0;
        res.num_nodes = nn;
        res.max_depth = md;
        res.max_stack = ms;
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) {
        cleanup();
        why = std::string("build: ") + hipGetErrorString(e);
        return RT_ERR_HIP;
    }
    res.num_leaves = 0;  // not tracked by the device builder
    res.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RT_OK;
}

int lbvh_binary(LbvhWorkspace *ws, const float *d_tri, uint32_t n, hipStream_t st, uint32_t *morton,
                uint32_t *order, uint32_t *left, uint32_t *right, float *boxes, std::string &why) {
    std::lock_guard<std::mutex> g(ws->mu);
    if (n == 0) return RT_OK;
    Binary b;
    int rc = build_binary(ws, d_tri, n, st, b, why);
    if (rc != RT_OK) return rc;
    LB_CHECK(hipMemcpyAsync(morton, b.codes, static_cast<size_t>(n) * 4, hipMemcpyDeviceToDevice, st));
    LB_CHECK(hipMemcpyAsync(order, b.order, static_cast<size_t>(n) * 4, hipMemcpyDeviceToDevice, st));
    if (n > 1) {
        LB_CHECK(hipMemcpyAsync(left, b.left, static_cast<size_t>(n - 1) * 4, hipMemcpyDeviceToDevice, st));
        LB_CHECK(hipMemcpyAsync(right, b.right, static_cast<size_t>(n - 1) * 4, hipMemcpyDeviceToDevice, st));
    }
    LB_CHECK(hipMemcpyAsync(boxes, b.boxes, (2 * static_cast<size_t>(n) - 1) * 6 * 4, hipMemcpyDeviceToDevice, st));
    LB_CHECK(hipStreamSynchronize(st));
    return RT_OK;
}

}  // namespace rtx
