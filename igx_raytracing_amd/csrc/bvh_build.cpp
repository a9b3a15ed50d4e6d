// Host BVH builders.  The reference has no acceleration structure at all (its
// traversal is the linear scan of res/shaders/trace.glsl:25-27 and
// "Raytracing optimization.md":16-18 is an empty heading); this file builds
// the binary SAH tree the query kernels traverse and its wide, quantised
// forms.  Results of a query never depend on the tree's shape: the kernels
// apply the scan's (t, lowest index) rule to whatever order they visit.
#include "bvh_build.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <utility>

namespace rtx {

namespace {

inline float tri_coord(const float *t, uint32_t i, int v, int k) {
    return t[static_cast<size_t>(i) * 9 + v * 3 + k];
}

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float *p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], p[k]);
            hi[k] = std::fmax(hi[k], p[k]);
        }
    }
    void grow(const Box &b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], b.lo[k]);
            hi[k] = std::fmax(hi[k], b.hi[k]);
        }
    }
    double half_area() const {
        if (!(lo[0] <= hi[0])) return 0.0;
        double dx = static_cast<double>(hi[0]) - lo[0], dy = static_cast<double>(hi[1]) - lo[1],
               dz = static_cast<double>(hi[2]) - lo[2];
        return dx * dy + dy * dz + dz * dx;
    }
};

struct Prim {
    Box box;
    float c[3];
    uint32_t id;  // triangle (or, with reference splitting, part) index
};

unsigned host_threads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(64u, hw ? hw : 1u));
}

// f(slice, lo, hi) over [0, n) in host_threads() slices, in parallel when n
// is at least min_par; f's effect must not depend on the slicing.
template <typename F>
void parallel_slices(size_t n, size_t min_par, F &&f) {
    const unsigned t = n >= min_par ? host_threads() : 1u;
    std::vector<std::thread> ts;
    for (unsigned j = 1; j < t; ++j)
        ts.emplace_back([&f, j, n, t] { f(j, n * j / t, n * (j + 1) / t); });
    f(0u, 0, n / t);
    for (std::thread &th : ts) th.join();
}

}  // namespace

bool resolve_opts(const rt_build_opts *in, rt_build_opts &o, std::string &err) {
    rt_build_opts_default(&o);
    if (in) o = *in;
    if (o.max_leaf == 0) o.max_leaf = 4;
    if (o.sah_bins == 0) o.sah_bins = 32;
    if (!(o.pad_rel > 0.f)) o.pad_rel = 1.0f / 65536.0f;
    if (o.node_width != 2 && o.node_width != 4 && o.node_width != 8) {
        err = "node_width must be 2, 4 or 8";
        return false;
    }
    if (o.max_leaf > kMaxLeafTris) {
        err = "max_leaf must be 1..8";
        return false;
    }
    if (o.sah_bins < 4 || o.sah_bins > 256) {
        err = "sah_bins must be 4..256";
        return false;
    }
    if (o.builder != RT_BUILD_SAH_CPU && o.builder != RT_BUILD_LBVH_GPU) {
        err = "unknown builder";
        return false;
    }
    if (!(o.pad_rel < 0.01f)) {
        err = "pad_rel must be in (0, 0.01)";
        return false;
    }
    if (o.presplit == 0) o.presplit = 1;
    if (o.presplit > 64) {
        err = "presplit must be 1..64";
        return false;
    }
    if (o.presplit > 1 && o.builder != RT_BUILD_SAH_CPU) {
        err = "presplit needs the host SAH builder";
        return false;
    }
    return true;
}

namespace {

// A convex polygon: a triangle clipped by axis-aligned planes (each plane
// adds at most one vertex; 64 pieces take at most 6 planes per piece, so 9
// vertices).  push() refuses a vertex beyond the capacity.
struct Poly {
    static constexpr int kCap = 12;
    double v[kCap][3];
    int n = 0;
    bool push(const double *p) {
        if (n == kCap) return false;
        for (int k = 0; k < 3; ++k) v[n][k] = p[k];
        ++n;
        return true;
    }
};

// Sutherland-Hodgman against the plane x_axis = s: the parts at or below
// and at or above it.  If rounding made a near-degenerate part cross the
// plane more often than a convex polygon can, both outputs are left empty
// (no cut: the caller keeps the part whole, whose box covers it).
void clip_poly(const Poly &in, int axis, double s, Poly &below, Poly &above) {
    below.n = above.n = 0;
    bool ok = true;
    for (int i = 0; i < in.n && ok; ++i) {
        const double *a = in.v[i], *b = in.v[(i + 1) % in.n];
        const double da = a[axis] - s, db = b[axis] - s;
        if (da <= 0.0) ok = ok && below.push(a);
        if (da >= 0.0) ok = ok && above.push(a);
        if ((da < 0.0 && db > 0.0) || (da > 0.0 && db < 0.0)) {
            const double t = da / (da - db);
            double p[3];
            for (int k = 0; k < 3; ++k) p[k] = a[k] + t * (b[k] - a[k]);
            p[axis] = s;
            ok = ok && below.push(p) && above.push(p);
        }
    }
    if (!ok) below.n = above.n = 0;
}

// The float box of a polygon, rounded outward and kept inside `cap`.
Box poly_box(const Poly &p, const Box &cap) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < p.n; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p.v[i][k]);
            hi[k] = std::max(hi[k], p.v[i][k]);
        }
    Box b;
    for (int k = 0; k < 3; ++k) {
        float fl = static_cast<float>(lo[k]), fh = static_cast<float>(hi[k]);
        if (static_cast<double>(fl) > lo[k]) fl = std::nextafter(fl, -INFINITY);
        if (static_cast<double>(fh) < hi[k]) fh = std::nextafter(fh, INFINITY);
        b.lo[k] = std::fmax(fl, cap.lo[k]);
        b.hi[k] = std::fmin(fh, cap.hi[k]);
    }
    return b;
}

// Cuts the part `p` (box `b`) of triangle `t` at the midpoint of the box's
// longest axis, recursively, while the box's half area exceeds `limit` and
// the budget allows, appending each final part's box.
void split_piece(const Poly &p, const Box &b, uint32_t budget, double limit, uint32_t t, std::vector<Prim> &out,
                 std::vector<uint32_t> &tri_of) {
    int axis = 0;
    for (int k = 1; k < 3; ++k)
        if (static_cast<double>(b.hi[k]) - b.lo[k] > static_cast<double>(b.hi[axis]) - b.lo[axis]) axis = k;
    const double s = 0.5 * (static_cast<double>(b.lo[axis]) + static_cast<double>(b.hi[axis]));
    Poly lo_p, hi_p;
    if (budget > 1 && b.half_area() > limit && s > b.lo[axis] && s < b.hi[axis]) clip_poly(p, axis, s, lo_p, hi_p);
    if (lo_p.n < 3 || hi_p.n < 3) {  // no cut: this part is final
        Prim q;
        q.box = b;
        for (int k = 0; k < 3; ++k) q.c[k] = 0.5f * b.lo[k] + 0.5f * b.hi[k];
        out.push_back(q);
        tri_of.push_back(t);
        return;
    }
    const uint32_t lb = budget / 2;
    split_piece(lo_p, poly_box(lo_p, b), lb, limit, t, out, tri_of);
    split_piece(hi_p, poly_box(hi_p, b), budget - lb, limit, t, out, tri_of);
}

}  // namespace

namespace {

using Node = BinTree::Node;
constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// The binned SAH build over prims[first, first + count), which it sorts in
// place.  The primitives themselves are partitioned (not an index array),
// so every pass over a node's range reads memory in order.  Subtrees of
// at least kSpawn primitives on both sides are built on threads of their
// own into separate node arrays ("parts") that are appended to their
// parent's at the end; ranges of at least kSplitPass primitives have their
// bounds and bins computed by several threads.  Neither changes the tree:
// bins merge by integer sums and fmin/fmax, which do not depend on the
// order, and every split decision is the sequential build's.  Only the
// numbering of the nodes depends on the threading; the emitters walk the
// tree from its root, so their output does not.
class SahBuilder {
   public:
    struct Part {
        std::vector<Node> nodes;
        // (node of this part whose right child is the graft's root, graft)
        std::vector<std::pair<uint32_t, std::unique_ptr<Part>>> grafts;
    };

    SahBuilder(std::vector<Prim> &prims, uint32_t bins, uint32_t maxleaf)
        : prims_(prims), tmp_(prims.size()), B_(bins), maxleaf_(maxleaf) {
        const unsigned hw = std::thread::hardware_concurrency();
        nthreads_ = std::max(1u, std::min(64u, hw ? hw : 1u));
        spare_ = static_cast<int>(nthreads_) - 1;
    }

    // Builds the whole tree into `out` (root = out[0]).
    void build(std::vector<Node> &out) {
        Part root;
        run(root, 0, static_cast<uint32_t>(prims_.size()));
        for (;;) {  // threads may start threads: join until none is left
            std::vector<std::thread> ts;
            {
                std::lock_guard<std::mutex> g(mu_);
                ts.swap(threads_);
            }
            if (ts.empty()) break;
            for (std::thread &t : ts) t.join();
        }
        out.clear();
        merge(out, root);
    }

   private:
    static constexpr uint32_t kSpawn = 1u << 13;
    static constexpr uint32_t kSplitPass = 1u << 16;
    static constexpr uint32_t kSmall = 64;  // nodes this small skip the bin arrays

    struct Bins {
        std::vector<uint32_t> cnt;
        std::vector<Box> box;
        explicit Bins(uint32_t B) : cnt(3 * B, 0u), box(3 * B) {}
        void add(const Bins &o) {
            for (size_t i = 0; i < cnt.size(); ++i) {
                cnt[i] += o.cnt[i];
                box[i].grow(o.box[i]);
            }
        }
    };

    // Runs f(lo, hi) over [first, first + count) in nthreads_ slices (f's
    // results must not depend on how the range is sliced).
    template <typename F>
    void slices(uint32_t first, uint32_t count, F &&f) {
        const uint32_t t = nthreads_;
        std::vector<std::thread> ts;
        for (uint32_t j = 1; j < t; ++j) {
            const uint32_t lo = first + static_cast<uint32_t>(static_cast<uint64_t>(count) * j / t);
            const uint32_t hi = first + static_cast<uint32_t>(static_cast<uint64_t>(count) * (j + 1) / t);
            ts.emplace_back([&f, j, lo, hi] { f(j, lo, hi); });
        }
        f(0u, first, first + static_cast<uint32_t>(static_cast<uint64_t>(count) / t));
        for (std::thread &th : ts) th.join();
    }

    void bounds(uint32_t first, uint32_t count, Box &nb, Box &cb) {
        auto pass = [&](uint32_t lo, uint32_t hi, Box &b, Box &c) {
            for (uint32_t i = lo; i < hi; ++i) {
                b.grow(prims_[i].box);
                c.grow(prims_[i].c);
            }
        };
        if (count < kSplitPass || nthreads_ == 1) {
            pass(first, first + count, nb, cb);
            return;
        }
        std::vector<Box> b(nthreads_), c(nthreads_);
        slices(first, count, [&](uint32_t j, uint32_t lo, uint32_t hi) { pass(lo, hi, b[j], c[j]); });
        for (uint32_t j = 0; j < nthreads_; ++j) {
            nb.grow(b[j]);
            cb.grow(c[j]);
        }
    }

    void bin(uint32_t first, uint32_t count, const Box &cb, const double *inv_extent, Bins &bins) {
        const uint32_t B = B_;
        auto pass = [&](uint32_t lo, uint32_t hi, Bins &out) {
            for (uint32_t i = lo; i < hi; ++i) {
                const Prim &p = prims_[i];
                for (int k = 0; k < 3; ++k) {
                    if (inv_extent[k] == 0.0) continue;
                    uint32_t b = static_cast<uint32_t>((p.c[k] - static_cast<double>(cb.lo[k])) * inv_extent[k]);
                    if (b >= B) b = B - 1;
                    out.cnt[k * B + b]++;
                    out.box[k * B + b].grow(p.box);
                }
            }
        };
        if (count < kSplitPass || nthreads_ == 1) {
            pass(first, first + count, bins);
            return;
        }
        std::vector<Bins> part(nthreads_, Bins(B));
        slices(first, count, [&](uint32_t j, uint32_t lo, uint32_t hi) { pass(lo, hi, part[j]); });
        for (uint32_t j = 0; j < nthreads_; ++j) bins.add(part[j]);
    }

    // The binned split of a small node without the B-bin arrays: only the
    // planes just above a non-empty bin can win, because the planes up to
    // the next non-empty bin have the same cost and the sweep keeps the
    // lowest plane of equal cost.  The boxes, counts and costs compared are
    // those of the full sweep, so the choice is the same.
    void best_split_small(uint32_t first, uint32_t count, const Box &cb, const double *inv_extent, double &best_cost,
                          int &best_axis, uint32_t &best_plane) const {
        const uint32_t B = B_;
        uint32_t binof[kSmall], idx[kSmall], gbin[kSmall], gcnt[kSmall];
        Box gbox[kSmall];
        double rc[kSmall + 1];
        for (int k = 0; k < 3; ++k) {
            if (inv_extent[k] == 0.0) continue;
            for (uint32_t i = 0; i < count; ++i) {
                uint32_t b = static_cast<uint32_t>((prims_[first + i].c[k] - static_cast<double>(cb.lo[k])) *
                                                   inv_extent[k]);
                if (b >= B) b = B - 1;
                binof[i] = b;
                idx[i] = i;
            }
            for (uint32_t i = 1; i < count; ++i) {  // insertion sort by bin
                const uint32_t v = idx[i];
                uint32_t j = i;
                while (j > 0 && binof[idx[j - 1]] > binof[v]) {
                    idx[j] = idx[j - 1];
                    --j;
                }
                idx[j] = v;
            }
            uint32_t G = 0;
            for (uint32_t i = 0; i < count; ++i) {
                const uint32_t b = binof[idx[i]];
                if (G == 0 || gbin[G - 1] != b) {
                    gbin[G] = b;
                    gcnt[G] = 0;
                    gbox[G] = Box();
                    ++G;
                }
                gcnt[G - 1]++;
                gbox[G - 1].grow(prims_[first + idx[i]].box);
            }
            Box acc;
            uint32_t cnt = 0;
            rc[G] = 0.0;
            for (uint32_t g = G; g-- > 0;) {
                acc.grow(gbox[g]);
                cnt += gcnt[g];
                rc[g] = cnt ? acc.half_area() * cnt : 0.0;
            }
            Box lacc;
            uint32_t lcnt = 0;
            for (uint32_t g = 0; g + 1 < G; ++g) {
                lacc.grow(gbox[g]);
                lcnt += gcnt[g];
                if (lcnt == 0 || lcnt == count) continue;
                const double c = lacc.half_area() * lcnt + rc[g + 1];
                if (c < best_cost) {
                    best_cost = c;
                    best_axis = k;
                    best_plane = gbin[g];
                }
            }
        }
    }

    bool take_thread() {
        int v = spare_.load();
        while (v > 0)
            if (spare_.compare_exchange_weak(v, v - 1)) return true;
        return false;
    }

    void run(Part &part, uint32_t first, uint32_t count) {
        const uint32_t B = B_;
        part.nodes.emplace_back();
        struct Work {
            uint32_t node, first, count;
        };
        std::vector<Work> stack;
        stack.push_back({0, first, count});
        Bins bins(B);
        std::vector<double> right_cost(B);
        while (!stack.empty()) {
            Work w = stack.back();
            stack.pop_back();
            Box nb, cb;
            bounds(w.first, w.count, nb, cb);
            {
                Node &nd = part.nodes[w.node];
                std::memcpy(nd.lo, nb.lo, sizeof nd.lo);
                std::memcpy(nd.hi, nb.hi, sizeof nd.hi);
            }
            auto make_leaf_node = [&]() {
                Node &nd = part.nodes[w.node];
                nd.first = w.first;
                nd.count = w.count;
            };
            if (w.count == 1) {
                make_leaf_node();
                continue;
            }
            // Best binned split over the three axes (ties: lowest axis, lowest plane).
            double best_cost = INFINITY;
            int best_axis = -1;
            uint32_t best_plane = 0;
            double inv_extent[3];
            for (int k = 0; k < 3; ++k) {
                double ext = static_cast<double>(cb.hi[k]) - cb.lo[k];
                inv_extent[k] = ext > 0.0 ? B / ext : 0.0;
            }
            if (w.count <= kSmall) {
                best_split_small(w.first, w.count, cb, inv_extent, best_cost, best_axis, best_plane);
            } else {
            std::fill(bins.cnt.begin(), bins.cnt.end(), 0u);
            std::fill(bins.box.begin(), bins.box.end(), Box());
            bin(w.first, w.count, cb, inv_extent, bins);
            for (int k = 0; k < 3; ++k) {
                if (inv_extent[k] == 0.0) continue;
                Box acc;
                uint32_t cnt = 0;
                for (uint32_t b = B - 1; b >= 1; --b) {
                    acc.grow(bins.box[k * B + b]);
                    cnt += bins.cnt[k * B + b];
                    right_cost[b] = cnt ? acc.half_area() * cnt : 0.0;
                }
                Box lacc;
                uint32_t lcnt = 0;
                for (uint32_t b = 0; b + 1 < B; ++b) {
                    lacc.grow(bins.box[k * B + b]);
                    lcnt += bins.cnt[k * B + b];
                    if (lcnt == 0 || lcnt == w.count) continue;
                    double c = lacc.half_area() * lcnt + right_cost[b + 1];
                    if (c < best_cost) {
                        best_cost = c;
                        best_axis = k;
                        best_plane = b;
                    }
                }
            }
            }
            double parent_area = nb.half_area();
            double split_cost = best_axis >= 0 && parent_area > 0.0 ? 1.0 + best_cost / parent_area : INFINITY;
            if (w.count <= maxleaf_ && static_cast<double>(w.count) <= split_cost) {
                make_leaf_node();
                continue;
            }
            uint32_t mid = 0;
            if (best_axis >= 0) {
                // Stable partition of the primitives (the left side keeps
                // its order, then the right side): the order a partition of
                // indices would give.
                const int k = best_axis;
                const double lo = cb.lo[k], inv = inv_extent[k];
                Prim *r = prims_.data() + w.first;
                Prim *t = tmp_.data() + w.first;
                uint32_t nl = 0, nr = 0;
                for (uint32_t i = 0; i < w.count; ++i) {
                    uint32_t b = static_cast<uint32_t>((r[i].c[k] - lo) * inv);
                    if (b >= B) b = B - 1;
                    if (b <= best_plane)
                        r[nl++] = r[i];
                    else
                        t[nr++] = r[i];
                }
                std::copy(t, t + nr, r + nl);
                mid = nl;
            }
            if (mid == 0 || mid == w.count) mid = w.count / 2;  // coincident centroids: split by position
            const uint32_t l = static_cast<uint32_t>(part.nodes.size());
            part.nodes[w.node].count = 0;
            if (mid >= kSpawn && w.count - mid >= kSpawn && take_thread()) {
                // The right subtree on a thread of its own, as a graft.
                part.nodes.emplace_back();
                part.nodes[w.node].left = l;
                part.nodes[w.node].right = kNoNode;
                auto g = std::make_unique<Part>();
                Part *gp = g.get();
                part.grafts.emplace_back(w.node, std::move(g));
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    threads_.emplace_back([this, gp, f = w.first + mid, c = w.count - mid] {
                        run(*gp, f, c);
                        spare_.fetch_add(1);
                    });
                }
                stack.push_back({l, w.first, mid});
            } else {
                part.nodes.emplace_back();
                part.nodes.emplace_back();
                part.nodes[w.node].left = l;
                part.nodes[w.node].right = l + 1;
                // Right pushed first so the left subtree is finished first.
                stack.push_back({l + 1, w.first + mid, w.count - mid});
                stack.push_back({l, w.first, mid});
            }
        }
    }

    // Appends part p (and, after it, its grafts in the order they were
    // made) to out; children always follow their parents.
    static void merge(std::vector<Node> &out, Part &p) {
        const uint32_t base = static_cast<uint32_t>(out.size());
        for (Node nd : p.nodes) {
            if (nd.count == 0) {
                nd.left += base;
                if (nd.right != kNoNode) nd.right += base;
            }
            out.push_back(nd);
        }
        for (auto &g : p.grafts) {
            const uint32_t root = static_cast<uint32_t>(out.size());
            merge(out, *g.second);
            out[base + g.first].right = root;
        }
    }

    std::vector<Prim> &prims_;
    std::vector<Prim> tmp_;
    const uint32_t B_, maxleaf_;
    unsigned nthreads_ = 1;
    std::atomic<int> spare_{0};
    std::mutex mu_;
    std::vector<std::thread> threads_;
};

}  // namespace

bool build_sah_tree(const float *tri, uint32_t n, const rt_build_opts &o, BinTree &tree,
                    std::string &err) {
    tree.nodes.clear();
    tree.order.clear();
    tree.piece.clear();
    if (n == 0) return true;
    if (n > kMaxLeafOffset) {
        err = "too many triangles for one scene";
        return false;
    }
    std::vector<Prim> prims(n);
    for (uint32_t i = 0; i < n; ++i) {
        Prim &p = prims[i];
        for (int v = 0; v < 3; ++v) p.box.grow(tri + static_cast<size_t>(i) * 9 + v * 3);
        for (int k = 0; k < 3; ++k) p.c[k] = 0.5f * p.box.lo[k] + 0.5f * p.box.hi[k];
        p.id = i;
    }
    // Reference splitting: triangles whose box is larger than a quarter of
    // the mean box (in half area) are cut into parts with tight boxes
    // (split_piece) until the parts are that small or the budget is used;
    // the build then runs over the parts, and part j belongs to triangle
    // tri_of[j].  On the displaced-icosphere path-tracing scene a quarter of
    // the mean measured 8% faster than the mean itself (28% against 18% over
    // no splitting, with 2.5x against 1.4x the records).
    std::vector<uint32_t> tri_of;
    const bool split = o.presplit > 1;
    if (split) {
        // The mean over finite boxes: one triangle with a NaN or infinite
        // coordinate (never split) must not switch splitting off for all.
        double limit = 0.0;
        uint32_t finite = 0;
        for (const Prim &p : prims) {
            const double a = p.box.half_area();
            if (std::isfinite(a)) {
                limit += a;
                ++finite;
            }
        }
        limit = finite ? 0.25 * limit / static_cast<double>(finite) : INFINITY;
        std::vector<Prim> parts;
        parts.reserve(n + n / 2);
        tri_of.reserve(n + n / 2);
        for (uint32_t i = 0; i < n; ++i) {
            const double area = prims[i].box.half_area();
            if (!(area > limit) || !std::isfinite(area)) {
                parts.push_back(prims[i]);
                tri_of.push_back(i);
                continue;
            }
            Poly p;
            for (int v = 0; v < 3; ++v) {
                double q[3];
                for (int k = 0; k < 3; ++k) q[k] = tri_coord(tri, i, v, k);
                p.push(q);
            }
            split_piece(p, prims[i].box, o.presplit, limit, i, parts, tri_of);
        }
        if (parts.size() > kMaxLeafOffset) {
            err = "too many triangle parts for one scene; lower presplit";
            return false;
        }
        for (size_t j = 0; j < parts.size(); ++j) parts[j].id = static_cast<uint32_t>(j);
        prims.swap(parts);
    }
    const uint32_t m = static_cast<uint32_t>(prims.size());
    {
        SahBuilder b(prims, o.sah_bins, o.max_leaf);
        b.build(tree.nodes);
    }
    if (!split) {
        tree.order.resize(m);
        for (uint32_t i = 0; i < m; ++i) tree.order[i] = prims[i].id;
        return true;
    }
    // Leaf positions as triangles, leaves in depth-first order (left child
    // first); parts of one triangle that share a leaf become one record
    // whose box is the union of theirs.
    std::vector<uint32_t> order;
    std::vector<float> piece;
    order.reserve(m);
    piece.reserve(static_cast<size_t>(m) * 6);
    std::vector<uint32_t> st = {0};
    while (!st.empty()) {
        Node &nd = tree.nodes[st.back()];
        st.pop_back();
        if (nd.count == 0) {
            st.push_back(nd.right);
            st.push_back(nd.left);
            continue;
        }
        const uint32_t first = static_cast<uint32_t>(order.size());
        for (uint32_t i = 0; i < nd.count; ++i) {
            const Prim &q = prims[nd.first + i];
            const uint32_t t = tri_of[q.id];
            uint32_t j = first;
            while (j < order.size() && order[j] != t) ++j;
            if (j == order.size()) {
                order.push_back(t);
                for (int k = 0; k < 3; ++k) piece.push_back(q.box.lo[k]);
                for (int k = 0; k < 3; ++k) piece.push_back(q.box.hi[k]);
            } else {
                for (int k = 0; k < 3; ++k) {
                    piece[j * 6 + k] = std::fmin(piece[j * 6 + k], q.box.lo[k]);
                    piece[j * 6 + 3 + k] = std::fmax(piece[j * 6 + 3 + k], q.box.hi[k]);
                }
            }
        }
        nd.first = first;
        nd.count = static_cast<uint32_t>(order.size()) - first;
    }
    tree.order.swap(order);
    tree.piece.swap(piece);
    return true;
}

void pad_and_refit(BinTree &tree, const float *tri, float pad_rel) {
    if (tree.nodes.empty()) return;
    // Scene scale from the root box (a reduction by fmin/fmax, so the
    // slicing does not change it).
    const size_t nord = tree.order.size();
    std::vector<Box> roots(host_threads());
    parallel_slices(nord, 1u << 16, [&](unsigned j, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i)
            for (int v = 0; v < 3; ++v) roots[j].grow(tri + static_cast<size_t>(tree.order[i]) * 9 + v * 3);
    });
    Box root;
    for (const Box &b : roots) root.grow(b);
    double scale = 0.0;
    for (int k = 0; k < 3; ++k) {
        scale = std::max(scale, std::fabs(static_cast<double>(root.lo[k])));
        scale = std::max(scale, std::fabs(static_cast<double>(root.hi[k])));
        scale = std::max(scale, static_cast<double>(root.hi[k]) - root.lo[k]);
    }
    const double delta = std::max(scale, 1e-30) * pad_rel;
    // Leaves, each from its own triangles (or pieces), in parallel.
    const size_t nn = tree.nodes.size();
    parallel_slices(nn, 1u << 14, [&](unsigned, size_t lo_i, size_t hi_i) {
        for (size_t idx = lo_i; idx < hi_i; ++idx) {
            BinTree::Node &nd = tree.nodes[idx];
            if (nd.count == 0) continue;
            Box b;
            for (uint32_t i = 0; i < nd.count; ++i) {
                if (!tree.piece.empty()) {
                    b.grow(&tree.piece[static_cast<size_t>(nd.first + i) * 6]);
                    b.grow(&tree.piece[static_cast<size_t>(nd.first + i) * 6 + 3]);
                    continue;
                }
                for (int v = 0; v < 3; ++v)
                    b.grow(tri + static_cast<size_t>(tree.order[nd.first + i]) * 9 + v * 3);
            }
            for (int k = 0; k < 3; ++k) {
                double lo = static_cast<double>(b.lo[k]) - delta;
                double hi = static_cast<double>(b.hi[k]) + delta;
                float flo = static_cast<float>(lo), fhi = static_cast<float>(hi);
                if (static_cast<double>(flo) > lo) flo = std::nextafter(flo, -INFINITY);
                if (static_cast<double>(fhi) < hi) fhi = std::nextafter(fhi, INFINITY);
                nd.lo[k] = flo;
                nd.hi[k] = fhi;
            }
        }
    });
    // Interior nodes in post-order (children have larger indices than
    // parents).
    for (size_t idx = nn; idx-- > 0;) {
        BinTree::Node &nd = tree.nodes[idx];
        if (nd.count > 0) continue;
        const BinTree::Node &a = tree.nodes[nd.left], &c = tree.nodes[nd.right];
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = std::fmin(a.lo[k], c.lo[k]);
            nd.hi[k] = std::fmax(a.hi[k], c.hi[k]);
        }
    }
}

namespace {

// Quantises child boxes of one wide node.  Returns false if a box cannot be
// represented (never for finite input).
template <int W, typename NodeT>
void quantise(NodeT &nd, const std::vector<const BinTree::Node *> &kids) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const BinTree::Node *c : kids)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], c->lo[k]);
            hi[k] = std::fmax(hi[k], c->hi[k]);
        }
    for (int k = 0; k < 3; ++k) {
        double ext = static_cast<double>(hi[k]) - lo[k];
        int e = -126;
        if (ext > 0.0) {
            // Boxes of finite coordinates need e <= 122; the clamps keep the
            // loops short whatever ext is.
            const double l = std::ceil(std::log2(ext / 255.0));
            e = l > 127.0 ? 127 : l < -126.0 ? -126 : static_cast<int>(l);
            while (e < 127 && std::ldexp(255.0, e) < ext) ++e;
            while (e > -126 && std::ldexp(255.0, e - 1) >= ext) --e;
        }
        nd.origin[k] = lo[k];
        nd.ex[k] = static_cast<uint8_t>(e + 127);
        const double s = std::ldexp(1.0, e);
        for (int c = 0; c < W; ++c) {
            if (c < static_cast<int>(kids.size())) {
                double ql = std::floor((static_cast<double>(kids[c]->lo[k]) - lo[k]) / s);
                double qh = std::ceil((static_cast<double>(kids[c]->hi[k]) - lo[k]) / s);
                ql = std::min(std::max(ql, 0.0), 255.0);
                qh = std::min(std::max(qh, 0.0), 255.0);
                nd.qlo[k][c] = static_cast<uint8_t>(ql);
                nd.qhi[k][c] = static_cast<uint8_t>(qh);
            } else {
                nd.qlo[k][c] = 255;
                nd.qhi[k][c] = 0;
            }
        }
    }
    nd.nchild = static_cast<uint8_t>(kids.size());
}

template <int W, typename NodeT>
void emit_wide(const BinTree &tree, HostBvh &out) {
    // Collapse: repeatedly open the interior child with the largest surface
    // area until W children (ties: lowest slot).
    auto collapse = [&](uint32_t b) {
        std::vector<uint32_t> kids = {tree.nodes[b].left, tree.nodes[b].right};
        while (static_cast<int>(kids.size()) < W) {
            int pick = -1;
            double best = -1.0;
            for (size_t i = 0; i < kids.size(); ++i) {
                const BinTree::Node &c = tree.nodes[kids[i]];
                if (c.count > 0) continue;
                Box bx;
                std::memcpy(bx.lo, c.lo, sizeof bx.lo);
                std::memcpy(bx.hi, c.hi, sizeof bx.hi);
                double a = bx.half_area();
                if (a > best) {
                    best = a;
                    pick = static_cast<int>(i);
                }
            }
            if (pick < 0) break;
            uint32_t c = kids[pick];
            kids[pick] = tree.nodes[c].left;
            kids.insert(kids.begin() + pick + 1, tree.nodes[c].right);
        }
        return kids;
    };
    // Structure first (the wide nodes' numbering, children and the binary
    // nodes they collapse), then the quantisation of each wide node, which
    // depends on that node alone, in parallel.
    std::vector<uint32_t> bin_of;               // wide node -> binary node
    std::vector<uint32_t> kids_of;              // W binary children per wide node (kNoKid: none)
    std::vector<uint32_t> refs;                 // W child refs per wide node
    constexpr uint32_t kNoKid = 0xFFFFFFFFu;
    bin_of.push_back(0);
    kids_of.resize(W, kNoKid);
    refs.resize(W, kRefNone);
    std::vector<uint32_t> stack = {0};
    while (!stack.empty()) {
        uint32_t wi = stack.back();
        stack.pop_back();
        std::vector<uint32_t> kids = collapse(bin_of[wi]);
        std::vector<uint32_t> fresh;
        for (int c = 0; c < static_cast<int>(kids.size()); ++c) {
            kids_of[static_cast<size_t>(wi) * W + c] = kids[c];
            const BinTree::Node &k = tree.nodes[kids[c]];
            if (k.count > 0) {
                refs[static_cast<size_t>(wi) * W + c] = make_leaf(k.first, k.count);
            } else {
                uint32_t idx = static_cast<uint32_t>(bin_of.size());
                bin_of.push_back(kids[c]);
                kids_of.resize(kids_of.size() + W, kNoKid);
                refs.resize(refs.size() + W, kRefNone);
                refs[static_cast<size_t>(wi) * W + c] = idx;
                fresh.push_back(idx);
            }
        }
        for (size_t i = fresh.size(); i-- > 0;) stack.push_back(fresh[i]);
    }
    std::vector<NodeT> nodes(bin_of.size());
    parallel_slices(nodes.size(), 1u << 12, [&](unsigned, size_t lo, size_t hi) {
        std::vector<const BinTree::Node *> kp;
        for (size_t wi = lo; wi < hi; ++wi) {
            kp.clear();
            for (int c = 0; c < W && kids_of[wi * W + c] != kNoKid; ++c) kp.push_back(&tree.nodes[kids_of[wi * W + c]]);
            NodeT nd{};
            quantise<W>(nd, kp);
            for (int c = 0; c < W; ++c) nd.child[c] = refs[wi * W + c];
            nodes[wi] = nd;
        }
    });
    out.num_nodes = static_cast<uint32_t>(nodes.size());
    out.node_bytes = sizeof(NodeT);
    out.nodes.resize(nodes.size() * sizeof(NodeT));
    std::memcpy(out.nodes.data(), nodes.data(), out.nodes.size());
    out.root_ref = 0;
    // Depth and stack bound: along any path, an interior node pushes at most
    // (children - 1) entries.
    struct Item {
        uint32_t node, depth, stack;
    };
    std::vector<Item> st = {{0, 1, 0}};
    out.max_depth = 0;
    out.max_stack = 0;
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        const NodeT &nd = nodes[it.node];
        uint32_t s = it.stack + (nd.nchild > 0 ? nd.nchild - 1u : 0u);
        out.max_depth = std::max(out.max_depth, it.depth);
        out.max_stack = std::max(out.max_stack, s);
        for (int c = 0; c < W; ++c)
            if (nd.child[c] != kRefNone && !ref_is_leaf(nd.child[c]))
                st.push_back({nd.child[c], it.depth + 1, s});
    }
}

void emit_binary(const BinTree &tree, HostBvh &out) {
    // Interior nodes numbered in preorder.
    std::vector<uint32_t> index(tree.nodes.size(), kRefNone);
    std::vector<uint32_t> pre;
    std::vector<uint32_t> st = {0};
    while (!st.empty()) {
        uint32_t b = st.back();
        st.pop_back();
        if (tree.nodes[b].count > 0) continue;
        index[b] = static_cast<uint32_t>(pre.size());
        pre.push_back(b);
        st.push_back(tree.nodes[b].right);
        st.push_back(tree.nodes[b].left);
    }
    auto ref_of = [&](uint32_t b) {
        const BinTree::Node &nd = tree.nodes[b];
        return nd.count > 0 ? make_leaf(nd.first, nd.count) : index[b];
    };
    std::vector<NodeB2> nodes(pre.size());
    for (size_t i = 0; i < pre.size(); ++i) {
        const BinTree::Node &nd = tree.nodes[pre[i]];
        const BinTree::Node &a = tree.nodes[nd.left], &c = tree.nodes[nd.right];
        NodeB2 &o = nodes[i];
        std::memset(&o, 0, sizeof o);
        for (int k = 0; k < 3; ++k) {
            o.lo0[k] = a.lo[k];
            o.hi0[k] = a.hi[k];
            o.lo1[k] = c.lo[k];
            o.hi1[k] = c.hi[k];
        }
        o.child[0] = ref_of(nd.left);
        o.child[1] = ref_of(nd.right);
    }
    out.num_nodes = static_cast<uint32_t>(nodes.size());
    out.node_bytes = sizeof(NodeB2);
    out.nodes.resize(nodes.size() * sizeof(NodeB2));
    std::memcpy(out.nodes.data(), nodes.data(), out.nodes.size());
    out.root_ref = 0;
    struct Item {
        uint32_t node, depth;
    };
    std::vector<Item> s2 = {{0, 1}};
    out.max_depth = 0;
    while (!s2.empty()) {
        Item it = s2.back();
        s2.pop_back();
        out.max_depth = std::max(out.max_depth, it.depth);
        for (int c = 0; c < 2; ++c)
            if (!ref_is_leaf(nodes[it.node].child[c])) s2.push_back({nodes[it.node].child[c], it.depth + 1});
    }
    out.max_stack = out.max_depth;
}

}  // namespace

bool emit_layout(const BinTree &tree, const float *tri, uint32_t n, const rt_build_opts &o, HostBvh &out,
                 std::string &err) {
    out.width = o.node_width;
    out.num_tris = n;
    out.nodes.clear();
    out.num_nodes = 0;
    out.node_bytes = o.node_width == 8 ? 128u : 64u;
    out.max_depth = out.max_stack = 0;
    out.num_leaves = 0;
    // One record per leaf position: n, or more when triangles were split.
    const uint32_t nrec = static_cast<uint32_t>(tree.order.size());
    out.tris.resize(nrec);
    for (uint32_t r = 0; r < nrec; ++r) {
        uint32_t id = tree.order[r];
        const float *t = tri + static_cast<size_t>(id) * 9;
        TriRec &rec = out.tris[r];
        for (int k = 0; k < 3; ++k) {
            rec.p0[k] = t[k];
            rec.e1[k] = t[3 + k] - t[k];
            rec.e2[k] = t[6 + k] - t[k];
        }
        rec.id = id;
        rec.pad1 = rec.pad2 = 0;
    }
    for (const BinTree::Node &nd : tree.nodes)
        if (nd.count > 0) out.num_leaves++;
    if (n == 0) {
        out.root_ref = kRefNone;
        for (int k = 0; k < 3; ++k) out.lo[k] = out.hi[k] = 0.f;
        return true;
    }
    for (int k = 0; k < 3; ++k) {
        out.lo[k] = tree.nodes[0].lo[k];
        out.hi[k] = tree.nodes[0].hi[k];
    }
    if (tree.nodes[0].count > 0) {  // the whole scene is one leaf
        out.root_ref = make_leaf(tree.nodes[0].first, tree.nodes[0].count);
        return true;
    }
    if (o.node_width == 2)
        emit_binary(tree, out);
    else if (o.node_width == 4)
        emit_wide<4, NodeB4>(tree, out);
    else if (o.node_width == 8)
        emit_wide<8, NodeB8>(tree, out);
    else {
        err = "bad node width";
        return false;
    }
    return true;
}

bool build_host(const float *tri, uint32_t n, const rt_build_opts &o, HostBvh &out, std::string &err) {
    auto t0 = std::chrono::steady_clock::now();
    // A box with an infinite side has no quantised form.  (NaN coordinates
    // are allowed: box arithmetic ignores them, and the triangle test
    // rejects the NaN t they give.)
    for (size_t i = 0; i < static_cast<size_t>(n) * 9; ++i)
        if (std::isinf(tri[i])) {
            err = "triangle " + std::to_string(i / 9) + " has an infinite coordinate";
            return false;
        }
    BinTree tree;
    if (!build_sah_tree(tri, n, o, tree, err)) return false;
    pad_and_refit(tree, tri, o.pad_rel);
    if (!emit_layout(tree, tri, n, o, out, err)) return false;
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

}  // namespace rtx

extern "C" void rt_build_opts_default(rt_build_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->builder = RT_BUILD_SAH_CPU;
    o->node_width = RT_NODE_BVH4;
    o->max_leaf = 4;
    o->sah_bins = 32;
    o->pad_rel = 1.0f / 65536.0f;
    o->keep_source = 1;
}
