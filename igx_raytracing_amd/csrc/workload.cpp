// Seeded synthetic workloads (scenes, rays).  Host code.
//
// The reference has no generator of its own (its one scene is hand-placed in
// test/scene/niels_scene.cpp:10-56); these generators define the benchmark
// inputs.  All arithmetic is double precision, rounded once to float at the
// store, so the same seed gives the same bits on any x86-64 host with the
// same libm.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <map>
#include <array>
#include <algorithm>
#include <thread>

#include "rt/rt_workload.h"

namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t seed, uint64_t stream) {
        inc = (stream << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
    }
    // [0, 1) with 32 bits of resolution.
    double uniform() { return next() * (1.0 / 4294967296.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    // Box-Muller, one normal per two draws; u1 in (0, 1].
    double normal(double mu, double sigma) {
        double u1 = (static_cast<double>(next()) + 1.0) * (1.0 / 4294967296.0);
        double u2 = uniform();
        double r = std::sqrt(-2.0 * std::log(u1));
        return mu + sigma * r * std::cos(6.283185307179586476925 * u2);
    }
};

// Runs fn(lo, hi) over [0, n) on the host's cores.  Every item draws from
// its own PCG stream, so the split never changes a bit of the output.
template <typename Fn>
void parallel_for(uint64_t n, Fn fn) {
    unsigned hw = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (n < 4096 || hw == 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    const uint64_t per = (n + hw - 1) / hw;
    for (unsigned t = 0; t < hw; ++t) {
        const uint64_t lo = t * per, hi = std::min<uint64_t>(n, lo + per);
        if (lo >= hi) break;
        th.emplace_back(fn, lo, hi);
    }
    for (auto &x : th) x.join();
}

inline void put_tri(float *dst, const double a[3], const double b[3], const double c[3]) {
    for (int k = 0; k < 3; ++k) dst[k] = static_cast<float>(a[k]);
    for (int k = 0; k < 3; ++k) dst[3 + k] = static_cast<float>(b[k]);
    for (int k = 0; k < 3; ++k) dst[6 + k] = static_cast<float>(c[k]);
}

}  // namespace

extern "C" uint32_t rtw_pcg32_nth(uint64_t seed, uint64_t stream, uint64_t skip) {
    Pcg32 g(seed, stream);
    for (uint64_t i = 0; i < skip; ++i) g.next();
    return g.next();
}

extern "C" uint32_t rtw_scene_clusters(uint64_t seed, uint32_t nclusters, uint32_t tpc, float range,
                                       float sigma, float edge_sigma, float *out) {
    auto one = [&](uint32_t c) {
        Pcg32 g(seed, c);
        double ctr[3];
        for (int k = 0; k < 3; ++k) ctr[k] = g.uniform(-range, range);
        for (uint32_t i = 0; i < tpc; ++i) {
            double m[3], p[3][3];
            for (int k = 0; k < 3; ++k) m[k] = ctr[k] + g.normal(0.0, sigma);
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) p[j][k] = m[k] + g.normal(0.0, edge_sigma);
            put_tri(out + (static_cast<size_t>(c) * tpc + i) * 9, p[0], p[1], p[2]);
        }
    };
    if (static_cast<uint64_t>(nclusters) * tpc < 65536)
        for (uint32_t c = 0; c < nclusters; ++c) one(c);
    else
        parallel_for(nclusters * 4096ull, [&](uint64_t lo, uint64_t hi) {
            // clusters [lo/4096, hi/4096): whole clusters per thread
            for (uint64_t c = (lo + 4095) / 4096; c < (hi + 4095) / 4096 && c < nclusters; ++c)
                one(static_cast<uint32_t>(c));
        });
    return nclusters * tpc;
}

extern "C" void rtw_rays_uniform(uint64_t seed, const float lo[3], const float hi[3], uint64_t first,
                                 uint32_t n, float tmax_lo, float tmax_hi, rt_ray *rays) {
    parallel_for(n, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i) {
            Pcg32 g(seed, 0x40000000ull + first + i);
            rt_ray r;
            for (int k = 0; k < 3; ++k) r.org[k] = static_cast<float>(g.uniform(lo[k], hi[k]));
            double z = 1.0 - 2.0 * g.uniform();
            double phi = 6.283185307179586476925 * g.uniform();
            double sn = std::sqrt(std::fmax(0.0, 1.0 - z * z));
            r.dir[0] = static_cast<float>(sn * std::cos(phi));
            r.dir[1] = static_cast<float>(sn * std::sin(phi));
            r.dir[2] = static_cast<float>(z);
            r.tmax = tmax_hi > 0.f ? static_cast<float>(g.uniform(tmax_lo, tmax_hi)) : RT_NO_HIT_T;
            r.prev_id = RT_MISS;
            rays[i] = r;
        }
    });
}

extern "C" uint32_t rtw_scene_mixed4k(uint64_t seed, float *out) {
    uint32_t nt = 0;
    // 64 axis-aligned cubes, 12 triangles each, outward counter-clockwise.
    {
        Pcg32 g(seed, 1);
        // Corner index = x | y<<1 | z<<2; each face as two triangles.
        static const int faces[6][4] = {
            {0, 2, 6, 4}, {1, 5, 7, 3},  // -x, +x
            {0, 4, 5, 1}, {2, 3, 7, 6},  // -y, +y
            {0, 1, 3, 2}, {4, 6, 7, 5},  // -z, +z
        };
        for (int c = 0; c < 64; ++c) {
            double ctr[3];
            for (int k = 0; k < 3; ++k) ctr[k] = g.uniform(-10.0, 10.0);
            double half = 0.5 * g.uniform(0.5, 2.0);
            double corner[8][3];
            for (int v = 0; v < 8; ++v)
                for (int k = 0; k < 3; ++k) corner[v][k] = ctr[k] + (((v >> k) & 1) ? half : -half);
            for (int f = 0; f < 6; ++f) {
                const int *q = faces[f];
                put_tri(out + 9 * nt++, corner[q[0]], corner[q[1]], corner[q[2]]);
                put_tri(out + 9 * nt++, corner[q[0]], corner[q[2]], corner[q[3]]);
            }
        }
    }
    // Ground grid: 32x32 quads over [-16,16]^2 at y = 0, heights N(0, 0.05).
    {
        Pcg32 g(seed, 2);
        double h[33][33];
        for (int j = 0; j < 33; ++j)
            for (int i = 0; i < 33; ++i) h[j][i] = g.normal(0.0, 0.05);
        auto vtx = [&](int i, int j, double p[3]) {
            p[0] = -16.0 + i;
            p[1] = h[j][i];
            p[2] = -16.0 + j;
        };
        for (int j = 0; j < 32; ++j)
            for (int i = 0; i < 32; ++i) {
                double a[3], b[3], c[3], d[3];
                vtx(i, j, a);
                vtx(i + 1, j, b);
                vtx(i + 1, j + 1, c);
                vtx(i, j + 1, d);
                put_tri(out + 9 * nt++, a, d, c);
                put_tri(out + 9 * nt++, a, c, b);
            }
    }
    // 1,280 loose triangles: centroid U[-10,10]^3, edges N(0,0.3)^3.
    {
        Pcg32 g(seed, 3);
        for (int t = 0; t < 1280; ++t) {
            double c[3], e1[3], e2[3], p0[3], p1[3], p2[3];
            for (int k = 0; k < 3; ++k) c[k] = g.uniform(-10.0, 10.0);
            for (int k = 0; k < 3; ++k) e1[k] = g.normal(0.0, 0.3);
            for (int k = 0; k < 3; ++k) e2[k] = g.normal(0.0, 0.3);
            for (int k = 0; k < 3; ++k) {
                p0[k] = c[k] - (e1[k] + e2[k]) / 3.0;
                p1[k] = p0[k] + e1[k];
                p2[k] = p0[k] + e2[k];
            }
            put_tri(out + 9 * nt++, p0, p1, p2);
        }
    }
    return nt;
}

extern "C" void rtw_rays_pinhole(uint32_t w, uint32_t hgt, const float eye[3], const float target[3],
                                 float vfov_deg, rt_ray *rays) {
    double e[3] = {eye[0], eye[1], eye[2]};
    double f[3] = {target[0] - e[0], target[1] - e[1], target[2] - e[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double &x : f) x /= fl;
    const double up[3] = {0.0, 1.0, 0.0};
    // right = up x f (image x grows to the right), v = f x right.
    double r[3] = {up[1] * f[2] - up[2] * f[1], up[2] * f[0] - up[0] * f[2], up[0] * f[1] - up[1] * f[0]};
    double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double &x : r) x /= rl;
    double u[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
    double th = std::tan(0.5 * vfov_deg * 3.14159265358979323846 / 180.0);
    double aspect = static_cast<double>(w) / static_cast<double>(hgt);
    for (uint32_t y = 0; y < hgt; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            double sx = (2.0 * (x + 0.5) / w - 1.0) * th * aspect;
            double sy = (1.0 - 2.0 * (y + 0.5) / hgt) * th;
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = f[k] + sx * r[k] + sy * u[k];
            double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            rt_ray ray;
            for (int k = 0; k < 3; ++k) {
                ray.org[k] = eye[k];
                ray.dir[k] = static_cast<float>(d[k] / dl);
            }
            ray.tmax = RT_NO_HIT_T;
            ray.prev_id = RT_MISS;
            rays[static_cast<size_t>(y) * w + x] = ray;
        }
}

extern "C" void rtw_bounds(const float *t, uint32_t n, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = INFINITY;
        hi[k] = -INFINITY;
    }
    for (size_t i = 0; i < static_cast<size_t>(n) * 3; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], t[3 * i + k]);
            hi[k] = std::fmax(hi[k], t[3 * i + k]);
        }
}

extern "C" uint32_t rtw_scene_icosphere(uint64_t seed, uint32_t subdiv, float radius, float disp,
                                        float *out) {
    // Unit icosahedron, then 1->4 midpoint subdivision with shared midpoints.
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<std::array<double, 3>> v = {
        {-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
        {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto &p : v) {
        double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (double &x : p) x /= l;
    }
    std::vector<std::array<uint32_t, 3>> f = {
        {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (uint32_t s = 0; s < subdiv; ++s) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? (static_cast<uint64_t>(a) << 32 | b) : (static_cast<uint64_t>(b) << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            std::array<double, 3> m;
            for (int k = 0; k < 3; ++k) m[k] = 0.5 * (v[a][k] + v[b][k]);
            double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            for (double &x : m) x /= l;
            v.push_back(m);
            uint32_t id = static_cast<uint32_t>(v.size() - 1);
            mid.emplace(key, id);
            return id;
        };
        std::vector<std::array<uint32_t, 3>> nf;
        nf.reserve(f.size() * 4);
        for (auto &tri : f) {
            uint32_t a = midpoint(tri[0], tri[1]), b = midpoint(tri[1], tri[2]), c = midpoint(tri[2], tri[0]);
            nf.push_back({tri[0], a, c});
            nf.push_back({tri[1], b, a});
            nf.push_back({tri[2], c, b});
            nf.push_back({a, b, c});
        }
        f.swap(nf);
    }
    Pcg32 g(seed, 7);
    std::vector<double> scale(v.size());
    for (size_t i = 0; i < v.size(); ++i) scale[i] = radius + disp * g.normal(0.0, 1.0);
    for (size_t i = 0; i < f.size(); ++i) {
        double p[3][3];
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) p[j][k] = v[f[i][j]][k] * scale[f[i][j]];
        put_tri(out + 9 * i, p[0], p[1], p[2]);
    }
    return static_cast<uint32_t>(f.size());
}
