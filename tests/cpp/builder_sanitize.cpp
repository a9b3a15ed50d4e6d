// The host BVH builders under AddressSanitizer and UndefinedBehaviorSanitizer
// (float-to-int overflow included): random soups of small and of long
// triangles, one with NaN coordinates and one with point triangles, at every
// node width and reference-splitting budget; then a scene with an infinite
// coordinate, which every builder refuses.  Prints one line per build and
// "ok" at the end; a sanitizer report aborts the run.
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "bvh_build.hpp"

int main() {
    std::mt19937 g(7);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::uniform_real_distribution<float> ud(-10.f, 10.f);
    const uint32_t n = 3000;
    for (int scene = 0; scene < 4; ++scene) {
        std::vector<float> t(static_cast<size_t>(n) * 9);
        const float edge = scene == 0 ? 0.05f : 3.f;  // small triangles, then long ones
        for (uint32_t i = 0; i < n; ++i) {
            const float c[3] = {ud(g), ud(g), ud(g)};
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) t[i * 9 + v * 3 + k] = c[k] + edge * nd(g);
        }
        if (scene == 2) {
            t[5] = NAN;
            t[77] = NAN;
        }
        if (scene == 3)  // 50 triangles collapsed to points
            for (uint32_t i = 0; i < 50; ++i)
                for (int v = 1; v < 3; ++v)
                    for (int k = 0; k < 3; ++k) t[i * 9 + v * 3 + k] = t[i * 9 + k];
        for (uint32_t w : {2u, 4u, 8u})
            for (uint32_t ps : {1u, 4u, 16u, 64u}) {
                rt_build_opts o;
                rt_build_opts_default(&o);
                o.node_width = w;
                o.presplit = ps;
                rt_build_opts r;
                std::string err;
                if (!rtx::resolve_opts(&o, r, err)) {
                    std::printf("options refused: %s\n", err.c_str());
                    return 1;
                }
                rtx::HostBvh h;
                if (!rtx::build_host(t.data(), n, r, h, err)) {
                    std::printf("build failed: %s\n", err.c_str());
                    return 1;
                }
                std::printf("scene %d width %u presplit %u: %u nodes, %zu records\n", scene, w, ps, h.num_nodes,
                            h.tris.size());
            }
    }
    std::vector<float> bad(9, 1.f);
    bad[4] = INFINITY;
    rt_build_opts o;
    rt_build_opts_default(&o);
    rt_build_opts r;
    std::string err;
    rtx::resolve_opts(&o, r, err);
    rtx::HostBvh h;
    const bool built = rtx::build_host(bad.data(), 1, r, h, err);
    if (built || err.find("infinite") == std::string::npos) {
        std::printf("an infinite coordinate was not refused\n");
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
