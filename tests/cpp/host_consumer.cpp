// A C++ consumer of the boundary: includes the C ABI and the C++ task
// mirror, links librt_hip.so, and runs the host-only builder (no device).
#include "rt/ray_query.hpp"
#include "rt/rt_workload.h"
#include <cstdio>
int main() {
    // Host-only part: builds and links against librt_hip.so; no device calls.
    std::vector<float> tris(4096 * 9);
    rtw_scene_mixed4k(0x5EED0000ull, tris.data());
    rt_bvh b = nullptr;
    rt_build_opts o; rt_build_opts_default(&o);
    igx::rt::hip::check(rt_bvh_build(tris.data(), 4096, &o, &b));
    rt_scene_info info; rt_bvh_info(b, &info);
    std::printf("nodes %u depth %u\n", info.num_nodes, info.max_depth);
    rt_bvh_free(b);
    return 0;
}
