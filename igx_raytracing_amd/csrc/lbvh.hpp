// GPU LBVH builder interface (implemented in lbvh.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "rt/rt_c.h"

namespace rtx {

struct LbvhWorkspace;

struct LbvhResult {
    void *d_nodes = nullptr;
    void *d_tris = nullptr;
    float *d_src = nullptr;
    uint32_t root_ref = 0xFFFFFFFFu;
    uint32_t num_nodes = 0;
    uint32_t max_depth = 0;
    uint32_t max_stack = 0;
    uint32_t num_leaves = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double build_ms = 0.0;
};

LbvhWorkspace *lbvh_workspace_new();
void lbvh_workspace_free(LbvhWorkspace *ws);

// Builds a scene from device-resident input-order triangles d_tri[n][9].
// Returns an RT_* status; on failure `why` says what failed.
int lbvh_build_device(LbvhWorkspace *ws, const float *d_tri, uint32_t n, const rt_build_opts &o,
                      hipStream_t stream, LbvhResult &res, std::string &why);

// Steps shared by every LBVH build (bounds, Morton codes, stable sort,
// radix tree, boxes), copied to caller buffers; see rt_lbvh_binary.
int lbvh_binary(LbvhWorkspace *ws, const float *d_tri, uint32_t n, hipStream_t stream, uint32_t *morton,
                uint32_t *order, uint32_t *left, uint32_t *right, float *boxes, std::string &why);

}  // namespace rtx
