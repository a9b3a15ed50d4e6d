// GPU LBVH builder.
#include "lbvh.hpp"

namespace rtx {

struct LbvhWorkspace {};

LbvhWorkspace *lbvh_workspace_new() { return new LbvhWorkspace; }
void lbvh_workspace_free(LbvhWorkspace *ws) { delete ws; }

int lbvh_build_device(LbvhWorkspace *, const float *, uint32_t, const rt_build_opts &, hipStream_t,
                      LbvhResult &, std::string &why) {
    why = "GPU LBVH builder not built yet";
    return RT_ERR_UNSUPPORTED;
}

}  // namespace rtx
