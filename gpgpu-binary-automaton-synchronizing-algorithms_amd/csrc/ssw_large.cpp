// ssw_large.cpp -- host driver of the device-wide level loop (placeholder until
// the device-wide kernels land).
#include "ssw_host.h"

namespace ssw {

void LargeEngine::release() {
    if (state.p) hipFree(state.p);
    if (pool.p) hipFree(pool.p);
    state = DevBuf();
    pool = DevBuf();
    if (graph) hipGraphExecDestroy(graph);
    graph = nullptr;
}

int LargeEngine::solve_indices(ssw_context &ctx, const DeviceBatch &, const uint8_t *,
                               const std::vector<uint32_t> &idx) {
    return ctx.fail(SSW_ERR_INTERNAL, "%zu automata need the device-wide path, not built yet", idx.size());
}

int LargeEngine::solve_all(ssw_context &ctx, const DeviceBatch &, const uint8_t *, int64_t *) {
    return ctx.fail(SSW_ERR_INTERNAL, "device-wide path not built yet");
}

}  // namespace ssw
