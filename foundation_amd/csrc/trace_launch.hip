// trace_launch.hip — instantiations of k_trace and the launcher that picks one.
#include "trace_coop.hpp"
#include "trace_kernel.hpp"

namespace fnd {

#define FND_TRACE_CONFIGS(X) \
    X(64, 16)                \
    X(64, 32)                \
    X(128, 8)                \
    X(128, 12)               \
    X(128, 16)               \
    X(128, 32)               \
    X(256, 8)                \
    X(256, 12)               \
    X(256, 16)

bool trace_config_supported(uint32_t block, uint32_t stack_lds)
{
#define FND_CHECK(B, S) \
    if (block == B && stack_lds == S) return true;
    FND_TRACE_CONFIGS(FND_CHECK)
#undef FND_CHECK
    return false;
}

template <int B, int S>
static const void* kernel_cfg(const TraceLaunch& l)
{
    if (l.coop) {
        if (l.mode != kModeMixedInternal) return nullptr;
        if (l.multi) return reinterpret_cast<const void*>(&k_trace_coop<true, true, S, B>);
        if (l.two_level) return reinterpret_cast<const void*>(&k_trace_coop<true, false, S, B>);
        return reinterpret_cast<const void*>(&k_trace_coop<false, false, S, B>);
    }
    if (l.multi) {
        if (l.mode == kModeClosestInternal) return reinterpret_cast<const void*>(&k_trace_render<kModeClosestInternal, true, true, false, S, B>);
        if (l.mode == kModeAnyInternal) return reinterpret_cast<const void*>(&k_trace_render<kModeAnyInternal, true, true, false, S, B>);
        if (l.mode == kModeMixedInternal) return reinterpret_cast<const void*>(&k_trace_render<kModeMixedInternal, true, true, false, S, B>);
        return nullptr;
    }
    if (l.mode == kModeMixedInternal) {
        // a render of one scene: its constants are kernel arguments
        if (l.two_level) return reinterpret_cast<const void*>(&k_trace_render<kModeMixedInternal, true, false, false, S, B>);
        return reinterpret_cast<const void*>(&k_trace_render<kModeMixedInternal, false, false, false, S, B>);
    }
    if (l.count) return nullptr;  // counting variants exist for the default configuration only
    if (l.two_level) {
        if (l.mode == kModeClosest) return reinterpret_cast<const void*>(&k_trace<kModeClosest, true, false, false, S, B>);
        if (l.mode == kModeAny) return reinterpret_cast<const void*>(&k_trace<kModeAny, true, false, false, S, B>);
        return nullptr;
    }
    if (l.mode == kModeClosest) return reinterpret_cast<const void*>(&k_trace<kModeClosest, false, false, false, S, B>);
    if (l.mode == kModeAny) return reinterpret_cast<const void*>(&k_trace<kModeAny, false, false, false, S, B>);
    return nullptr;
}

static const void* kernel_count(const TraceLaunch& l)
{
    if (l.block != 128 || l.stack_lds != 16) return nullptr;
    if (l.two_level)
        return l.mode == kModeClosest ? reinterpret_cast<const void*>(&k_trace<kModeClosest, true, false, true, 16, 128>)
                                      : reinterpret_cast<const void*>(&k_trace<kModeAny, true, false, true, 16, 128>);
    return l.mode == kModeClosest ? reinterpret_cast<const void*>(&k_trace<kModeClosest, false, false, true, 16, 128>)
                                  : reinterpret_cast<const void*>(&k_trace<kModeAny, false, false, true, 16, 128>);
}

static const void* kernel_for(const TraceLaunch& l)
{
    if (l.count && !l.multi) return kernel_count(l);
#define FND_PICK(B, S) \
    if (l.block == B && l.stack_lds == S) return kernel_cfg<B, S>(l);
    FND_TRACE_CONFIGS(FND_PICK)
#undef FND_PICK
    return nullptr;
}

uint32_t trace_blocks_per_cu(const TraceLaunch& l)
{
    const void* k = kernel_for(l);
    int nb = 0;
    if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, int(l.block), 0) != hipSuccess) return 0;
    return uint32_t(nb);
}

hipError_t launch_trace(const TraceLaunch& l, const TraceArgs& args, hipStream_t stream)
{
    const void* k = kernel_for(l);
    if (!k) return hipErrorInvalidValue;
    TraceArgs a = args;
    void* params[] = {&a};
    return hipLaunchKernel(k, dim3(l.grid), dim3(l.block), params, 0, stream);
}

}  // namespace fnd
