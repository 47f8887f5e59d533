// trace_launch.hip — instantiations of k_trace and the launcher that picks one.
#include "trace_kernel.hpp"

namespace fnd {

#define FND_TRACE_CONFIGS(X) \
    X(64, 16)                \
    X(64, 32)                \
    X(128, 8)                \
    X(128, 16)               \
    X(128, 32)               \
    X(256, 8)                \
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
static hipError_t launch_cfg(const TraceLaunch& l, const TraceArgs& args, hipStream_t stream)
{
    const dim3 grid(l.grid), block(B);
    if (l.multi) {
        if (l.mode == kModeClosestInternal)
            hipLaunchKernelGGL((k_trace<kModeClosestInternal, true, true, false, S, B>), grid, block, 0, stream, args);
        else if (l.mode == kModeAnyInternal)
            hipLaunchKernelGGL((k_trace<kModeAnyInternal, true, true, false, S, B>), grid, block, 0, stream, args);
        else if (l.mode == kModeMixedInternal)
            hipLaunchKernelGGL((k_trace<kModeMixedInternal, true, true, false, S, B>), grid, block, 0, stream, args);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (l.count) {
        // counting variants exist for the default configuration only
        return hipErrorInvalidValue;
    }
    if (l.two_level) {
        if (l.mode == kModeClosest)
            hipLaunchKernelGGL((k_trace<kModeClosest, true, false, false, S, B>), grid, block, 0, stream, args);
        else if (l.mode == kModeAny)
            hipLaunchKernelGGL((k_trace<kModeAny, true, false, false, S, B>), grid, block, 0, stream, args);
        else
            return hipErrorInvalidValue;
    } else {
        if (l.mode == kModeClosest)
            hipLaunchKernelGGL((k_trace<kModeClosest, false, false, false, S, B>), grid, block, 0, stream, args);
        else if (l.mode == kModeAny)
            hipLaunchKernelGGL((k_trace<kModeAny, false, false, false, S, B>), grid, block, 0, stream, args);
        else
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static hipError_t launch_count(const TraceLaunch& l, const TraceArgs& args, hipStream_t stream)
{
    const dim3 grid(l.grid), block(128);
    if (l.two_level) {
        if (l.mode == kModeClosest)
            hipLaunchKernelGGL((k_trace<kModeClosest, true, false, true, 16, 128>), grid, block, 0, stream, args);
        else
            hipLaunchKernelGGL((k_trace<kModeAny, true, false, true, 16, 128>), grid, block, 0, stream, args);
    } else {
        if (l.mode == kModeClosest)
            hipLaunchKernelGGL((k_trace<kModeClosest, false, false, true, 16, 128>), grid, block, 0, stream, args);
        else
            hipLaunchKernelGGL((k_trace<kModeAny, false, false, true, 16, 128>), grid, block, 0, stream, args);
    }
    return hipGetLastError();
}

hipError_t launch_trace(const TraceLaunch& l, const TraceArgs& args, hipStream_t stream)
{
    if (l.count && !l.multi) {
        if (l.block != 128 || l.stack_lds != 16) return hipErrorInvalidValue;
        return launch_count(l, args, stream);
    }
#define FND_DISPATCH(B, S) \
    if (l.block == B && l.stack_lds == S) return launch_cfg<B, S>(l, args, stream);
    FND_TRACE_CONFIGS(FND_DISPATCH)
#undef FND_DISPATCH
    return hipErrorInvalidValue;
}

}  // namespace fnd
