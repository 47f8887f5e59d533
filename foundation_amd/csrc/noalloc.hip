// noalloc.hip — device scratch of the allocation-free trace entry point
// (fnd_pt_trace_noalloc), for callers that cannot allocate device memory.
//
// The traversal needs memory beyond its rays and hits: a work counter that
// the persistent waves take ray blocks from, and a spill area for the stack
// entries past the LDS part.  fnd_pt_trace_buffers allocates both with the
// context; this entry point uses one fixed array per GPU instead, resolved
// to a device address once per context by fnd_pt_create.  Every call on a
// GPU shares it, so calls that overlap on one GPU race; a call zeroes the
// counter's shards on its stream before the launch (with a kernel, not a
// memset) and needs nothing else cleared.
#include "ctx.hpp"

namespace fnd {

// [0, kNoAllocSpillOffset): the work counter's shards; [kNoAllocSpillOffset,
// kNoAllocScratchBytes): the spill area.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

__global__ void k_zero_counter(uint32_t* counter)
{
    zero_trace_counter(counter, threadIdx.x);
}

}  // namespace

hipError_t noalloc_scratch_address(void** out)
{
    *out = nullptr;
    return hipGetSymbolAddress(out, HIP_SYMBOL(scratch));
}

void launch_zero_counter(uint32_t* counter, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, counter);
}

}  // namespace fnd
