// arena.cpp — bump allocation over one block of memory (host bookkeeping).
//
// The semantics are those of the reference's Core::StackAllocator
// (src/Core/Allocator/StackAllocator.hpp:9-63, StackAllocator.cpp:5-29):
// the aligned form rounds both the start and the end of a block up to the
// alignment, the unaligned form takes exactly `size` bytes, and a request that
// is empty, has no memory to come from or would pass the end returns null and
// takes nothing.  Used memory is the distance from the start of the memory to
// the next free byte.
//
// Allocation is one compare-and-swap on `current`, so concurrent callers get
// disjoint blocks (the reference's multi-threaded counter loads, compares and
// stores in three steps, StackAllocator.cpp:8-15, and two threads can be given
// the same block).  The addresses are only computed, never dereferenced, so
// the block may live in device memory: the path tracer carves its per-frame
// ray queues this way.
#include <cstdint>

#include "../../include/fnd_pt.h"

namespace {

// Start and end of a block of `size` bytes at `cur`, or false when it does
// not fit in [cur, end) or its arithmetic would wrap.  `mask` = alignment - 1
// (0 for the unaligned form).
inline bool place(uint64_t cur, uint64_t end, uint64_t size, uint64_t mask, uint64_t* start, uint64_t* stop)
{
    if (cur > UINT64_MAX - mask) return false;
    const uint64_t s = (cur + mask) & ~mask;
    if (size > UINT64_MAX - mask - s) return false;
    const uint64_t e = (s + size + mask) & ~mask;
    if (e > end) return false;
    *start = s;
    *stop = e;
    return true;
}

void* take(fnd_arena* a, uint64_t size, uint64_t mask)
{
    if (!a || size == 0 || a->memory == 0) return nullptr;
    uint64_t cur = __atomic_load_n(&a->current, __ATOMIC_RELAXED);
    for (;;) {
        uint64_t s, e;
        if (!place(cur, a->end, size, mask, &s, &e)) return nullptr;
        // on failure `cur` is reloaded with the value another caller stored
        if (__atomic_compare_exchange_n(&a->current, &cur, e, true, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED))
            return reinterpret_cast<void*>(static_cast<uintptr_t>(s));
    }
}

}  // namespace

extern "C" {

void fnd_arena_reset(fnd_arena* a, void* memory, uint64_t size)
{
    if (!a) return;
    const uint64_t m = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(memory));
    if (m == 0) {
        a->memory = a->end = a->current = 0;
        return;
    }
    a->memory = m;
    a->end = size > UINT64_MAX - m ? UINT64_MAX : m + size;
    __atomic_store_n(&a->current, m, __ATOMIC_RELEASE);
}

void* fnd_arena_alloc(fnd_arena* a, uint64_t size) { return take(a, size, 0); }

void* fnd_arena_alloc_aligned(fnd_arena* a, uint64_t size, uint64_t alignment)
{
    if (alignment == 0 || (alignment & (alignment - 1)) != 0) return nullptr;
    return take(a, size, alignment - 1);
}

uint64_t fnd_arena_used(const fnd_arena* a)
{
    if (!a || a->memory == 0) return 0;
    return __atomic_load_n(&a->current, __ATOMIC_ACQUIRE) - a->memory;
}

}  // extern "C"
