// C entry points over the reference's Core::StackAllocator, which
// oracle/ref_stack.mk compiles from the reference's own
// src/Core/Allocator/StackAllocator.cpp.  Test infrastructure only: the tests
// use it to pin the restatement in oracle/stack_alloc.py.  The allocator only
// computes addresses, so the memory given here may be any address.
#include <cstdint>
#include <new>

#include <Core/Allocator/StackAllocator.hpp>

using Foundation::Core::Allocator;
using Foundation::Core::StackAllocatorSingleThreaded;

extern "C" {

void* ref_stack_new(void* memory, uint64_t size)
{
    return new (std::nothrow) StackAllocatorSingleThreaded(Allocator::Arena{memory, size});
}

void ref_stack_reset(void* h, void* memory, uint64_t size)
{
    static_cast<StackAllocatorSingleThreaded*>(h)->Reset(Allocator::Arena{memory, size});
}

void* ref_stack_alloc(void* h, uint64_t size) { return static_cast<StackAllocatorSingleThreaded*>(h)->Allocate(size); }

void* ref_stack_alloc_aligned(void* h, uint64_t size, uint64_t alignment)
{
    return static_cast<StackAllocatorSingleThreaded*>(h)->Allocate(size, alignment);
}

uint64_t ref_stack_used(void* h) { return static_cast<StackAllocatorSingleThreaded*>(h)->GetUsedMemory(); }

void ref_stack_free(void* h) { delete static_cast<StackAllocatorSingleThreaded*>(h); }

}  // extern "C"
