"""CPU restatement of the reference's bump allocator, Core::StackAllocator.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the frame
arena (fnd_arena_* and the per-frame ray queues carved from it).

Follows src/Core/Allocator/StackAllocator.hpp:18-29,54-56 (Reset, Reset(),
GetUsedMemory) and src/Core/Allocator/StackAllocator.cpp:5-29 (the aligned and
unaligned Allocate), with AlignUp of src/Core/Allocator/Allocator.hpp:9-11.
Addresses are plain integers; None stands for the reference's nullptr.

The reference is single-threaded here: the product's compare-and-swap makes
no difference to the sequence of blocks one thread is given.  Alignments are
powers of two (the reference's mask arithmetic is only defined for them) and
addresses stay below 2^64, the ranges the product and the reference agree on.

Pinned by the reference itself, compiled from its own sources into
oracle/_ref/libref_stack.so by oracle/ref_stack.mk, on random sequences, and by
hand-worked cases (tests/test_arena.py).
"""

MASK64 = (1 << 64) - 1


def align_up(value, alignment):
    """Allocator.hpp:9-11."""
    return (value + alignment - 1) & ~(alignment - 1) & MASK64


class StackAllocator:
    def __init__(self, memory=0, size=0):
        self.reset(memory, size)

    def reset(self, memory=0, size=0):
        """StackAllocator.hpp:18-23 (memory 0 is Reset(), hpp:27-29)."""
        self.memory = memory
        self.current = memory
        self.end = memory + size

    def allocate(self, size, alignment=None):
        """StackAllocator.cpp:5-16 with an alignment, :18-29 without."""
        if size == 0 or self.memory == 0:
            return None
        if alignment is None:
            start = self.current
            end = start + size
        else:
            start = align_up(self.current, alignment)
            end = align_up(start + size, alignment)
        if end > self.end:
            return None
        self.current = end
        return start

    def used(self):
        """StackAllocator.hpp:54-56."""
        return self.current - self.memory
