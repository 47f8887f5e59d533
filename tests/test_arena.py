"""Frame arena: bump allocation with the reference's Core::StackAllocator
semantics (fnd_arena_*), and the per-frame ray queues carved from it.

CPU: the restatement in oracle/stack_alloc.py against the reference's own
allocator compiled from its sources (oracle/_ref/libref_stack.so) and against
hand-worked cases; the C ABI against the restatement; concurrent allocations
never overlap; the queue footprint per path.
GPU: renders whose queues come from the caller's device memory, at full and at
reduced batch sizes, give the same image bit for bit as the library's own
block; a block too small for the minimum batch fails cleanly.
"""
import ctypes as C
import os
import random
import threading

import numpy as np
import pytest

from foundation_amd import _abi as A
from foundation_amd import pt
from oracle.stack_alloc import StackAllocator as OracleStack

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_stack.so")


class RefStack:
    """The reference's StackAllocatorSingleThreaded through oracle/ref_stack_shim.cpp."""
    _lib = None

    def __init__(self, memory, size):
        if RefStack._lib is None:
            lib = C.CDLL(REF_LIB)
            lib.ref_stack_new.restype = C.c_void_p
            lib.ref_stack_new.argtypes = [C.c_void_p, C.c_uint64]
            lib.ref_stack_reset.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
            lib.ref_stack_alloc.restype = C.c_void_p
            lib.ref_stack_alloc.argtypes = [C.c_void_p, C.c_uint64]
            lib.ref_stack_alloc_aligned.restype = C.c_void_p
            lib.ref_stack_alloc_aligned.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
            lib.ref_stack_used.restype = C.c_uint64
            lib.ref_stack_used.argtypes = [C.c_void_p]
            lib.ref_stack_free.argtypes = [C.c_void_p]
            RefStack._lib = lib
        self.lib = RefStack._lib
        self.h = self.lib.ref_stack_new(memory or None, size)

    def reset(self, memory=0, size=0):
        self.lib.ref_stack_reset(self.h, memory or None, size)

    def allocate(self, size, alignment=None):
        if alignment is None:
            return self.lib.ref_stack_alloc(self.h, size)
        return self.lib.ref_stack_alloc_aligned(self.h, size, alignment)

    def used(self):
        return self.lib.ref_stack_used(self.h)

    def close(self):
        self.lib.ref_stack_free(self.h)


def random_program(rng, n_ops):
    """A sequence of resets and allocations that stays far from address wrap
    (the reference's arithmetic is only defined there) and uses power-of-two
    alignments, sizes 0 and requests larger than the block."""
    ops = []
    for _ in range(n_ops):
        r = rng.random()
        if r < 0.03:
            base = rng.choice([0, rng.randrange(1, 1 << 40), rng.randrange(1, 1 << 30) * 4096])
            ops.append(("reset", base, rng.choice([0, 1, 100, rng.randrange(1, 1 << 20)])))
        else:
            size = rng.choice([0, 1, 3, 8, rng.randrange(1, 4096), rng.randrange(1, 1 << 18), 1 << 30])
            align = rng.choice([None, None, 1, 2, 4, 8, 16, 64, 256, 4096])
            ops.append(("alloc", size, align))
    return ops


def run_program(alloc, ops):
    trace = []
    for op in ops:
        if op[0] == "reset":
            alloc.reset(op[1], op[2])
            trace.append(("reset", alloc.used()))
        else:
            trace.append((alloc.allocate(op[1], op[2]), alloc.used()))
    return trace


# ---------------------------------------------------------------- the oracle


def test_oracle_hand_worked_cases():
    a = OracleStack(1000, 100)  # unaligned block [1000, 1100)
    assert a.allocate(10, 16) == 1008 and a.used() == 24  # start 1008, end rounded to 1024
    assert a.allocate(5) == 1024 and a.used() == 29  # unaligned: exactly 5 bytes
    assert a.allocate(1, 8) == 1032 and a.used() == 40
    assert a.allocate(60, 4) == 1040 and a.used() == 100  # ends exactly at the end
    assert a.allocate(1) is None and a.used() == 100
    assert a.allocate(0) is None
    b = OracleStack(0x1000, 20)
    assert b.allocate(17, 16) is None  # 17 bytes fit, the rounded end (0x1020) does not
    assert b.used() == 0  # nothing taken
    assert b.allocate(20) == 0x1000
    b.reset()
    assert b.allocate(1) is None and b.used() == 0
    c = OracleStack(0x2000, 0)
    assert c.allocate(1) is None


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref/libref_stack.so not built (needs the reference's sources)")
def test_oracle_matches_the_reference_build():
    rng = random.Random(7)
    for trial in range(40):
        base = rng.choice([4096, 1000, rng.randrange(1, 1 << 40)])
        size = rng.choice([0, 64, 1 << 16, 1 << 24, rng.randrange(1, 1 << 32)])
        ops = random_program(rng, 300)
        ref = RefStack(base, size)
        try:
            got_ref = run_program(ref, ops)
        finally:
            ref.close()
        got_orc = run_program(OracleStack(base, size), ops)
        assert got_orc == got_ref, trial


# ---------------------------------------------------------------- the C ABI


def test_arena_matches_oracle():
    rng = random.Random(11)
    for trial in range(40):
        base = rng.choice([4096, 1000, rng.randrange(1, 1 << 40)])
        size = rng.choice([0, 64, 1 << 16, 1 << 24, rng.randrange(1, 1 << 32)])
        ops = random_program(rng, 300)
        assert run_program(pt.StackAllocator(base, size), ops) == run_program(OracleStack(base, size), ops), trial


def test_arena_rejects_what_the_reference_leaves_undefined():
    a = pt.StackAllocator(4096, 1 << 20)
    assert a.allocate(8, 0) is None and a.allocate(8, 3) is None and a.allocate(8, 48) is None
    assert a.used() == 0
    top = (1 << 64) - 4096
    b = pt.StackAllocator(top, 4095)  # the block ends at 2^64 - 1
    assert b.allocate(4096) is None  # would wrap
    assert b.allocate(8, 1 << 63) is None  # rounding the start up wraps
    assert b.allocate(4000) == top and b.used() == 4000
    with pytest.raises(RuntimeError):
        a.reallocate(0, 16)
    a.deallocate(4096, 8)  # no-op
    lib = A.load()
    assert lib.fnd_arena_alloc(None, 8) is None
    assert lib.fnd_arena_used(None) == 0


def test_concurrent_allocations_are_disjoint():
    """Eight threads allocate from one arena until it runs out (ctypes calls
    release the GIL, so the compare-and-swap is contended)."""
    lib = A.load()
    base, size = 1 << 32, 1 << 24
    arena = A.Arena()
    lib.fnd_arena_reset(C.byref(arena), base, size)
    got = [[] for _ in range(8)]

    def worker(k):
        rng = random.Random(100 + k)
        out = got[k]
        for _ in range(6000):
            n = rng.randrange(1, 2048)
            al = 1 << rng.randrange(0, 9)
            p = lib.fnd_arena_alloc_aligned(C.byref(arena), n, al)
            if p is not None:
                out.append((p, n, al))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    blocks = sorted(b for g in got for b in g)
    assert len(blocks) > 1000
    for p, n, al in blocks:
        assert p % al == 0 and base <= p and p + n <= base + size
    for (p0, n0, _), (p1, _, _) in zip(blocks, blocks[1:]):
        assert p0 + n0 <= p1
    assert lib.fnd_arena_used(C.byref(arena)) >= blocks[-1][0] + blocks[-1][1] - base
    assert lib.fnd_arena_used(C.byref(arena)) <= size


# ---------------------------------------------------------------- per-frame ray queues

# Bytes per slot of the queues: path state (job, pixel, sample, train slot,
# window index, bounce 4 each; throughput and radiance 16 each) 56; two
# extension queues of a 32-byte ray, a path id and a scene slot, 80;
# extension hits 16; the shadow queue (ray, path, scene, contribution 16,
# hit 16) 72; cache queries (48-byte query, path 4, weight 16, prediction 16)
# 84; the window's radiance, 16 B for each of 8 window paths per slot, 128;
# two free lists, 8.
QUEUE_BYTES_PER_PATH = 56 + 80 + 16 + 72 + 84 + 128 + 8
# counters (16 slots, one 256-byte line each), the traversal work counter (16
# shards, one 256-byte line each) and the ray total (one 256-byte slot)
QUEUE_BYTES_FIXED = 16 * 256 + 16 * 256 + 256


def test_queue_bytes_per_path():
    for k in (1, 2, 5, 64, 1000, 32768, 65536):
        cap = 64 * k  # every buffer a multiple of 256 bytes: no padding
        assert pt.queue_bytes(cap) == QUEUE_BYTES_PER_PATH * cap + QUEUE_BYTES_FIXED, cap
    prev = 0
    for cap in range(1, 600, 7):  # padding adds at most 255 bytes to each of the 29 per-slot buffers
        q = pt.queue_bytes(cap)
        assert QUEUE_BYTES_PER_PATH * cap + QUEUE_BYTES_FIXED <= q <= QUEUE_BYTES_PER_PATH * cap + QUEUE_BYTES_FIXED + 29 * 256
        assert q % 256 == 0 and q >= prev
        prev = q
    with pytest.raises(pt.FndError):
        pt.queue_bytes(0)


def test_frame_arena_entry_points_check_arguments(lib):
    assert lib.fnd_pt_set_frame_arena(None, None, 0) == A.E_INVALID_ARG
    assert lib.fnd_pt_frame_arena_info(None, None) == A.E_INVALID_ARG
    assert lib.fnd_pt_queue_bytes(16, None) == A.E_INVALID_ARG


# ---------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def cornell(ctx):
    g = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    s = ctx.scene_create(g.desc)
    yield g, s
    ctx.scene_destroy(s)


def _render(ctx, s, cam, w, h, spp, depth, seed):
    fb = ctx.framebuffer_create(w, h)
    try:
        ctx.render([pt.make_job(s, fb, cam, spp, depth, seed=seed)])
        return ctx.framebuffer_read(fb, w, h)[0]
    finally:
        ctx.framebuffer_destroy(fb)


def largest_batch(nbytes, want):
    """The batch a caller's block of `nbytes` holds: the largest cap <= want
    whose queues fit (queue_bytes never decreases with the cap)."""
    if pt.queue_bytes(want) <= nbytes:
        return want
    lo, hi = 0, want
    while hi - lo > 1:
        mid = (lo + hi) // 2
        lo, hi = (mid, hi) if pt.queue_bytes(mid) <= nbytes else (lo, mid)
    return lo


@pytest.mark.gpu
def test_caller_arena_gives_the_same_images(ctx, cornell):
    g, s = cornell
    W, H, spp, depth = 40, 24, 4, 4
    ctx.set_frame_arena(0, 0)
    own = _render(ctx, s, g.camera, W, H, spp, depth, seed=3)
    info = ctx.frame_arena_info()
    full = info.batch_paths  # the default batch holds the whole call: W x H padded to whole tiles, x spp
    assert info.caller_owned == 0 and W * H * spp <= full < (W + 16) * (H + 16) * spp
    assert info.used == pt.queue_bytes(full) <= info.size
    try:
        for nbytes in [pt.queue_bytes(full) + 4096, pt.queue_bytes(3000), pt.queue_bytes(1024) + 200]:
            with pt.DeviceMemory(nbytes) as mem:
                ctx.set_frame_arena(mem.ptr, nbytes)
                try:
                    img = _render(ctx, s, g.camera, W, H, spp, depth, seed=3)
                    info = ctx.frame_arena_info()
                finally:
                    ctx.set_frame_arena(0, 0)  # before the memory goes back
            batch = largest_batch(nbytes, full)
            assert info.caller_owned == 1 and info.size == nbytes
            assert info.batch_paths == batch and batch >= 1024, (nbytes, info.batch_paths)
            assert info.used == pt.queue_bytes(batch)
            assert np.array_equal(img.view(np.uint32), own.view(np.uint32)), nbytes
    finally:
        ctx.set_frame_arena(0, 0)


@pytest.mark.gpu
def test_too_small_caller_arena_fails_cleanly(ctx, cornell):
    g, s = cornell
    nbytes = pt.queue_bytes(1024) - 256  # fewer than 1024 paths fit
    mem = pt.DeviceMemory(nbytes)
    try:
        ctx.set_frame_arena(mem.ptr, nbytes)
        with pytest.raises(pt.FndError) as e:
            _render(ctx, s, g.camera, 40, 24, 2, 2, seed=1)  # 3072 paths: a batch must hold 1024
        assert e.value.status == A.E_OOM_DEVICE and "at least 1024" in e.value.message
        img = _render(ctx, s, g.camera, 16, 16, 1, 2, seed=1)  # a call of 256 paths needs a batch of 256
        assert ctx.frame_arena_info().batch_paths == 256
        with pytest.raises(pt.FndError):
            ctx.set_frame_arena(mem.ptr + 8, 1024)  # not 256-byte aligned
    finally:
        ctx.set_frame_arena(0, 0)
        mem.free()
    ref = _render(ctx, s, g.camera, 16, 16, 1, 2, seed=1)
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_library_block_is_allocated_once_per_size(ctx, cornell):
    g, s = cornell
    ctx.set_frame_arena(0, 0)
    _render(ctx, s, g.camera, 64, 64, 2, 2, seed=1)
    before = ctx.stats().device_bytes
    size = ctx.frame_arena_info().size
    for seed in range(2, 6):  # same and smaller batches reuse the block
        _render(ctx, s, g.camera, 64, 64, 2, 2, seed=seed)
        _render(ctx, s, g.camera, 32, 32, 1, 2, seed=seed)
        assert ctx.frame_arena_info().size == size
        assert ctx.stats().device_bytes == before


@pytest.mark.gpu
def test_device_memory_budget_fails_cleanly():
    """A context's device-memory budget is enforced as FND_E_OOM_DEVICE, with
    no crash, and the context stays usable: a scene larger than the budget
    is refused, and so is a frame whose ray queues exceed it, after which a
    small frame still renders."""
    g = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    with pt.Context(0, device_memory_budget=256 << 10) as c:  # 10,034 triangles x 48 B do not fit
        with pytest.raises(pt.FndError) as e:
            c.scene_create(g.desc)
        assert e.value.status == A.E_OOM_DEVICE and "budget" in e.value.message
    # 128 MiB: the scene, a 1080p framebuffer (24.9 MB) and the traversal
    # stacks' spill area for the whole persistent grid (about 60 MiB here) fit;
    # the ray queues of a 1080p 1-spp frame (878 MiB) do not
    with pt.Context(0, device_memory_budget=128 << 20) as c:
        s = c.scene_create(g.desc)
        fb = c.framebuffer_create(1920, 1080)
        with pytest.raises(pt.FndError) as e:
            c.render([pt.make_job(s, fb, g.camera, 1, 4)])
        assert e.value.status == A.E_OOM_DEVICE
        small = c.framebuffer_create(32, 32)
        c.render([pt.make_job(s, small, g.camera, 1, 4)])
        img, n = c.framebuffer_read(small, 32, 32)
        assert n == 1 and np.all(np.isfinite(img)) and img.max() > 0
