"""fnd_pt_trace_noalloc: the allocation-free variant of fnd_pt_trace_buffers.

It traces caller device memory with the traversal's scratch (work counter
and stack spill area) in the library's fixed per-GPU array, so the context
allocates nothing (its device-memory count does not move).  Two calls in a
row, the second with fewer rays, must give the records fnd_pt_trace_buffers
gives for the same rays, bit for bit, closest and any hit; the second call
leaves the hits past its own range untouched.  On the 10.08M-triangle scene
the spill of the full persistent grid does not fit the fixed array, so the
call runs fewer persistent blocks, with the same records.
"""
import numpy as np
import pytest

from foundation_amd import _abi as A
from foundation_amd import pt
from foundation_amd import rays as R

pytestmark = pytest.mark.gpu


def _words(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _reference(ctx, scene, rays, flags):
    n = rays.shape[0]
    rb, hb = ctx.buffer_create(n * 32), ctx.buffer_create(n * 16)
    try:
        ctx.buffer_write(rb, rays)
        ctx.trace_buffers(scene, rb, hb, n, flags)
        return ctx.buffer_read(hb, np.zeros((n, 4), np.float32)), ctx.stats().last_grid
    finally:
        ctx.buffer_destroy(rb)
        ctx.buffer_destroy(hb)


@pytest.mark.parametrize("kind,kw", [(A.GEN_CORNELL_SOUP, {}), (A.GEN_INSTANCED_SPHERES, {})])
def test_noalloc_trace_equals_trace_buffers(ctx, kind, kw):
    gen = pt.GeneratedScene(kind, **kw)
    scene = ctx.scene_create(gen.desc)
    mem_r = mem_h = None
    try:
        prim = R.primary_rays(gen.camera, 320, 180)
        ph = ctx.trace(scene, prim)
        bounce, _ = R.bounce_rays(prim, ph, gen.triangle_lookup(), seed=3)
        first = np.concatenate([prim, bounce], 0)
        second = bounce[: bounce.shape[0] // 3].copy()
        second[:, 0:3] += np.float32(1e-3)   # other rays than the first call's
        n1, n2 = first.shape[0], second.shape[0]
        print(f"noalloc: kind {kind}, {n1} then {n2} rays", flush=True)
        mem_r, mem_h = pt.DeviceMemory(n1 * 32), pt.DeviceMemory(n1 * 16)
        for flags in (0, A.TRACE_ANYHIT):
            ref1, full_grid = _reference(ctx, scene, first, flags)
            ref2, _ = _reference(ctx, scene, second, flags)
            before = ctx.stats().device_bytes
            mem_r.write(first)
            ctx.trace_noalloc(scene, mem_r.ptr, mem_h.ptr, n1, flags)
            got1 = mem_h.read().view(np.float32).reshape(n1, 4)
            grid = ctx.stats().last_grid
            mem_r.write(second)
            ctx.trace_noalloc(scene, mem_r.ptr, mem_h.ptr, n2, flags)
            got2 = mem_h.read().view(np.float32).reshape(n1, 4)
            assert ctx.stats().device_bytes == before  # nothing allocated
            if flags & A.TRACE_ANYHIT:
                assert np.array_equal(_words(got1)[:, 1] != A.NO_HIT, _words(ref1)[:, 1] != A.NO_HIT)
                assert np.array_equal(_words(got2)[:n2, 1] != A.NO_HIT, _words(ref2)[:, 1] != A.NO_HIT)
            else:
                assert np.array_equal(_words(got1), _words(ref1))
                assert np.array_equal(_words(got2[:n2]), _words(ref2))
            assert np.array_equal(_words(got2[n2:]), _words(got1[n2:]))  # past the second call's range: untouched
            print(f"noalloc: flags {flags}, grid {grid} (trace_buffers {full_grid})", flush=True)
            if kind == A.GEN_INSTANCED_SPHERES:
                assert grid < full_grid  # the full grid's spill does not fit the fixed scratch
    finally:
        if mem_r is not None:
            mem_r.free()
            mem_h.free()
        ctx.scene_destroy(scene)


def test_noalloc_trace_refuses_bad_arguments(ctx):
    gen = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    scene = ctx.scene_create(gen.desc)
    mem = pt.DeviceMemory(4096)
    try:
        for args in [(scene, mem.ptr, mem.ptr, 4, A.TRACE_COUNT), (scene, mem.ptr + 4, mem.ptr, 4, 0),
                     (scene, 0, mem.ptr, 4, 0), (A.INVALID_HANDLE, mem.ptr, mem.ptr, 4, 0)]:
            with pytest.raises(pt.FndError) as e:
                ctx.trace_noalloc(*args)
            assert e.value.status in (A.E_INVALID_ARG, A.E_NOT_FOUND)
        ctx.trace_noalloc(scene, 0, 0, 0)  # nothing to do
    finally:
        mem.free()
        ctx.scene_destroy(scene)
