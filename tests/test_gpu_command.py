"""Compute command lists (fnd_pt_cmd_*, fnd_pt_submit, fences) on the GPU.

A frame recorded as commands (clear, path-trace dispatch, present dispatch),
submitted with a fence and waited on, must give the same image and the same
R8G8B8A8 bytes, bit for bit, as the direct calls (fnd_pt_framebuffer_clear,
fnd_pt_render, fnd_pt_framebuffer_present).  Consecutive path-trace
dispatches run as one batched launch and equal separate renders; trace
dispatches equal fnd_pt_trace_buffers; recording refuses calls that do not
fit the bound pipeline.
"""
import numpy as np
import pytest

from foundation_amd import _abi as A
from foundation_amd import pt
from foundation_amd import rays as R

pytestmark = pytest.mark.gpu

AS, BUF, IMG = A.DESCRIPTOR_ACCELERATION_STRUCTURE, A.DESCRIPTOR_STORAGE_BUFFER, A.DESCRIPTOR_STORAGE_IMAGE


@pytest.fixture(scope="module")
def scenes(ctx):
    gc = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    gi = pt.GeneratedScene(A.GEN_INSTANCED_SPHERES, instances=12, subdivision=10)
    sc, si = ctx.scene_create(gc.desc), ctx.scene_create(gi.desc)
    yield (gc, sc), (gi, si)
    ctx.scene_destroy(sc)
    ctx.scene_destroy(si)


def _words(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _record_frame(cl, scene, fb, cam, w, h, spp, depth, rr, seed, offset=0, present=None, exposure=1.0):
    k = pt.path_trace_constants(cam, depth, rr, seed=seed, sample_offset=offset)
    cl.set_pipeline(A.PIPELINE_PATH_TRACE).bind((0, AS, scene), (1, IMG, fb)).push_constants(k).dispatch(w, h, spp)
    if present is not None:
        ptr, pitch = present
        cl.set_pipeline(A.PIPELINE_PRESENT).bind((0, IMG, fb), (1, IMG, A.INVALID_HANDLE, 0, pitch, ptr))
        cl.push_constants(np.float32(exposure).tobytes()).dispatch(w, h, 1)


def test_recorded_frame_equals_direct_calls(ctx, scenes):
    (g, s), _ = scenes
    w, h, pitch = 96, 64, 512
    fa, fb = ctx.framebuffer_create(w, h), ctx.framebuffer_create(w, h)
    ma, mb = pt.DeviceMemory(pitch * h), pt.DeviceMemory(pitch * h)
    cl = ctx.command_list()
    fence = ctx.fence_create()
    try:
        # direct calls: two progressive frames, then the present
        for off in (0, 3):
            ctx.render([pt.make_job(s, fa, g.camera, 3, 5, rr_depth=2, seed=7, sample_offset=off)])
        ctx.framebuffer_present(fa, ma.ptr, pitch, exposure=1.5)
        # the same as a recording: clear, two path-trace dispatches into one image, present
        fill = np.full(pitch * h, 0x5A, np.uint8)
        ma_bytes = ma.read()
        mb.write(fill)
        cl.begin()
        cl.set_pipeline(A.PIPELINE_CLEAR).bind((0, IMG, fb)).dispatch(w, h, 1)
        _record_frame(cl, s, fb, g.camera, w, h, 3, 5, 2, 7, offset=0)
        _record_frame(cl, s, fb, g.camera, w, h, 3, 5, 2, 7, offset=3, present=(mb.ptr, pitch), exposure=1.5)
        cl.end()
        assert cl.size() == 4
        ctx.submit([cl], fence)
        ctx.fence_wait(fence)
        a, na = ctx.framebuffer_read(fa, w, h)
        b, nb = ctx.framebuffer_read(fb, w, h)
        assert na == nb == 6
        assert np.array_equal(_words(a), _words(b))
        got = mb.read().reshape(h, pitch)
        want = ma_bytes.reshape(h, pitch)
        assert np.array_equal(got[:, :4 * w], want[:, :4 * w])
        assert np.all(got[:, 4 * w:] == 0x5A)  # bytes between rows untouched
        # the list can be submitted again: clear + the same samples give the same image
        ctx.submit([cl], fence)
        ctx.fence_wait(fence)
        assert np.array_equal(_words(ctx.framebuffer_read(fb, w, h)[0]), _words(a))
    finally:
        cl.destroy()
        ctx.fence_destroy(fence)
        ma.free()
        mb.free()
        ctx.framebuffer_destroy(fa)
        ctx.framebuffer_destroy(fb)


def test_consecutive_dispatches_run_as_one_batch(ctx, scenes):
    (gc, sc), (gi, si) = scenes
    specs = [(sc, gc.camera, 48, 40, 4, 4, 11), (si, gi.camera, 64, 32, 2, 6, 12), (sc, gc.camera, 20, 36, 3, 3, 13)]
    alone = []
    for s, cam, w, h, spp, d, seed in specs:
        f = ctx.framebuffer_create(w, h)
        ctx.render([pt.make_job(s, f, cam, spp, d, seed=seed)])
        alone.append(ctx.framebuffer_read(f, w, h)[0])
        ctx.framebuffer_destroy(f)
    fbs = [ctx.framebuffer_create(w, h) for _, _, w, h, _, _, _ in specs]
    cl = ctx.command_list()
    try:
        ctx.reset_stats()
        cl.begin()
        for (s, cam, w, h, spp, d, seed), f in zip(specs, fbs):
            _record_frame(cl, s, f, cam, w, h, spp, d, d, seed)
        cl.end()
        ctx.submit([cl])
        ctx.wait()
        assert ctx.stats().paths >= sum(w * h * spp for _, _, w, h, spp, _, _ in specs)
        for (s, cam, w, h, spp, d, seed), f, ref in zip(specs, fbs, alone):
            img, n = ctx.framebuffer_read(f, w, h)
            assert n == spp
            assert np.array_equal(_words(img), _words(ref))
    finally:
        cl.destroy()
        for f in fbs:
            ctx.framebuffer_destroy(f)


def test_trace_dispatch_equals_trace_buffers(ctx, scenes, oracle):
    _, (g, s) = scenes
    rays = R.primary_rays(g.camera, 64, 48)
    n = rays.shape[0]
    off = 4096  # rays start past an offset inside the buffer
    rb = ctx.buffer_create(off + n * 32)
    hb = ctx.buffer_create(n * 16)
    hb2 = ctx.buffer_create(n * 16)
    cl = ctx.command_list()
    try:
        ctx.buffer_write(rb, rays, offset=off)
        for pipe, anyhit in ((A.PIPELINE_TRACE_CLOSEST, False), (A.PIPELINE_TRACE_ANYHIT, True)):
            cl.begin().set_pipeline(pipe).bind((0, AS, s), (1, BUF, rb, off), (2, BUF, hb)).dispatch(n).end()
            ctx.submit([cl])
            ctx.wait()
            got = ctx.buffer_read(hb, np.zeros((n, 4), np.float32))
            direct = ctx.trace(s, rays, anyhit=anyhit)
            if anyhit:
                assert np.array_equal(_words(got)[:, 1] != A.NO_HIT, _words(direct)[:, 1] != A.NO_HIT)
            else:
                assert np.array_equal(_words(got), _words(direct))
                assert np.array_equal(_words(got), _words(oracle.scene_from_generated(g).trace(rays)))
        # a buffer too small for the dispatch is refused while recording
        cl.begin().set_pipeline(A.PIPELINE_TRACE_CLOSEST).bind((0, AS, s), (1, BUF, rb, off), (2, BUF, hb2, 0, 16))
        with pytest.raises(pt.FndError) as e:
            cl.dispatch(n)
        assert e.value.status == A.E_INVALID_ARG
    finally:
        cl.destroy()
        for b in (rb, hb, hb2):
            ctx.buffer_destroy(b)


def test_recording_refuses_what_does_not_fit(ctx, scenes):
    (g, s), _ = scenes
    w, h = 16, 8
    f = ctx.framebuffer_create(w, h)
    cl = ctx.command_list()
    fence = ctx.fence_create()
    k = pt.path_trace_constants(g.camera, 3)
    try:
        ctx.fence_wait(fence)  # never submitted: already signalled
        with pytest.raises(pt.FndError):
            cl.dispatch(1)                                     # not recording
        cl.begin()
        bad = [
            lambda: cl.set_pipeline(99),
            lambda: cl.dispatch(w, h, 1),                      # no pipeline
            lambda: cl.set_pipeline(A.PIPELINE_PATH_TRACE).bind((1, IMG, f)).push_constants(k).dispatch(w, h, 1),
            lambda: cl.bind((0, AS, s), (1, IMG, f)).push_constants(b"1234").dispatch(w, h, 1),
            lambda: cl.push_constants(k).dispatch(w + 1, h, 1),
            lambda: cl.dispatch(w, h, 0),
            lambda: cl.bind((0, IMG, f), (0, AS, s)),           # one binding twice
            lambda: cl.bind((0, 9, s)),                         # unknown type
            lambda: cl.set_pipeline(A.PIPELINE_PRESENT).bind((0, IMG, f), (1, IMG, A.INVALID_HANDLE, 0, 4 * w - 4, 256))
            .push_constants(np.float32(1).tobytes()).dispatch(w, h, 1),
        ]
        for call in bad:
            with pytest.raises(pt.FndError) as e:
                call()
            assert e.value.status == A.E_INVALID_ARG
        with pytest.raises(pt.FndError):
            cl.end()                                            # a refused call taints the recording
        with pytest.raises(pt.FndError):
            ctx.submit([cl], fence)
        # a fresh recording is accepted
        cl.begin().set_pipeline(A.PIPELINE_PATH_TRACE).bind((0, AS, s), (1, IMG, f)).push_constants(k)
        cl.dispatch(w, h, 2).end()
        ctx.submit([cl], fence)
        ctx.fence_wait(fence)
        assert ctx.framebuffer_read(f, w, h)[1] == 2
    finally:
        cl.destroy()
        ctx.fence_destroy(fence)
        ctx.framebuffer_destroy(f)
