"""Radiance cache on the GPU vs the FP64 CPU oracle, and what it gains.

Tolerances, about twice the agreement measured on the MI355X: the GPU
forward stores weights, grid values and activations in FP16 (the oracle
rounds the same quantities with round_fp16=True): |y_gpu - y_ref| <= 6e-4 +
1e-3 |y_ref| (measured 3.1e-4 at most).  The backward multiplies in BF16
(8-bit mantissa) with FP32 accumulation, so gradients are compared as whole
vectors per layer, |g_gpu - g_ref| / |g_ref| < 1e-2 (measured <= 0.4%).  The
hash grid's gradient is < 2.5e-2 off whether its sums use packed BF16
atomics (the default, measured 1.85%) or FP32 atomics (FND_CACHE_GRAD_FP32,
measured 1.84%): the summation is not what limits it, the BF16 operands of
the backward are.
"""
import ctypes as C

import numpy as np
import pytest

from foundation_amd import _abi as A
from foundation_amd import pt

pytestmark = pytest.mark.gpu

LAYERS = [(0, 3072), (3072, 7168), (7168, 11264), (11264, 15360), (15360, 19456), (19456, 20480)]


def configure(ctx, lo=(0, 0, 0), hi=(1, 1, 1), **kw):
    d = A.CacheDesc()
    assert ctx.lib.fnd_pt_cache_default_desc(C.byref(d)) == 0
    for i in range(3):
        d.bounds_lo[i] = lo[i]
        d.bounds_hi[i] = hi[i]
    for k, v in kw.items():
        setattr(d, k, v)
    ctx._ck(ctx.lib.fnd_pt_cache_configure(ctx.h, C.byref(d)))


def set_params(ctx, mlp, grid):
    ctx._ck(ctx.lib.fnd_pt_cache_set_params(ctx.h, mlp.ctypes.data, grid.ctypes.data))


def gradients(ctx, q, t):
    loss = C.c_float()
    dm = np.zeros(20480, np.float32)
    dg = np.zeros(16 * (1 << 19) * 2, np.float32)
    ctx._ck(ctx.lib.fnd_pt_cache_gradients(ctx.h, q.ctypes.data, t.ctypes.data, q.shape[0], C.byref(loss),
                                           dm.ctypes.data, dg.ctypes.data))
    return loss.value, dm, dg


def infer(ctx, q):
    out = np.zeros((q.shape[0], 3), np.float32)
    ctx._ck(ctx.lib.fnd_pt_cache_infer(ctx.h, q.ctypes.data, q.shape[0], out.ctypes.data))
    return out


@pytest.fixture(scope="module")
def params(oracle):
    rng = np.random.default_rng(1)
    mlp = rng.uniform(-0.25, 0.25, 20480).astype(np.float32)
    grid = rng.uniform(-0.5, 0.5, oracle.CACHE_GRID_PARAMS).astype(np.float32)
    n = 2048
    pos = rng.uniform(0.02, 0.98, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    # half the queries are glossy: outgoing direction and roughness in [0.5, 1]
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rough = rng.uniform(0.5, 1.0, (n, 1))
    glossy = (np.arange(n) % 2 == 1)[:, None]
    q = np.concatenate([pos, nrm, np.where(glossy, d, 0.0), np.where(glossy, rough, 0.0)], 1).astype(np.float32)
    t = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    return mlp, grid, q, t


def test_forward_matches_oracle(ctx, oracle, params):
    mlp, grid, q, t = params
    configure(ctx)
    set_params(ctx, mlp, grid)
    got = infer(ctx, q)
    _, ref, _, _ = oracle.cache_loss_grad(mlp, grid, q, t, [0, 0, 0], 1.0, round_fp16=True, want_grid=False)
    err = np.abs(got - ref)
    print(f"forward max err {err.max():.3e}, ref scale {np.abs(ref).mean():.3e}", flush=True)
    assert np.all(err <= 6e-4 + 1e-3 * np.abs(ref)), err.max()


@pytest.mark.parametrize("mode,grid_tol", [(A.CACHE_GRAD_BF16, 2.5e-2), (A.CACHE_GRAD_FP32, 2.5e-2)])
def test_gradients_match_oracle(ctx, oracle, params, mode, grid_tol):
    mlp, grid, q, t = params
    configure(ctx, grid_gradient=mode)
    set_params(ctx, mlp, grid)
    loss, dm, dg = gradients(ctx, q, t)
    rloss, _, rdm, rdg = oracle.cache_loss_grad(mlp, grid, q, t, [0, 0, 0], 1.0, round_fp16=True)
    print(f"loss gpu {loss:.6g} oracle {rloss:.6g}", flush=True)
    assert abs(loss - rloss) <= 5e-3 * abs(rloss)
    for a, b in LAYERS:
        rel = np.linalg.norm(dm[a:b] - rdm[a:b]) / np.linalg.norm(rdm[a:b])
        print(f"layer [{a}:{b}) relative gradient error {rel:.3e}", flush=True)
        assert rel < 1e-2, (a, b, rel)
    rel = np.linalg.norm(dg - rdg) / np.linalg.norm(rdg)
    print(f"grid relative gradient error ({'fp32' if mode == A.CACHE_GRAD_FP32 else 'bf16'} sums) {rel:.3e}",
          flush=True)
    assert rel < grid_tol
    # no gradient outside the touched entries
    assert np.all(dg[rdg == 0] == 0)


def test_adam_first_step(ctx, params):
    mlp, grid, q, t = params
    configure(ctx, learning_rate=1e-2)
    set_params(ctx, mlp, grid)
    _, dm, dg = gradients(ctx, q, t)
    loss = C.c_float()
    ctx._ck(ctx.lib.fnd_pt_cache_train(ctx.h, q.ctypes.data, t.ctypes.data, q.shape[0], 1, C.byref(loss)))
    new_m = np.zeros(20480, np.float32)
    new_g = np.zeros_like(grid)
    ctx._ck(ctx.lib.fnd_pt_cache_get_params(ctx.h, new_m.ctypes.data, new_g.ctypes.data))
    sel = np.abs(dm) > 1e-6
    np.testing.assert_allclose((new_m - mlp)[sel], -1e-2 * np.sign(dm[sel]), rtol=0, atol=2e-6)
    assert np.all(new_m[dm == 0] == mlp[dm == 0])
    gs = np.abs(dg) > 1e-6
    step = (new_g - grid)[gs]
    ok = np.abs(step + 1e-2 * np.sign(dg[gs])) <= 2e-6
    # The bf16 atomic sums of the two calls add in different orders, so an
    # entry whose gradient nearly cancels can change sign, or cancel to exactly
    # zero (the lazy update then skips it), between them: allow that for at
    # most 1e-4 of the entries, each within 1e-3 of the largest gradient.
    # Every other entry moves by exactly -lr * sign(gradient).
    bad = ~ok
    assert bad.sum() <= 1e-4 * gs.sum(), bad.sum()
    assert np.all(np.abs(dg[gs][bad]) < 1e-3 * np.abs(dg).max())
    assert np.all((np.abs(step[bad]) == 0) | (np.abs(np.abs(step[bad]) - 1e-2) <= 2e-6))


def test_training_fits_a_smooth_field(ctx):
    rng = np.random.default_rng(3)
    configure(ctx, seed=7)
    n = 1 << 16
    pos = rng.uniform(0, 1, (n, 3))
    nrm = np.tile([0.0, 0.0, 1.0], (n, 1))
    q = np.concatenate([pos, nrm, np.zeros((n, 4))], 1).astype(np.float32)
    # targets in [0.5, 2.5] so the relative error below is well defined
    t = np.stack([1.5 + np.sin(6 * pos[:, 0]), 1.5 + np.cos(5 * pos[:, 1]), 1 + pos[:, 2] ** 2], 1).astype(np.float32)
    loss = C.c_float()
    ctx._ck(ctx.lib.fnd_pt_cache_train(ctx.h, q.ctypes.data, t.ctypes.data, n, 1, C.byref(loss)))
    first = loss.value
    ctx._ck(ctx.lib.fnd_pt_cache_train(ctx.h, q.ctypes.data, t.ctypes.data, n, 200, C.byref(loss)))
    last = loss.value
    print(f"loss {first:.4g} -> {last:.4g}", flush=True)
    assert last < 0.05 * first
    pred = infer(ctx, q[:4096])
    err = float(np.mean(np.abs(pred - t[:4096]) / t[:4096]))
    print(f"mean relative error {err:.4f}", flush=True)
    # measured 0.005-0.025 over ten runs, once 0.051: the grid gradients are
    # summed with BF16 atomics in no fixed order, and an unlucky order now and
    # then leaves the 200 steps with a higher loss
    assert err < 0.08


def rel_mse(x, y):
    return float(np.mean((x.astype(np.float64) - y) ** 2 / (y.astype(np.float64) ** 2 + 0.01)))


def test_cache_queries_on_two_streams_match_one_window(ctx, params):
    """Paths that end in the radiance cache (fixed parameters, no training)
    have most of their queries answered on a second stream while the frame's
    last bounces trace, and the rest after the bounce loop.  The image is the
    same bit for bit from run to run, and with 1024 slots (twelve windows that
    reuse the query buffers) as with one window; and it differs from plain
    path tracing, so the cache was used."""
    mlp, grid, _, _ = params
    g = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    scene = ctx.scene_create(g.desc)
    w, h = 64, 48
    old = ctx.tuning()

    def draw(flags):
        fb = ctx.framebuffer_create(w, h)
        try:
            ctx.render([pt.make_job(scene, fb, g.camera, 4, 6, seed=7, flags=flags)])
            return ctx.framebuffer_read(fb, w, h)[0]
        finally:
            ctx.framebuffer_destroy(fb)

    try:
        configure(ctx, lo=(0, 0, 0), hi=(0, 0, 0))  # bounds from the scene
        set_params(ctx, mlp, grid)
        a = draw(A.JOB_USE_CACHE)
        b = draw(A.JOB_USE_CACHE)
        ctx.set_tuning(queue_capacity=1024)
        c = draw(A.JOB_USE_CACHE)
        ctx.set_tuning(queue_capacity=old.queue_capacity)
        p = draw(0)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
        assert not np.array_equal(a.view(np.uint32), p.view(np.uint32))
    finally:
        ctx.set_tuning(queue_capacity=old.queue_capacity)
        ctx.scene_destroy(scene)


def test_cache_beats_plain_path_tracing_at_equal_samples(ctx):
    """Cornell box + soup: 32 frames of 1 spp with the cache (trained every
    frame on paths the tracer generates) against 32 spp of plain path
    tracing, both measured against a 1024-spp plain render in FP64."""
    g = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    scene = ctx.scene_create(g.desc)
    w = h = 96
    try:
        ref_fb = ctx.framebuffer_create(w, h, fp64=True)
        ctx.render([pt.make_job(scene, ref_fb, g.camera, 1024, 6, seed=99)])
        ref = ctx.framebuffer_read(ref_fb, w, h, fp64=True)[0]
        ctx.framebuffer_destroy(ref_fb)
        plain_fb = ctx.framebuffer_create(w, h)
        ctx.render([pt.make_job(scene, plain_fb, g.camera, 32, 6, seed=5)])
        plain = ctx.framebuffer_read(plain_fb, w, h)[0]
        ctx.framebuffer_destroy(plain_fb)
        configure(ctx, lo=(0, 0, 0), hi=(0, 0, 0), train_paths=1 << 14, train_depth=5)  # bounds from the scene
        cache_fb = ctx.framebuffer_create(w, h)
        for f in range(32):
            ctx.render([pt.make_job(scene, cache_fb, g.camera, 1, 6, seed=5, sample_offset=f,
                                    flags=A.JOB_USE_CACHE | A.JOB_TRAIN_CACHE)])
        cached = ctx.framebuffer_read(cache_fb, w, h)[0]
        ctx.framebuffer_destroy(cache_fb)
        e_plain, e_cache = rel_mse(plain, ref), rel_mse(cached, ref)
        print(f"relMSE plain 32 spp {e_plain:.4e}, cache 32 frames {e_cache:.4e}", flush=True)
        assert np.all(np.isfinite(cached))
        assert e_cache < e_plain
    finally:
        ctx.scene_destroy(scene)


def test_next_frame_reads_the_weights_after_deferred_training(ctx, params):
    """A frame's training runs after its image, on its own stream, and the
    next frame waits for it where it first reads the cache.  So a cached
    frame drawn right after a training frame must equal, bit for bit, the
    same frame drawn with the trained weights read back and set again."""
    mlp, grid, _, _ = params
    g = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    scene = ctx.scene_create(g.desc)
    w, h = 64, 48
    fb = ctx.framebuffer_create(w, h)

    def draw(flags, offset):
        ctx.framebuffer_clear(fb)
        ctx.render([pt.make_job(scene, fb, g.camera, 1, 6, seed=11, sample_offset=offset, flags=flags)])
        return ctx.framebuffer_read(fb, w, h)[0]

    try:
        configure(ctx, lo=(0, 0, 0), hi=(0, 0, 0), train_paths=1 << 14)  # bounds from the scene
        set_params(ctx, mlp, grid)
        draw(A.JOB_USE_CACHE | A.JOB_TRAIN_CACHE, 0)
        after = draw(A.JOB_USE_CACHE, 1)
        mlp1 = np.zeros_like(mlp)
        grid1 = np.zeros_like(grid)
        ctx._ck(ctx.lib.fnd_pt_cache_get_params(ctx.h, mlp1.ctypes.data, grid1.ctypes.data))
        assert not np.array_equal(mlp1, mlp)
        set_params(ctx, mlp1, grid1)
        again = draw(A.JOB_USE_CACHE, 1)
        assert np.all(np.isfinite(after))
        assert np.array_equal(after.view(np.uint32), again.view(np.uint32))
    finally:
        ctx.framebuffer_destroy(fb)
        ctx.scene_destroy(scene)
