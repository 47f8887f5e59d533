"""GPU traversal parity: libfnd_pt.so closest-hit / any-hit vs the CPU oracle.

Closest-hit records (t, prim, u, v) must match the oracle bit for bit: both
sides evaluate the same FP32 expressions (explicit fma, no contraction) and
resolve equal distances by the smallest primitive id, so the answer does not
depend on the tree.  The tolerance for ties is zero; a mismatch is reported
with the oracle's own distance to the GPU's primitive so a near-tie can be
told apart from a wrong answer.
"""
import numpy as np
import pytest

from foundation_amd import pt
from foundation_amd import rays as R

pytestmark = pytest.mark.gpu

NO_HIT = 0xFFFFFFFF


def _words(h):
    return np.ascontiguousarray(h).view(np.uint32)


def _assert_same(gpu, ref, what):
    g, r = _words(gpu), _words(ref)
    bad = np.nonzero(np.any(g != r, axis=1))[0]
    if bad.size:
        i = bad[0]
        raise AssertionError(
            f"{what}: {bad.size}/{g.shape[0]} records differ; first ray {i}: gpu t={gpu[i,0]!r} prim={g[i,1]} "
            f"u={gpu[i,2]!r} v={gpu[i,3]!r} / oracle t={ref[i,0]!r} prim={r[i,1]} u={ref[i,2]!r} v={ref[i,3]!r}")


def _occluded(h):
    return _words(h)[:, 1] != NO_HIT


@pytest.fixture(scope="module")
def cornell(ctx, oracle):
    gen = pt.GeneratedScene(1)
    scene = ctx.scene_create(gen.desc)
    ref = oracle.scene_from_generated(gen)
    yield gen, scene, ref
    ctx.scene_destroy(scene)


@pytest.fixture(scope="module")
def instanced(ctx, oracle):
    gen = pt.GeneratedScene(3, instances=24, subdivision=16)
    scene = ctx.scene_create(gen.desc)
    ref = oracle.scene_from_generated(gen)
    yield gen, scene, ref
    ctx.scene_destroy(scene)


def _ray_sets(gen, ref, w, h, light):
    prim = R.primary_rays(gen.camera, w, h)
    ph = ref.trace(prim)
    wt = gen.world_triangles()
    bounce, _ = R.bounce_rays(prim, ph, wt, seed=1)
    shadow, _ = R.shadow_rays(prim, ph, wt, light)
    return prim, bounce, shadow


@pytest.mark.parametrize("which", ["cornell", "instanced"])
def test_closest_and_anyhit_match_oracle(ctx, request, which):
    gen, scene, ref = request.getfixturevalue(which)
    light = [0.0, 0.0, 0.9] if which == "cornell" else [0.0, 0.0, 9.0]
    for name, rays in zip(("primary", "bounce", "shadow"), _ray_sets(gen, ref, 160, 120, light)):
        print(f"[{which}] {name}: {rays.shape[0]} rays", flush=True)
        _assert_same(ctx.trace(scene, rays), ref.trace(rays), f"{which}/{name} closest")
        assert np.array_equal(_occluded(ctx.trace(scene, rays, anyhit=True)), _occluded(ref.trace(rays, anyhit=True))), \
            f"{which}/{name} any-hit"


@pytest.mark.parametrize("block,stack", [(64, 16), (64, 32), (128, 8), (128, 32), (256, 8), (256, 16)])
@pytest.mark.parametrize("refill", [1, 16, 64])
def test_launch_shapes_give_identical_records(ctx, instanced, block, stack, refill):
    gen, scene, ref = instanced
    _, bounce, _ = _ray_sets(gen, ref, 96, 64, [0, 0, 9])
    base = ref.trace(bounce)
    old = ctx.tuning()
    try:
        ctx.set_tuning(block_size=block, stack_lds=stack, refill=refill, waves_per_cu=8)
        _assert_same(ctx.trace(scene, bounce), base, f"block {block} stack {stack} refill {refill}")
    finally:
        ctx.set_tuning(block_size=old.block_size, stack_lds=old.stack_lds, refill=old.refill,
                       waves_per_cu=old.waves_per_cu)


def test_counts_match_host_transcription(ctx, instanced, oracle):
    gen, scene, ref = instanced
    _, bounce, _ = _ray_sets(gen, ref, 96, 64, [0, 0, 9])
    enc = pt.encode_scene(gen.desc)
    _, cnt = oracle.trace_wide(enc, bounce)
    ctx.reset_stats()
    ctx.trace(scene, bounce, count=True)
    s = ctx.stats()
    assert s.node_visits == int(cnt[:, 0].sum())
    assert s.tri_tests == int(cnt[:, 1].sum())
    assert s.instance_visits == int(cnt[:, 2].sum())


def test_export_matches_host_encoding(ctx, instanced):
    import ctypes as C
    gen, scene, _ = instanced
    enc = pt.encode_scene(gen.desc)
    nb, tb = C.c_uint64(0), C.c_uint64(0)
    assert ctx.lib.fnd_pt_scene_export(ctx.h, scene, None, C.byref(nb), None, C.byref(tb)) == 0
    nodes = np.zeros(nb.value, np.uint8)
    tris = np.zeros(tb.value, np.uint8)
    assert ctx.lib.fnd_pt_scene_export(ctx.h, scene, nodes.ctypes.data, C.byref(nb), tris.ctypes.data, C.byref(tb)) == 0
    assert nodes.tobytes() == enc["nodes"].tobytes()
    assert tris.tobytes() == enc["tris"].tobytes()


def test_degenerate_rays(ctx, cornell):
    gen, scene, ref = cornell
    rays = R.primary_rays(gen.camera, 8, 8)
    rays[0, 4:7] = np.nan          # NaN direction: no hit
    rays[1, 4:7] = 0.0             # zero direction: no hit
    rays[2, 3], rays[2, 7] = 5.0, 1.0   # empty interval
    rays[3, 4:7] = [0.0, 1.0, 0.0]      # axis-aligned
    rays[4, 4:7] = [0.0, 0.0, -1.0]
    g = ctx.trace(scene, rays)
    _assert_same(g, ref.trace(rays), "degenerate rays")
    assert (_words(g)[:3, 1] == NO_HIT).all()
    assert ctx.trace(scene, rays[:0]).shape == (0, 4)


@pytest.mark.parametrize("which", ["cornell", "instanced"])
def test_refused_rays_are_misses(ctx, request, which):
    """A ray with a non-finite origin or direction component, or a NaN range
    bound, is answered as a miss (t = tmax as given, no primitive) without a
    walk, as in the oracle; whole waves of such rays (the first 300) leave the
    other rays' records untouched."""
    gen, scene, ref = request.getfixturevalue(which)
    rays = R.primary_rays(gen.camera, 32, 32)
    bad = rays.copy()
    bad[:300, 4] = np.nan
    bad[300:310, 0] = np.inf
    bad[310:320, 7] = np.nan
    bad[320:330, 5] = -np.inf
    bad[330:340, 3] = np.nan
    bad[340:350, 2] = np.nan
    refused = np.zeros(len(bad), bool)
    refused[:350] = True
    g = ctx.trace(scene, bad)
    _assert_same(g, ref.trace(bad), f"{which} refused rays")
    assert np.all(_words(g)[refused, 1] == NO_HIT)
    assert np.array_equal(_words(g)[refused, 0], bad[refused, 7].view(np.uint32))
    clean = ctx.trace(scene, rays)
    assert np.array_equal(_words(g)[~refused], _words(clean)[~refused])
    occ = _occluded(ctx.trace(scene, bad, anyhit=True))
    assert not occ[refused].any()
    assert np.array_equal(occ, _occluded(ref.trace(bad, anyhit=True)))


def test_repeat_is_bitwise_identical(ctx, instanced):
    gen, scene, ref = instanced
    _, bounce, _ = _ray_sets(gen, ref, 96, 64, [0, 0, 9])
    a = ctx.trace(scene, bounce)
    b = ctx.trace(scene, bounce)
    assert np.array_equal(_words(a), _words(b))


def test_device_buffers_path(ctx, cornell):
    gen, scene, ref = cornell
    rays = R.primary_rays(gen.camera, 64, 64)
    n = rays.shape[0]
    rb = ctx.buffer_create(n * 32)
    hb = ctx.buffer_create(n * 16)
    try:
        ctx.buffer_write(rb, rays)
        ctx.trace_buffers(scene, rb, hb, n)
        out = ctx.buffer_read(hb, np.zeros((n, 4), np.float32))
        _assert_same(out, ref.trace(rays), "device buffers")
        assert ctx.stats().last_kernel_ms > 0
    finally:
        ctx.buffer_destroy(rb)
        ctx.buffer_destroy(hb)


def test_spatially_ordered_batches_match_oracle(ctx, instanced):
    """Batches of at least 131072 rays are taken in the order of their
    origins' cells of a coarse grid over the scene (ray_order.hip); every
    record must still be the ray's own.  Incoherent rays (bounces, with
    refused rays and origins far outside the scene's bounds mixed in) use the
    order; camera rays from one eye fall in one cell and keep the caller's
    order.  Closest-hit records bit for bit, any-hit occlusion."""
    gen, scene, ref = instanced
    _, bounce, _ = _ray_sets(gen, ref, 256, 192, [0, 0, 9])
    reps = -(-140000 // bounce.shape[0])
    rays = np.concatenate([bounce] * reps)
    rng = np.random.default_rng(5)
    idx = rng.choice(rays.shape[0], 2000, replace=False)
    rays[idx[:500], 0] = np.nan                 # refused
    rays[idx[500:1000], 0:3] += 1e4             # far outside the bounds: border cells
    rays[idx[1000:1500], 2] = -np.inf           # refused
    rays[idx[1500:], 4:7] *= -1.0
    assert rays.shape[0] >= 131072
    _assert_same(ctx.trace(scene, rays), ref.trace(rays), "ordered incoherent batch")
    assert np.array_equal(_occluded(ctx.trace(scene, rays, anyhit=True)), _occluded(ref.trace(rays, anyhit=True)))
    camera = R.primary_rays(gen.camera, 512, 288)
    assert camera.shape[0] >= 131072
    _assert_same(ctx.trace(scene, camera), ref.trace(camera), "camera batch")


def test_back_to_back_async_ordered_batches(ctx, instanced):
    """Asynchronous ray queries of 131072 rays or more make their order on a
    second stream while the previous batch traces, in two slots used in turn:
    three different batches queued back to back, then one wait, give the same
    records as each batch traced alone."""
    from foundation_amd import _abi as A
    gen, scene, ref = instanced
    prim = R.primary_rays(gen.camera, 256, 192)
    ph = ref.trace(prim)
    sets = []
    for seed in (1, 2, 3):
        b, _ = R.bounce_rays(prim, ph, gen.world_triangles(), seed=seed)
        sets.append(np.ascontiguousarray(np.concatenate([b] * -(-140000 // b.shape[0]))))
    bufs = []
    try:
        for rays in sets:
            n = rays.shape[0]
            rb, hb = ctx.buffer_create(n * 32), ctx.buffer_create(n * 16)
            ctx.buffer_write(rb, rays)
            bufs.append((rb, hb, n))
        for rb, hb, n in bufs:
            ctx.trace_buffers(scene, rb, hb, n, A.ASYNC)
        ctx.wait()
        for (rb, hb, n), rays in zip(bufs, sets):
            got = ctx.buffer_read(hb, np.zeros((n, 4), np.float32))
            _assert_same(got, ctx.trace(scene, rays), "async ordered batch")
    finally:
        for rb, hb, _ in bufs:
            ctx.buffer_destroy(rb)
            ctx.buffer_destroy(hb)
