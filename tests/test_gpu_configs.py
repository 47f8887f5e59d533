"""The benchmark configurations at their stated workloads, on the GPU.

* C2 traversal microbenchmark: the 1.31M-triangle displaced icosphere +
  200k soup; every 1920x1080 coherent primary ray, every incoherent cosine
  bounce ray from the first hits (PCG32 seed 1, stream = pixel) and every
  shadow ray to the first light's centre: closest-hit records bit-exact and
  any-hit occlusion equal, GPU vs oracle.
* C4: all 64 scenes of the batch generator (triangles log-uniform in
  [10k, 2M], square images of side 256, 512 or 1024), 16 spp, depth 6, with
  launch knobs from the autotuner: the one batched call gives each scene
  the image it gets rendered alone, bit for bit; one of them is also checked
  against the oracle bit for bit.
* C5: the radiance cache on the C2 geometry (4 quad lights, 50/50
  Lambert/GGX): 32 one-sample frames with the cache, trained every frame on
  the frame's own paths, land closer to a 1024-spp FP64 reference than 32
  samples of plain path tracing.
(C1 and C3 are compared with the oracle in test_gpu_parity.py.)
"""
import ctypes as C

import numpy as np
import pytest

from foundation_amd import _abi as A
from foundation_amd import pt
from foundation_amd import rays as R
from foundation_amd.metrics import rel_mse

pytestmark = pytest.mark.gpu


def _words(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_c2_full_ray_sets_match_oracle(ctx, oracle):
    gen = pt.GeneratedScene(A.GEN_DISPLACED_SPHERE)
    scene = ctx.scene_create(gen.desc)
    try:
        ref = oracle.scene_from_generated(gen)
        primary = R.primary_rays(gen.camera, 1920, 1080)
        ph = ctx.trace(scene, primary)
        oph = ref.trace(primary)
        assert np.array_equal(_words(ph), _words(oph)), int(np.any(_words(ph) != _words(oph), axis=1).sum())
        lookup = gen.triangle_lookup()
        incoherent, _ = R.bounce_rays(primary, ph, lookup, seed=1)
        lt = gen.desc.lights[0]
        centre = [lt.corner[k] + 0.5 * (lt.edge0[k] + lt.edge1[k]) for k in range(3)]
        shadow, _ = R.shadow_rays(primary, ph, lookup, centre)
        print(f"C2: {primary.shape[0]} primary, {incoherent.shape[0]} incoherent, {shadow.shape[0]} shadow rays",
              flush=True)
        assert incoherent.shape[0] > 1_000_000
        g, o = ctx.trace(scene, incoherent), ref.trace(incoherent)
        bad = int(np.any(_words(g) != _words(o), axis=1).sum())
        assert bad == 0, f"{bad} incoherent records differ"
        occ = _words(ctx.trace(scene, shadow, anyhit=True))[:, 1] != A.NO_HIT
        oocc = _words(ref.trace(shadow, anyhit=True))[:, 1] != A.NO_HIT
        assert np.array_equal(occ, oocc), int((occ != oocc).sum())
        assert 0.05 < occ.mean() < 0.95
        ref.close()
    finally:
        ctx.scene_destroy(scene)


def test_c4_batched_64_scenes_equal_alone(ctx, oracle):
    scenes, fbs, jobs, sizes = [], [], [], []
    kept = None  # (index, generated scene) of the first 256^2 scene, for the oracle
    old = ctx.tuning()
    try:
        tris = 0
        for i in range(64):
            g, side = pt.batch_scene(i)
            scenes.append(ctx.scene_create(g.desc))
            fbs.append(ctx.framebuffer_create(side, side))
            jobs.append(pt.make_job(scenes[-1], fbs[-1], g.camera, 16, 6, seed=i))
            sizes.append(side)
            tris += g.num_prims()
            if kept is None and side == 256:
                kept = (i, g)
            else:
                g.close()
        print(f"C4: 64 scenes, {tris} triangles, {sum(s * s for s in sizes)} pixels", flush=True)
        assert 10_000_000 < tris < 60_000_000 and set(sizes) == {256, 512, 1024}
        res = ctx.autotune(jobs, force=True)
        print(f"C4 autotune: {res.candidates} configurations, best {res.best_ms:.1f} ms", flush=True)
        ctx.render(jobs)
        batched = [ctx.framebuffer_read(f, s, s)[0] for f, s in zip(fbs, sizes)]
        for k, (f, s) in enumerate(zip(fbs, sizes)):
            ctx.framebuffer_clear(f)
            ctx.render([jobs[k]])
            alone = ctx.framebuffer_read(f, s, s)[0]
            assert np.array_equal(_words(alone), _words(batched[k])), k
        # one 256^2 scene of the batch against the oracle, bit for bit
        k, gk = kept
        osc = oracle.scene_from_generated(gk, shading=True)
        acc, _, _ = osc.render(oracle.camera10(gk.camera), 256, 256, 16, 6, seed=k)
        refimg = acc * (np.float32(1.0) / np.float32(16))
        assert np.array_equal(_words(batched[k]), _words(refimg)), int(np.any(_words(batched[k]) != _words(refimg),
                                                                             axis=-1).sum())
    finally:
        ctx.set_tuning(block_size=old.block_size, stack_lds=old.stack_lds, waves_per_cu=old.waves_per_cu,
                       refill=old.refill, queue_capacity=old.queue_capacity, tile=old.tile)
        for f in fbs:
            ctx.framebuffer_destroy(f)
        for s in scenes:
            ctx.scene_destroy(s)


def test_c5_cache_beats_plain_on_the_c2_scene(ctx):
    g = pt.GeneratedScene(A.GEN_DISPLACED_SPHERE)
    scene = ctx.scene_create(g.desc)
    w, h, depth, rr = 640, 360, 8, 3
    try:
        ref_fb = ctx.framebuffer_create(w, h, fp64=True)
        ctx.render([pt.make_job(scene, ref_fb, g.camera, 1024, depth, rr_depth=rr, seed=99)])
        ref = ctx.framebuffer_read(ref_fb, w, h, fp64=True)[0]
        ctx.framebuffer_destroy(ref_fb)
        plain_fb = ctx.framebuffer_create(w, h)
        ctx.render([pt.make_job(scene, plain_fb, g.camera, 32, depth, rr_depth=rr, seed=5)])
        plain = ctx.framebuffer_read(plain_fb, w, h)[0]
        ctx.framebuffer_destroy(plain_fb)
        d = A.CacheDesc()
        assert ctx.lib.fnd_pt_cache_default_desc(C.byref(d)) == 0
        ctx._ck(ctx.lib.fnd_pt_cache_configure(ctx.h, C.byref(d)))
        cache_fb = ctx.framebuffer_create(w, h)
        for f in range(32):
            ctx.render([pt.make_job(scene, cache_fb, g.camera, 1, depth, rr_depth=rr, seed=5, sample_offset=f,
                                    flags=A.JOB_USE_CACHE | A.JOB_TRAIN_CACHE)])
        cached = ctx.framebuffer_read(cache_fb, w, h)[0]
        ctx.framebuffer_destroy(cache_fb)
        e_plain, e_cache = rel_mse(plain, ref), rel_mse(cached, ref)
        print(f"C5 scene: relMSE plain 32 spp {e_plain:.4e}, cache 32 frames {e_cache:.4e}", flush=True)
        assert np.all(np.isfinite(cached))
        assert e_cache < 0.75 * e_plain
    finally:
        d = A.CacheDesc()
        d.enable = 0
        ctx.lib.fnd_pt_cache_configure(ctx.h, C.byref(d))
        ctx.scene_destroy(scene)
