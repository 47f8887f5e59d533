"""Autotuner and the multi-scene batch workload.

CPU: the batch generator's triangle counts are log-uniform in [10k, 2M] and
its image sides come from {256, 512, 1024}, deterministically per index.
GPU: the tuned configuration is no slower than the first round's median on
a held-out run, the cache file round-trips, the caller's framebuffers are
untouched, and a batch of batch-generator scenes renders bit-identically to
each scene alone under two different launch configurations.
"""
import os

import numpy as np
import pytest

from foundation_amd import _abi as A
from foundation_amd import pt


def test_batch_generator_distribution():
    counts, sides = [], []
    for i in range(64):
        g, side = pt.batch_scene(i)
        n = 20 * g.params.subdivision ** 2 + g.params.soup_triangles
        counts.append(n)
        sides.append(side)
        assert g.params.soup_triangles == n // 8 or abs(g.params.soup_triangles - n // 8) <= n // 50
        g.close()
    counts = np.array(counts)
    assert counts.min() >= 9000 and counts.max() <= 2_100_000
    # log-uniform: the logs spread over the whole range, about evenly
    u = (np.log(counts) - np.log(1e4)) / (np.log(2e6) - np.log(1e4))
    assert u.min() < 0.15 and u.max() > 0.85 and 0.35 < u.mean() < 0.65
    assert set(sides) == {256, 512, 1024}
    g2, side2 = pt.batch_scene(5)
    assert side2 == sides[5] and 20 * g2.params.subdivision ** 2 + g2.params.soup_triangles == counts[5]


def _small_batch(ctx, indices, side, spp, depth):
    scenes, fbs, jobs = [], [], []
    for i in indices:
        g, _ = pt.batch_scene(i)
        scenes.append(ctx.scene_create(g.desc))
        fbs.append(ctx.framebuffer_create(side, side))
        jobs.append(pt.make_job(scenes[-1], fbs[-1], g.camera, spp, depth, seed=i))
        g.close()
    return scenes, fbs, jobs


def _free(ctx, scenes, fbs):
    for f in fbs:
        ctx.framebuffer_destroy(f)
    for s in scenes:
        ctx.scene_destroy(s)


def _time(ctx, jobs, tuning, offset, reps=5):
    ctx.set_tuning(**{k: getattr(tuning, k) for k, _ in tuning._fields_})
    ms = []
    for r in range(reps + 1):
        for j in jobs:
            j.sample_offset = offset + r
        ctx.render(jobs)
        if r:
            ms.append(ctx.stats().last_call_ms)
    return float(np.median(ms))


@pytest.mark.gpu
def test_autotune_held_out_and_cache_round_trip(ctx, tmp_path):
    scenes, fbs, jobs = _small_batch(ctx, [1, 3, 6, 9], 192, 1, 4)
    cache = str(tmp_path / "tune.txt")
    try:
        ctx.set_tuning(block_size=0, stack_lds=0, waves_per_cu=0, refill=0, queue_capacity=0, tile=0)
        r = ctx.autotune(jobs, cache_path=cache, force=True)
        print(f"key {r.key.decode()}: {r.candidates} configs, {r.renders} renders, best {r.best_ms:.3f} ms, "
              f"median {r.median_ms:.3f} ms, default {r.default_ms:.3f} ms", flush=True)
        assert r.candidates >= 20 and r.renders > r.candidates and not r.from_cache
        t = ctx.tuning()
        assert all(getattr(t, k) == getattr(r.tuning, k) for k, _ in t._fields_)
        # the callers' framebuffers were not written
        for f in fbs:
            img, n = ctx.framebuffer_read(f, 192, 192)
            assert n == 0 and not img.any()
        # held out: other samples, repeated timings
        tuned = _time(ctx, jobs, r.tuning, 1000)
        median = _time(ctx, jobs, r.median, 1000)
        print(f"held out: tuned {tuned:.3f} ms, median config {median:.3f} ms", flush=True)
        assert tuned <= 1.10 * median
        # the cache round-trips
        ctx.set_tuning(block_size=0, stack_lds=0, waves_per_cu=0, refill=0, queue_capacity=0, tile=0)
        r2 = ctx.autotune(jobs, cache_path=cache)
        assert r2.from_cache and r2.key == r.key
        assert all(getattr(r2.tuning, k) == getattr(r.tuning, k) for k, _ in t._fields_)
        lines = open(cache).read().split("\n")
        assert sum(1 for ln in lines if ln.startswith(r.key.decode() + " ")) == 1
    finally:
        ctx.set_tuning(block_size=0, stack_lds=0, waves_per_cu=0, refill=0, queue_capacity=0, tile=0)
        _free(ctx, scenes, fbs)


@pytest.mark.gpu
def test_batch_scenes_equal_alone_under_two_configs(ctx):
    idx = [1, 4, 7]
    side = 96
    out = {}
    try:
        for tag, tun in (("default", {}), ("other", dict(block_size=64, stack_lds=32, waves_per_cu=12, refill=8,
                                                          queue_capacity=1 << 20, tile=4))):
            ctx.set_tuning(**{"block_size": 0, "stack_lds": 0, "waves_per_cu": 0, "refill": 0, "queue_capacity": 0,
                              "tile": 0, **tun})
            scenes, fbs, jobs = _small_batch(ctx, idx, side, 3, 5)
            ctx.render(jobs)
            batched = [ctx.framebuffer_read(f, side, side)[0] for f in fbs]
            alone = []
            for j, f in zip(jobs, fbs):
                ctx.framebuffer_clear(f)
                ctx.render([j])
                alone.append(ctx.framebuffer_read(f, side, side)[0])
            _free(ctx, scenes, fbs)
            for a, b in zip(batched, alone):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
                assert a.mean() > 0
            out[tag] = batched
        for a, b in zip(out["default"], out["other"]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        ctx.set_tuning(block_size=0, stack_lds=0, waves_per_cu=0, refill=0, queue_capacity=0, tile=0)
