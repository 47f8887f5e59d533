"""The C++ host side (csrc/PathTracerPass.hpp) through its headless driver.

CPU: the driver is built beside the library and refuses a bad command line.
GPU: frames rendered by PathTracerPass::Execute are bit-identical to the same
jobs submitted through the Python wrapper, and every byte the pass took from
the borrowed host allocator is returned (the driver exits 3 otherwise).
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "foundation_amd", "pt_headless")


def test_driver_is_built_and_checks_its_arguments():
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "usage" in r.stderr


@pytest.mark.gpu
def test_pass_matches_python_jobs_bitwise(ctx, tmp_path):
    from foundation_amd import _abi as A
    from foundation_amd import pt

    w, h, frames, spp, depth, rr, seed = 64, 48, 3, 2, 5, 3, 11
    out = tmp_path / "img.f32"
    r = subprocess.run([DRIVER, "1", str(w), str(h), str(frames), str(spp), str(depth), str(rr), str(seed), "0",
                        str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    stats = json.loads(r.stdout.strip().splitlines()[-1])
    assert stats["host_alloc_calls"] > 0 and stats["host_live_bytes_after"] == 0
    got = np.fromfile(out, np.float32).reshape(h, w, 3)

    g = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    scene = ctx.scene_create(g.desc)
    fb = ctx.framebuffer_create(w, h)
    try:
        for f in range(frames):
            ctx.render([pt.make_job(scene, fb, g.camera, spp, depth, rr_depth=rr, seed=seed, sample_offset=f * spp)])
        ref, n = ctx.framebuffer_read(fb, w, h)
    finally:
        ctx.framebuffer_destroy(fb)
        ctx.scene_destroy(scene)
    assert n == frames * spp
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
