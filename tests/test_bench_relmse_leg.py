"""The time-to-relMSE leg of bench.py, driven by a fake context on the CPU.

The fake renders nothing: frame f's image is 1 + sqrt(1.01 / f) everywhere
and the reference is 1, so the image's relMSE is exactly 1 / f (eps = 0.01).
The leg must count frames up to the first whose relMSE reaches the target
(and no further), stop at max_frames when the target is not reached, sum
the per-frame device times of the counted frames only, time the extra
back-to-back frames separately, and report whether the device evaluation
agrees with the host evaluation of the last frame.
"""
import os
import sys
import types

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from foundation_amd import pt  # noqa: E402


class _Stats:
    def __init__(self, ms):
        self.last_call_ms = ms


class FakeContext:
    """Frame f (1-based) takes f ms of device time; its relMSE is 1 / f up to
    the FP32 rounding of the image, and the device evaluation is off by
    `device_error` when that is set."""

    def __init__(self, device_error=0.0):
        self.frames = 0
        self.renders = 0
        self.device_error = device_error
        self.h = None
        self.lib = types.SimpleNamespace(fnd_pt_cache_default_desc=lambda *a: 0,
                                         fnd_pt_cache_configure=lambda *a: 0)

    def _ck(self, st):
        pass

    def scene_create(self, desc):
        return 1

    def scene_destroy(self, s):
        pass

    def framebuffer_create(self, w, h, fp64=False):
        if not fp64:
            self.frames = 0
        return 2 if fp64 else 1

    def framebuffer_destroy(self, fb):
        pass

    def render(self, jobs):
        self.frames += 1
        self.renders += 1

    def stats(self):
        return _Stats(float(self.frames))

    def framebuffer_read(self, fb, w, h, fp64=False):
        if fp64:
            return np.ones((h, w, 3)), 0
        x = 1.0 + np.sqrt(1.01 / max(self.frames, 1))
        return np.full((h, w, 3), x, np.float32), self.frames

    def framebuffer_rel_mse(self, fb, ref, eps=0.01):
        # the relMSE of the same FP32 image the readback gives, as the device evaluates it
        x = self.framebuffer_read(fb, 4, 2)[0].astype(np.float64)
        return float(np.mean((x - 1.0) ** 2 / (1.0 + eps))) + self.device_error


@pytest.fixture
def fake_scene(monkeypatch):
    monkeypatch.setattr(pt, "GeneratedScene", lambda kind: types.SimpleNamespace(desc=None, camera=None))
    monkeypatch.setattr(pt, "make_job", lambda *a, **k: None)


def _leg(ctx, target, max_frames, back_to_back=5):
    return bench.time_to_rel_mse(ctx, 16, 8, 8, 3, target, 256, max_frames, eval_threads=2,
                                 back_to_back=back_to_back)


def test_counts_frames_up_to_the_first_that_reaches_the_target(fake_scene):
    ctx = FakeContext()
    out = _leg(ctx, 0.0401, 100)  # 1 / 25 = 0.04 is the first relMSE <= 0.0401
    for mode in ("plain", "cache"):
        r = out[mode]
        assert r["frames"] == 25 and r["reached"]
        assert r["seconds"] == pytest.approx(sum(range(1, 26)) / 1e3)
        assert r["ms_per_frame"] == pytest.approx(13.0)
        assert r["rel_mse"] == pytest.approx(0.04, abs=1e-5)
        assert r["rel_mse_host"] == pytest.approx(0.04, abs=1e-5)
        assert r["rel_mse_device_host_agree"]
        # back-to-back frames: a fresh framebuffer, frames 1..5 of it
        assert r["frames_back_to_back"] == 5
        assert r["ms_per_frame_back_to_back"] == pytest.approx(3.0)
    assert out["speedup_cache_vs_plain"] == pytest.approx(1.0)
    # one reference render (256 spp in chunks of 256), then 25 + 5 frames per mode
    assert ctx.renders == 1 + 2 * 30


def test_stops_at_max_frames_when_the_target_is_not_reached(fake_scene):
    out = _leg(FakeContext(), 0.001, 12, back_to_back=0)
    for mode in ("plain", "cache"):
        r = out[mode]
        assert r["frames"] == 12 and not r["reached"]
        assert "ms_per_frame_back_to_back" not in r
    assert "speedup_cache_vs_plain" not in out


def test_flags_a_device_evaluation_that_disagrees_with_the_host(fake_scene):
    out = _leg(FakeContext(device_error=1e-6), 0.0401, 100)
    assert not out["plain"]["rel_mse_device_host_agree"]
