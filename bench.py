#!/usr/bin/env python3
"""Headline benchmark: Mrays/s of incoherent closest-hit rays at 10M triangles.

Workload (one "step"): one closest-hit launch over the incoherent ray set of
the 10M-triangle scene — 100 instances of distinct 100,820-triangle displaced
icospheres (random rotation, scale in [0.5, 2], placement), two-level BVH,
plus 4 quad lights; the rays are one cosine-weighted bounce ray about the
geometric normal at every primary hit of a 1920x1080 pinhole view (PCG32,
seed 1, stream = pixel).  Rays and hits stay in HBM across the timed steps.

Printed (stdout, one JSON line): value = rays * steps / wall time of the timed
steps; roofline of the traversal kernel from its algorithmic bytes (32 B ray +
16 B hit + 64 B per visited node + 48 B per tested triangle + 64 B per entered
instance, counted by the kernel's counting variant over the same rays) over
its mean launch time from HIP events; the CPU oracle timed on the same rays
on the host cores (cpu_baseline); and hit-record parity GPU vs oracle.

extra.launch_scaling: kernel ms of one launch over the incoherent set and
over four copies of it in one array; the linear fit splits a launch into a
fixed part (start, and the drain while the longest rays finish) and a
steady per-ray rate.

extra.traversal_1_5M: the traversal microbenchmark configuration (1.31M-
triangle displaced icosphere + 200k soup; 1080p coherent, incoherent and
shadow ray sets) with its own hit-record parity against the oracle.

`value` is the first half of the metric (Mrays/s).  The second half is in
extra.time_to_rel_mse_0_01: progressive 1 spp/frame path tracing of the
displaced-sphere scene (1.31M-triangle icosphere + 200k soup, 4 quad lights,
1920x1080, depth 8, RR from 3) until relMSE <= 0.01 against a 4096-spp FP64
GPU reference, plain and with the radiance cache trained every frame; device
seconds of the frames are summed.  extra.path_tracing_10M is samples/s of
full path tracing on the 10M-triangle scene at 1920x1080, 64 spp in one call,
depth 8, RR from 3 (NEE + MIS on the 4 quad lights).  extra.frame_arena_10M: one
1-spp 1080p frame of that scene with its per-frame ray queues carved from the
library's own block and from caller device memory of 1024/256/64/16 MiB
(smaller blocks run smaller batches): ms per frame, batch, and bitwise
equality of the images.  --no-render skips these legs.
extra.batch_64_scenes: 64 scenes of the batch generator (triangles
log-uniform in [10k, 2M], images 256/512/1024 square) rendered at 16 spp,
depth 6, by one fnd_pt_render call with launch knobs from the autotuner
(--no-batch skips it).
Progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_gpu_set(ctx, scene, rays, steps, warmup, flags):
    import numpy as np
    from foundation_amd import _abi as A
    n = rays.shape[0]
    rb = ctx.buffer_create(n * 32)
    hb = ctx.buffer_create(n * 16)
    ctx.buffer_write(rb, rays)
    for _ in range(warmup):
        ctx.trace_buffers(scene, rb, hb, n, flags | A.ASYNC)
    ctx.wait()
    ctx.reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.trace_buffers(scene, rb, hb, n, flags | A.ASYNC)
    ctx.wait()
    wall = time.perf_counter() - t0
    st = ctx.stats()
    kernel_ms = st.trace_kernel_ms_total / max(1, st.trace_launches)
    hits = ctx.buffer_read(hb, np.zeros((n, 4), np.float32))
    # algorithmic traffic from the counting variant over the same rays
    ctx.reset_stats()
    ctx.trace_buffers(scene, rb, hb, n, flags | A.TRACE_COUNT)
    cs = ctx.stats()
    ctx.buffer_destroy(rb)
    ctx.buffer_destroy(hb)
    counts = dict(nodes=cs.node_visits, tris=cs.tri_tests, instances=cs.instance_visits)
    return wall, kernel_ms, hits, counts


def launch_scaling(ctx, scene, rays, copies=4, steps=5):
    """Kernel ms of one launch over the ray set and over `copies` copies of it
    back to back in one array: a linear fit splits a launch into a fixed part
    (start, and the drain while its longest rays finish) and a per-ray part."""
    import numpy as np
    from foundation_amd import _abi as A
    out = {}
    for k in (1, copies):
        rr = np.ascontiguousarray(np.tile(rays, (k, 1)))
        n = rr.shape[0]
        rb = ctx.buffer_create(n * 32)
        hb = ctx.buffer_create(n * 16)
        ctx.buffer_write(rb, rr)
        ctx.trace_buffers(scene, rb, hb, n, A.ASYNC)
        ctx.wait()
        ctx.reset_stats()
        for _ in range(steps):
            ctx.trace_buffers(scene, rb, hb, n, A.ASYNC)
        ctx.wait()
        st = ctx.stats()
        out[k] = st.trace_kernel_ms_total / max(1, st.trace_launches)
        ctx.buffer_destroy(rb)
        ctx.buffer_destroy(hb)
    n1 = rays.shape[0]
    per = (out[copies] - out[1]) / (copies - 1)
    return {"rays": n1, "copies": copies, "ms_1x": round(out[1], 4), f"ms_{copies}x": round(out[copies], 4),
            "mrays_s_1x": round(n1 / out[1] / 1e3, 1), f"mrays_s_{copies}x": round(copies * n1 / out[copies] / 1e3, 1),
            "fixed_ms": round(out[1] - per, 4), "ms_per_ray_set": round(per, 4),
            "steady_mrays_s": round(n1 / per / 1e3, 1)}


def render_throughput(ctx, scene, camera, width, height, spp, depth, rr):
    """Samples/s of full path tracing (NEE + MIS + RR) on a scene already built."""
    from foundation_amd import pt
    fb = ctx.framebuffer_create(width, height)
    ctx.render([pt.make_job(scene, fb, camera, 1, depth, rr_depth=rr, seed=1)])  # warm-up
    ctx.reset_stats()
    t = time.perf_counter()
    ctx.render([pt.make_job(scene, fb, camera, spp, depth, rr_depth=rr, seed=1, sample_offset=1)])
    wall = time.perf_counter() - t
    st = ctx.stats()
    ctx.framebuffer_destroy(fb)
    ms = st.last_call_ms
    return {"workload": f"{width}x{height}, {spp} spp, depth {depth}, RR from {rr}, NEE+MIS on 4 quad lights",
            "msamples_s": round(width * height * spp / ms / 1e3, 2), "ms_per_spp": round(ms / spp, 3),
            "render_mrays_s": round(st.render_rays / ms / 1e3, 1), "wall_ms": round(wall * 1e3, 1)}


def frame_arena_sweep(ctx, scene, camera, width, height, depth, rr, frames=3):
    """The same 1-spp frame with its per-frame ray queues carved from the
    library's own block, then from caller device memory of decreasing size
    (a smaller block runs smaller wavefront batches): device ms per frame
    (median of `frames`), the batch, and whether the image equals the library
    block's bit for bit."""
    import numpy as np
    from foundation_amd import pt
    fb = ctx.framebuffer_create(width, height)
    rows, ref = [], None
    try:
        for mib in (None, 1024, 256, 64, 16):
            mem = None
            if mib is None:
                ctx.set_frame_arena(0, 0)
            else:
                mem = pt.DeviceMemory(mib << 20)
                ctx.set_frame_arena(mem.ptr, mib << 20)
            times = []
            for k in range(frames + 1):  # the first call may grow the library's block
                ctx.framebuffer_clear(fb)
                ctx.render([pt.make_job(scene, fb, camera, 1, depth, rr_depth=rr, seed=1)])
                if k:
                    times.append(ctx.stats().last_call_ms)
            img = ctx.framebuffer_read(fb, width, height)[0]
            if ref is None:
                ref = img
            info = ctx.frame_arena_info()
            ms = sorted(times)[len(times) // 2]
            rows.append({"block": "library" if mib is None else f"caller {mib} MiB",
                         "block_mib": round(info.size / 2**20, 1), "used_mib": round(info.used / 2**20, 1),
                         "batch_paths": int(info.batch_paths), "ms_per_frame": round(ms, 3),
                         "msamples_s": round(width * height / ms / 1e3, 1),
                         "same_image": bool(np.array_equal(img.view(np.uint32), ref.view(np.uint32)))})
            ctx.set_frame_arena(0, 0)
            if mem is not None:
                mem.free()
    finally:
        ctx.set_frame_arena(0, 0)
        ctx.framebuffer_destroy(fb)
    return {"workload": f"{width}x{height}, 1 spp, depth {depth}, RR from {rr}, 10.08M tris",
            "queue_bytes_per_path": int(pt.queue_bytes(1 << 16) - pt.queue_bytes(1 << 15)) >> 15, "rows": rows}


def traversal_1_5m(ctx, steps, warmup, threads, with_oracle):
    """The traversal microbenchmark configuration: a 1.31M-triangle displaced
    icosphere (8 subdivisions, 3-octave value-noise displacement of amplitude
    0.1) plus 200k random-soup triangles, single-level BVH; 1920x1080 coherent
    primary rays, the incoherent cosine bounce rays from their first hits, and
    shadow rays from those hits to the centre of the first quad light.  GPU
    Mrays/s per set (closest hit for primary and incoherent, any hit for
    shadow) and, with the oracle, hit-record parity on every ray."""
    import numpy as np
    from foundation_amd import _abi as A
    from foundation_amd import pt
    from foundation_amd import rays as R
    gen = pt.GeneratedScene(A.GEN_DISPLACED_SPHERE)
    scene = ctx.scene_create(gen.desc)
    try:
        info = ctx.scene_info(scene)
        primary = R.primary_rays(gen.camera, 1920, 1080)
        prim_hits = ctx.trace(scene, primary)
        lookup = gen.triangle_lookup()
        incoherent, _ = R.bounce_rays(primary, prim_hits, lookup, seed=1)
        lt = gen.desc.lights[0]
        centre = [lt.corner[k] + 0.5 * (lt.edge0[k] + lt.edge1[k]) for k in range(3)]
        shadow, _ = R.shadow_rays(primary, prim_hits, lookup, centre)
        out = {"scene": "displaced icosphere (1.31M tris) + 200k soup, single-level BVH",
               "triangles": int(info.num_triangles)}
        hits = {}
        for name, rays, flags in (("coherent", primary, 0), ("incoherent", incoherent, 0),
                                  ("shadow_anyhit", shadow, A.TRACE_ANYHIT)):
            wall, kernel_ms, h, _ = time_gpu_set(ctx, scene, rays, steps, warmup, flags)
            out[f"{name}_rays"] = int(rays.shape[0])
            out[f"{name}_mrays_s"] = round(rays.shape[0] * steps / wall / 1e6, 1)
            hits[name] = h
        if with_oracle:
            from oracle import pt_oracle as O
            oscene = O.scene_from_generated(gen, threads=threads)
            par = {}
            for name, rays in (("coherent", primary), ("incoherent", incoherent)):
                ref = oscene.trace(rays, threads=threads)
                par[f"{name}_record_mismatches"] = int(np.any(hits[name].view(np.uint32) != ref.view(np.uint32),
                                                              axis=1).sum())
            occ = hits["shadow_anyhit"].view(np.uint32)[:, 1] != 0xFFFFFFFF
            oocc = oscene.trace(shadow, anyhit=True, threads=threads).view(np.uint32)[:, 1] != 0xFFFFFFFF
            par["anyhit_mismatches"] = int((occ != oocc).sum())
            out["parity"] = par
        return out
    finally:
        ctx.scene_destroy(scene)


def batch64(ctx, count, spp, depth, cache_path):
    """The multi-scene workload: `count` scenes of the batch generator
    (displaced sphere + soup, triangles log-uniform in [10k, 2M], square
    images of side 256/512/1024), all rendered by one fnd_pt_render call at
    `spp` spp and depth `depth`, with launch knobs from the autotuner."""
    from foundation_amd import pt
    t = time.perf_counter()
    gens, scenes, fbs, jobs = [], [], [], []
    tris = pixels = 0
    try:
        for i in range(count):
            g, side = pt.batch_scene(i)
            gens.append(g)
            scenes.append(ctx.scene_create(g.desc))
            fbs.append(ctx.framebuffer_create(side, side))
            jobs.append(pt.make_job(scenes[-1], fbs[-1], g.camera, spp, depth, seed=i))
            tris += g.num_prims()
            pixels += side * side
            g.close()
        build_s = time.perf_counter() - t
        log(f"batch: {count} scenes, {tris} tris, {pixels} pixels built in {build_s:.1f} s")
        ctx.set_tuning(block_size=0, stack_lds=0, waves_per_cu=0, refill=0, queue_capacity=0, tile=0)
        t = time.perf_counter()
        res = ctx.autotune(jobs, cache_path=cache_path, force=cache_path is None)
        tune_s = time.perf_counter() - t
        tun = res.tuning
        log(f"batch autotune: {res.candidates} configs, {res.renders} renders in {tune_s:.1f} s; best {res.best_ms:.1f} ms"
            f" vs default {res.default_ms:.1f} ms (1 spp)")
        ctx.reset_stats()
        ctx.render(jobs)
        st = ctx.stats()
        ms = st.last_call_ms
        # Held-out check of the tuner: the full workload (every scene at spp
        # samples) on samples the search never rendered, under the tuned
        # configuration, the library default and the first round's median.
        held = {}
        for name, t in (("tuned", tun), ("default", None), ("median", res.median)):
            if t is None:
                ctx.set_tuning(block_size=0, stack_lds=0, waves_per_cu=0, refill=0, queue_capacity=0, tile=0)
            else:
                ctx.set_tuning(**{k: getattr(t, k) for k, _ in t._fields_})
            for j in jobs:
                j.sample_offset = spp
            ctx.render(jobs)
            held[f"{name}_ms"] = round(ctx.stats().last_call_ms, 2)
        held["tuned_vs_default"] = round(held["default_ms"] / held["tuned_ms"], 3)
        held["tuned_vs_median"] = round(held["median_ms"] / held["tuned_ms"], 3)
        log(f"batch held-out: {held}")
        return {"scenes": count, "triangles": int(tris), "pixels": int(pixels), "spp": spp, "max_depth": depth,
                "paths": int(pixels * spp), "build_s": round(build_s, 1), "seconds": round(ms / 1e3, 3),
                "msamples_s": round(pixels * spp / ms / 1e3, 1), "render_mrays_s": round(st.render_rays / ms / 1e3, 1),
                "autotune": {"key": res.key.decode(), "configs": res.candidates, "renders": res.renders,
                             "seconds": round(tune_s, 1), "from_cache": bool(res.from_cache),
                             "best_ms_1spp": round(res.best_ms, 2), "default_ms_1spp": round(res.default_ms, 2),
                             "median_ms_1spp": round(res.median_ms, 2),
                             "held_out_full_workload": held,
                             "tuning": {k: getattr(tun, k) for k, _ in tun._fields_}}}
    finally:
        for f in fbs:
            ctx.framebuffer_destroy(f)
        for s_ in scenes:
            ctx.scene_destroy(s_)
        ctx.set_tuning(block_size=0, stack_lds=0, waves_per_cu=0, refill=0, queue_capacity=0, tile=0)


def frames_only(frames, width, height, depth, rr, grid_gradient=0):
    """The frames of the time-to-relMSE leg without its reference: `frames`
    plain frames, then `frames` frames with the radiance cache, 1 spp each;
    prints the mean device ms per frame of each (for kernel-trace profiles)."""
    import ctypes as C
    from foundation_amd import _abi as A
    from foundation_amd import pt
    ctx = pt.Context(0)
    gen = pt.GeneratedScene(A.GEN_DISPLACED_SPHERE)
    scene = ctx.scene_create(gen.desc)
    fb = ctx.framebuffer_create(width, height)
    for mode in ("plain", "cache"):
        flags = 0
        if mode == "cache":
            d = A.CacheDesc()
            ctx.lib.fnd_pt_cache_default_desc(C.byref(d))
            d.grid_gradient = grid_gradient
            ctx._ck(ctx.lib.fnd_pt_cache_configure(ctx.h, C.byref(d)))
            flags = A.JOB_USE_CACHE | A.JOB_TRAIN_CACHE
        ctx.framebuffer_clear(fb)
        ms = 0.0
        for f in range(frames):
            ctx.render([pt.make_job(scene, fb, gen.camera, 1, depth, rr_depth=rr, seed=5, sample_offset=f, flags=flags)])
            ms += ctx.stats().last_call_ms
        log(f"{mode}: {ms / frames:.3f} ms per frame over {frames} frames")
    ctx.framebuffer_destroy(fb)
    ctx.scene_destroy(scene)
    ctx.close()


def time_to_rel_mse(ctx, width, height, depth, rr, target, ref_spp, max_frames, grid_gradient=0, eval_threads=8,
                    back_to_back=40):
    """Progressive 1 spp/frame rendering of the displaced-sphere scene, plain and
    with the radiance cache (trained every frame on the frame's own paths),
    until relMSE <= target against an FP64 GPU reference of ref_spp samples.
    Device time of the frames is summed (relMSE evaluation excluded).

    Each frame's relMSE is evaluated on the device against the FP64
    reference framebuffer (fnd_pt_framebuffer_rel_mse), so only 8 bytes come
    back between frames: a frame queued after the device sat idle runs slower
    than one queued right behind the previous frame (plain: +0.13 ms after a
    2 ms host pause, +0.08 ms after an image readback).  The last frame of
    each mode is also read back and evaluated on the host (rel_mse_host), as
    a check of the device evaluation.  Each mode also times `back_to_back`
    frames rendered one after another, as a render loop queues them
    (ms_per_frame_back_to_back)."""
    import ctypes as C
    from foundation_amd import _abi as A
    from foundation_amd import pt
    from foundation_amd.metrics import RelMseEvaluator
    gen = pt.GeneratedScene(A.GEN_DISPLACED_SPHERE)
    scene = ctx.scene_create(gen.desc)
    out = {"scene": "displaced icosphere (1.31M tris) + 200k soup, 4 quad lights, 50/50 Lambert/GGX",
           "width": width, "height": height, "max_depth": depth, "rr_depth": rr, "target": target,
           "cache_grid_gradient": "bf16" if grid_gradient == 0 else "fp32"}
    evaluator = None
    ref_fb = None
    try:
        ref_fb = ctx.framebuffer_create(width, height, fp64=True)
        t = time.perf_counter()
        chunk = 256
        for k in range(ref_spp // chunk):
            ctx.render([pt.make_job(scene, ref_fb, gen.camera, chunk, depth, rr_depth=rr, seed=99,
                                    sample_offset=k * chunk)])
        ref = ctx.framebuffer_read(ref_fb, width, height, fp64=True)[0]
        out["reference"] = {"spp": ref_spp, "accumulation": "fp64", "seconds": round(time.perf_counter() - t, 1)}
        log(f"relMSE reference: {ref_spp} spp FP64 in {out['reference']['seconds']} s")
        evaluator = RelMseEvaluator(ref, threads=eval_threads)
        for mode in ("plain", "cache"):
            flags = 0
            if mode == "cache":
                d = A.CacheDesc()
                ctx.lib.fnd_pt_cache_default_desc(C.byref(d))
                d.grid_gradient = grid_gradient
                ctx._ck(ctx.lib.fnd_pt_cache_configure(ctx.h, C.byref(d)))
                flags = A.JOB_USE_CACHE | A.JOB_TRAIN_CACHE
            fb = ctx.framebuffer_create(width, height)
            gpu_ms = 0.0
            wall = 0.0  # host wall clock around each render call (relMSE evaluation excluded)
            e = float("inf")
            f = 0
            while f < max_frames and e > target:
                t0 = time.perf_counter()
                ctx.render([pt.make_job(scene, fb, gen.camera, 1, depth, rr_depth=rr, seed=5, sample_offset=f,
                                        flags=flags)])
                wall += time.perf_counter() - t0
                gpu_ms += ctx.stats().last_call_ms
                f += 1
                e = ctx.framebuffer_rel_mse(fb, ref_fb)
            e_host = evaluator(ctx.framebuffer_read(fb, width, height)[0])
            ctx.framebuffer_destroy(fb)
            out[mode] = {"frames": f, "seconds": round(gpu_ms / 1e3, 3), "ms_per_frame": round(gpu_ms / f, 3),
                         "wall_seconds": round(wall, 3), "rel_mse": round(e, 5), "reached": bool(e <= target),
                         "rel_mse_host": round(e_host, 5), "rel_mse_device_host_agree": bool(abs(e - e_host) <= 1e-9 * e_host)}
            if back_to_back:
                fb = ctx.framebuffer_create(width, height)
                b2b = 0.0
                for k in range(back_to_back):
                    ctx.render([pt.make_job(scene, fb, gen.camera, 1, depth, rr_depth=rr, seed=5, sample_offset=k,
                                            flags=flags)])
                    b2b += ctx.stats().last_call_ms
                ctx.framebuffer_destroy(fb)
                out[mode]["ms_per_frame_back_to_back"] = round(b2b / back_to_back, 3)
                out[mode]["frames_back_to_back"] = back_to_back
            log(f"relMSE {mode}: {e:.4g} after {f} frames, {gpu_ms / 1e3:.2f} s; "
                f"{out[mode].get('ms_per_frame_back_to_back')} ms per frame back to back")
        if out["plain"]["reached"] and out["cache"]["reached"] and out["cache"]["seconds"] > 0:
            out["speedup_cache_vs_plain"] = round(out["plain"]["seconds"] / out["cache"]["seconds"], 3)
            if out["cache"]["wall_seconds"] > 0:
                out["speedup_cache_vs_plain_wall"] = round(out["plain"]["wall_seconds"] / out["cache"]["wall_seconds"], 3)
    finally:
        if evaluator is not None:
            evaluator.close()
        if ref_fb is not None:
            ctx.framebuffer_destroy(ref_fb)
        d = A.CacheDesc()
        d.enable = 0
        ctx.lib.fnd_pt_cache_configure(ctx.h, C.byref(d))
        ctx.scene_destroy(scene)
    return out


CPU_SHARE = 16  # host threads a one-GPU job may use on the GPU box (its CPU share)


def kernel_source_sha():
    """SHA-256 (16 hex digits) of the traversal kernel's sources and of the
    ray order it takes a batch in: a PMC summary made from other sources no
    longer describes the kernel that runs."""
    import hashlib
    h = hashlib.sha256()
    for f in ("trace_kernel.hpp", "pt_device.hpp", "pt_layout.hpp", "trace_launch.hip", "ray_order.hip"):
        with open(os.path.join(ROOT, "foundation_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def host_facts(threads):
    """What the CPU baseline ran on: threads used, CPUs online and in this
    process's affinity mask, and the processor model from /proc/cpuinfo."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"threads": threads, "cpus_online": os.cpu_count(), "cpus_in_affinity": affinity, "model": model,
            "note": f"threads = min({CPU_SHARE}, CPUs in the affinity mask): the CPU share of one GPU on this box"}


def c1_leg(ctx, threads, with_cpu):
    """The CPU configuration C1 in full on both sides: Cornell box + 10,000
    soup triangles, 256x256, 16 spp, depth 4, PCG seed 1.  GPU device time
    and oracle wall time (1 warm-up + median of 3), samples/s of each, and
    whether the two images are equal bit for bit (relMSE between them)."""
    import numpy as np
    from foundation_amd import _abi as A
    from foundation_amd import pt
    from foundation_amd.metrics import rel_mse
    gen = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    scene = ctx.scene_create(gen.desc)
    w = h = 256
    spp, depth = 16, 4
    try:
        fb = ctx.framebuffer_create(w, h)
        times = []
        for _ in range(4):
            ctx.framebuffer_clear(fb)
            ctx.render([pt.make_job(scene, fb, gen.camera, spp, depth, seed=1)])
            times.append(ctx.stats().last_call_ms)
        gpu_img = ctx.framebuffer_read(fb, w, h)[0]
        ctx.framebuffer_destroy(fb)
        gms = statistics.median(times[1:])
        out = {"workload": "Cornell box + 10k soup, 256x256, 16 spp, depth 4, seed 1",
               "gpu_ms": round(gms, 3), "gpu_msamples_s": round(w * h * spp / gms / 1e3, 1)}
        if with_cpu:
            from oracle import pt_oracle as O
            osc = O.scene_from_generated(gen, shading=True, threads=threads)
            cam = O.camera10(gen.camera)
            ctimes = []
            for k in range(4):
                t = time.perf_counter()
                acc, _, rays = osc.render(cam, w, h, spp, depth, seed=1, threads=threads)
                if k:
                    ctimes.append(time.perf_counter() - t)
            cs = statistics.median(ctimes)
            ref = acc * (np.float32(1.0) / np.float32(spp))
            out.update({"cpu_s": round(cs, 3), "cpu_msamples_s": round(w * h * spp / cs / 1e6, 3),
                        "cpu_mrays_s": round(rays / cs / 1e6, 2), "cpu_threads": threads,
                        "gpu_vs_cpu": round((w * h * spp / gms / 1e3) / (w * h * spp / cs / 1e6), 1),
                        "images_bit_exact": bool(np.array_equal(gpu_img.view(np.uint32), ref.view(np.uint32))),
                        "rel_mse_gpu_vs_cpu": float(rel_mse(gpu_img, ref.astype(np.float64)))})
            osc.close()
        return out
    finally:
        ctx.scene_destroy(scene)


def c3_cpu_subset(gen, threads):
    """The oracle's samples/s on a bounded sample of C3: the 10.08M-triangle
    scene at 1 spp, 480x270 (the 1080p view at a quarter of the width and
    height), depth 8, roulette from 3, NEE + MIS on the 4 quad lights."""
    from oracle import pt_oracle as O
    osc = O.scene_from_generated(gen, shading=True, threads=threads)
    cam = O.camera10(gen.camera)
    w, h = 480, 270
    osc.render(cam, 96, 54, 1, 8, rr_depth=3, seed=1, threads=threads)  # warm-up
    t = time.perf_counter()
    _, _, rays = osc.render(cam, w, h, 1, 8, rr_depth=3, seed=1, threads=threads)
    s = time.perf_counter() - t
    osc.close()
    return {"sample": "10.08M-triangle scene, 480x270, 1 spp, depth 8, RR from 3", "seconds": round(s, 3),
            "msamples_s": round(w * h / s / 1e6, 4), "mrays_s": round(rays / s / 1e6, 3), "threads": threads}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--instances", type=int, default=100)
    ap.add_argument("--subdivision", type=int, default=71)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: min(16, CPUs in the affinity mask), the CPU share of one GPU on the box")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity check")
    ap.add_argument("--no-render", action="store_true",
                    help="skip the path-tracing legs (samples/s and time to relMSE 0.01, plain vs cache)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched launch leg")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--autotune-cache", default=None,
                    help="autotuner cache file for the batch leg (default: search every run)")
    ap.add_argument("--relmse-width", type=int, default=1920)
    ap.add_argument("--relmse-height", type=int, default=1080)
    ap.add_argument("--relmse-ref-spp", type=int, default=4096)
    ap.add_argument("--cache-grid-gradient", choices=("bf16", "fp32"), default="bf16",
                    help="how the radiance cache sums its hash-grid gradient (FP32 or packed BF16 atomics)")
    ap.add_argument("--relmse-only", action="store_true",
                    help="run only the time-to-relMSE leg (plain vs cache) and print its result")
    ap.add_argument("--pmc-summary", default=os.path.join(ROOT, "profiles", "trace_incoherent_10M_pmc.json"))
    ap.add_argument("--only-incoherent", action="store_true",
                    help="profiling runs: time only the incoherent set (no coherent/shadow sets, no CPU leg)")
    ap.add_argument("--frames-only", type=int, default=0, metavar="N",
                    help="profiling runs: render N plain then N cache frames of the relMSE scene (no reference, "
                         "no traversal legs), print ms per frame, exit")
    ap.add_argument("--pt-only", action="store_true",
                    help="profiling runs: build the 10M-triangle scene, run only the 64-spp path-tracing leg, "
                         "print its result, exit")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="profiling runs: first copy this many MiB device-to-device (a plain copy kernel of "
                         "known size that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only; the path shards by rays, samples or image tiles with no "
              "exchange inside a step, so N GPUs would each run this same single-GPU workload. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    try:
        cpu_threads = min(CPU_SHARE, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        cpu_threads = min(CPU_SHARE, os.cpu_count() or 1)

    grid_gradient = 0 if args.cache_grid_gradient == "bf16" else 1
    if args.frames_only:
        frames_only(args.frames_only, args.relmse_width, args.relmse_height, 8, 3, grid_gradient)
        return
    if args.relmse_only:
        from foundation_amd import pt
        ctx = pt.Context(0)
        r = time_to_rel_mse(ctx, args.relmse_width, args.relmse_height, 8, 3, 0.01, args.relmse_ref_spp, 1500,
                            grid_gradient, eval_threads=min(8, cpu_threads))
        print(json.dumps(r), flush=True)
        ctx.close()
        return

    if args.calibrate_copy_mib:
        import torch
        n = args.calibrate_copy_mib * 2**20 // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mib} MiB device-to-device")
        del src, dst

    from foundation_amd import _abi as A
    from foundation_amd import pt
    from foundation_amd import rays as R

    t = time.perf_counter()
    gen = pt.GeneratedScene(A.GEN_INSTANCED_SPHERES, instances=args.instances, subdivision=args.subdivision)
    nprims = gen.num_prims()
    log(f"scene generated: {nprims} primitives in {time.perf_counter() - t:.1f}s")
    ctx = pt.Context(0)
    scene = ctx.scene_create(gen.desc)
    info = ctx.scene_info(scene)
    log(f"scene built: {info.num_triangles} tris, {info.num_nodes} nodes, depth {info.max_depth}, "
        f"{info.build_ms:.0f} ms, {info.device_bytes / 2**20:.0f} MiB")
    if args.pt_only:
        log(f"path tracing @10M tris: {render_throughput(ctx, scene, gen.camera, args.width, args.height, 64, 8, 3)}")
        ctx.close()
        return

    primary = R.primary_rays(gen.camera, args.width, args.height)
    prim_hits = ctx.trace(scene, primary)
    lookup = gen.triangle_lookup()
    incoherent, _ = R.bounce_rays(primary, prim_hits, lookup, seed=1)
    # shadow rays toward the centre of the plane that holds the scene's four quad lights (z = 7.5)
    shadow, _ = R.shadow_rays(primary, prim_hits, lookup, [0.0, 0.0, 7.5])
    nrays = incoherent.shape[0]
    log(f"ray sets: {primary.shape[0]} primary, {nrays} incoherent, {shadow.shape[0]} shadow")

    wall, kernel_ms, gpu_hits, counts = time_gpu_set(ctx, scene, incoherent, args.steps, args.warmup, 0)
    value = nrays * args.steps / wall / 1e6
    log(f"incoherent: {value:.1f} Mrays/s, kernel {kernel_ms:.3f} ms/launch, counts {counts}")
    algo_bytes = nrays * (32 + 16) + 64 * counts["nodes"] + 48 * counts["tris"] + 64 * counts["instances"]
    achieved = algo_bytes / (kernel_ms * 1e-3) / 1e9

    coherent = shadow_rate = None
    ccounts = {"nodes": 0}
    gpu_shadow = None
    scaling = None
    if not args.only_incoherent:
        scaling = launch_scaling(ctx, scene, incoherent)
        log(f"launch scaling: {scaling}")
        csteps = max(2, args.steps // 2)
        cwall, ckms, _, ccounts = time_gpu_set(ctx, scene, primary, csteps, 1, 0)
        swall, skms, gpu_shadow, _ = time_gpu_set(ctx, scene, shadow, csteps, 1, A.TRACE_ANYHIT)
        coherent = primary.shape[0] * csteps / cwall / 1e6
        shadow_rate = shadow.shape[0] * csteps / swall / 1e6
        log(f"coherent: {coherent:.1f} Mrays/s, shadow any-hit: {shadow_rate:.1f} Mrays/s")

    # Counter figures come from the committed rocprof PMC summary of this
    # kernel; the summary records the hash of the kernel sources it was
    # measured on, so a summary of an older kernel is reported as stale.
    traffic = None
    valu = None
    pmc_source = None
    if os.path.exists(args.pmc_summary):
        try:
            pmc = json.load(open(args.pmc_summary))
            traffic = pmc.get("hbm_bytes_per_launch")
            v = pmc.get("valu") or {}
            valu = {k: v[k] for k in ("valu_busy", "lane_utilization") if k in v} or None
            if valu and "valu_busy" in valu and "lane_utilization" in valu:
                valu["useful_lane_frac"] = round(valu["valu_busy"] * valu["lane_utilization"], 4)
            cur = kernel_source_sha()
            pmc_source = {"file": os.path.relpath(args.pmc_summary, ROOT),
                          "kernel_source_sha": pmc.get("kernel_source_sha"), "current_kernel_source_sha": cur,
                          "stale": pmc.get("kernel_source_sha") != cur}
            if pmc_source["stale"]:
                log(f"PMC summary {pmc_source['file']} was measured on other kernel sources: traffic and VALU "
                    "figures are reported as stale")
        except (OSError, ValueError):
            traffic = None

    cpu = None
    parity = None
    if not args.no_cpu and not args.only_incoherent:
        from oracle import pt_oracle as O
        threads = args.cpu_threads or cpu_threads
        t = time.perf_counter()
        oscene = O.scene_from_generated(gen, threads=threads)
        log(f"oracle built in {time.perf_counter() - t:.1f}s")
        t = time.perf_counter()
        ref = oscene.trace(incoherent, threads=threads)          # warm-up run, also the parity reference
        first = time.perf_counter() - t
        reps = max(1, min(5, int(25.0 / max(first, 1e-3))))
        times = []
        for _ in range(reps):
            t = time.perf_counter()
            oscene.trace(incoherent, threads=threads)
            times.append(time.perf_counter() - t)
        med = statistics.median(times)
        cpu = {"value": round(nrays / med / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"all {nrays} incoherent rays, binary-BVH oracle, 1 warm-up + median of {reps}",
               "host": host_facts(threads)}
        g = gpu_hits.view(np.uint32)
        r = ref.view(np.uint32)
        mism = int(np.any(g != r, axis=1).sum())
        socc = gpu_shadow.view(np.uint32)[:, 1] != 0xFFFFFFFF
        oocc = oscene.trace(shadow, anyhit=True, threads=threads).view(np.uint32)[:, 1] != 0xFFFFFFFF
        parity = {"incoherent_rays": nrays, "closest_hit_record_mismatches": mism,
                  "shadow_rays": int(shadow.shape[0]), "anyhit_mismatches": int((socc != oocc).sum())}
        log(f"cpu: {cpu['value']} Mrays/s on {threads} threads; parity {parity}")

    micro = None
    if not args.only_incoherent:
        micro = traversal_1_5m(ctx, max(2, args.steps // 2), args.warmup,
                               args.cpu_threads or cpu_threads, not args.no_cpu)
        log(f"traversal 1.5M tris: {micro}")

    c1 = None
    if not args.only_incoherent:
        c1 = c1_leg(ctx, args.cpu_threads or cpu_threads, not args.no_cpu)
        log(f"C1: {c1}")

    render = relmse = batch = arena = None
    if not args.no_render and not args.only_incoherent:
        render = render_throughput(ctx, scene, gen.camera, args.width, args.height, 64, 8, 3)
        if not args.no_cpu:
            render["cpu_subset"] = c3_cpu_subset(gen, args.cpu_threads or cpu_threads)
            render["gpu_vs_cpu_samples"] = round(render["msamples_s"] / render["cpu_subset"]["msamples_s"], 1)
        log(f"path tracing @10M tris: {render}")
        arena = frame_arena_sweep(ctx, scene, gen.camera, args.width, args.height, 8, 3)
        log(f"frame arena: {arena}")
        ctx.scene_destroy(scene)  # free the 10M-triangle scene before the next one
        scene = None
        relmse = time_to_rel_mse(ctx, args.relmse_width, args.relmse_height, 8, 3, 0.01, args.relmse_ref_spp, 1500,
                                 grid_gradient, eval_threads=min(8, cpu_threads))
    if not args.no_batch and not args.only_incoherent:
        if scene is not None:
            ctx.scene_destroy(scene)
            scene = None
        batch = batch64(ctx, args.batch_scenes, 16, 6, args.autotune_cache)
        log(f"batch: {batch}")

    line = {
        "metric": "Mrays/s incoherent @10M tris; time-to-relMSE 0.01 (cache vs plain)",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "incoherent closest-hit rays, 100 instanced displaced spheres (10.08M tris), "
                               "two-level 64-byte wide BVH", "width": args.width, "height": args.height,
                   "rays_per_step": nrays, "triangles": int(nprims)},
        # frac prices the algorithmic bytes against the HBM peak (the bench
        # contract); what binds the kernel is VALU issue with few useful lanes
        # and dependent-fetch latency (the PMC summary), named in `limiter`
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "limiter": "VALU issue (few useful lanes per instruction) and dependent-fetch latency",
                     "useful_lane_frac": valu.get("useful_lane_frac") if valu else None,
                     "pmc_source": pmc_source,
                     "algorithmic_bytes_per_launch": int(algo_bytes), "kernel_ms": round(kernel_ms, 4)},
        "cpu_baseline": cpu,
        "extra": {"coherent_mrays_s": coherent and round(coherent, 2),
                  "shadow_anyhit_mrays_s": shadow_rate and round(shadow_rate, 2),
                  "nodes_per_ray": round(counts["nodes"] / nrays, 3), "tris_per_ray": round(counts["tris"] / nrays, 3),
                  "instances_per_ray": round(counts["instances"] / nrays, 3),
                  "coherent_nodes_per_ray": round(ccounts["nodes"] / primary.shape[0], 3),
                  "scene_nodes": int(info.num_nodes), "scene_build_ms": round(info.build_ms, 1),
                  "scene_mib": round(info.device_bytes / 2**20, 1), "parity": parity,
                  "launch_scaling": scaling,
                  "trace_valu_issue": valu,
                  "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
                  "traversal_1_5M": micro, "c1_cornell_256": c1, "path_tracing_10M": render,
                  "frame_arena_10M": arena, "time_to_rel_mse_0_01": relmse, "batch_64_scenes": batch},
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    # every parity check of the run must hold; a mismatch fails the run after the line is printed
    bad = []
    if parity and (parity["closest_hit_record_mismatches"] or parity["anyhit_mismatches"]):
        bad.append(f"10M-triangle sets {parity}")
    if micro and micro.get("parity") and any(micro["parity"].values()):
        bad.append(f"traversal microbenchmark {micro['parity']}")
    if c1 and c1.get("images_bit_exact") is False:
        bad.append("C1 image differs from the oracle's")
    if arena and not all(r["same_image"] for r in arena["rows"]):
        bad.append("frame-arena images differ")
    if bad:
        log("PARITY FAILED: " + "; ".join(bad))
        sys.exit(1)


if __name__ == "__main__":
    main()
