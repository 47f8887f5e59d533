"""ctypes binding of the CPU oracle (liboracle_pt.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed host baseline.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_pt.so")
NO_HIT = 0xFFFFFFFF

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    lib.orc_version.restype = C.c_uint32
    lib.orc_pcg32_seq.argtypes = [C.c_uint64, C.c_uint64, P, C.c_uint32]
    lib.orc_pcg32_seq.restype = None
    lib.orc_intersect.argtypes = [P, P, P]
    lib.orc_intersect.restype = C.c_int
    lib.orc_scene_new.restype = P
    lib.orc_scene_free.argtypes = [P]
    lib.orc_scene_free.restype = None
    lib.orc_scene_add_mesh.argtypes = [P, P, C.c_uint32]
    lib.orc_scene_add_mesh.restype = C.c_int
    lib.orc_scene_add_instance.argtypes = [P, C.c_uint32, P, P]
    lib.orc_scene_add_instance.restype = C.c_int
    lib.orc_scene_build.argtypes = [P, C.c_int, C.c_int, C.c_int]
    lib.orc_scene_build.restype = C.c_int
    lib.orc_scene_check.argtypes = [P]
    lib.orc_scene_check.restype = C.c_int
    lib.orc_scene_num_prims.argtypes = [P]
    lib.orc_scene_num_prims.restype = C.c_uint64
    lib.orc_trace.argtypes = [P, P, C.c_uint32, C.c_int, C.c_int, C.c_int, P, P]
    lib.orc_trace.restype = None
    lib.orc_trace_wide.argtypes = [P, P, P, C.c_uint32, C.c_int, C.c_uint32, P, C.c_uint32, C.c_int, C.c_int, P, P]
    lib.orc_trace_wide.restype = None
    lib.orc_scene_set_mesh_attrs.argtypes = [P, C.c_uint32, P, P]
    lib.orc_scene_set_mesh_attrs.restype = C.c_int
    lib.orc_scene_set_materials.argtypes = [P, P, C.c_uint32]
    lib.orc_scene_set_materials.restype = C.c_int
    lib.orc_scene_set_lights.argtypes = [P, P, C.c_uint32]
    lib.orc_scene_set_lights.restype = C.c_int
    lib.orc_scene_set_background.argtypes = [P, P]
    lib.orc_scene_set_background.restype = None
    lib.orc_render.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                               C.c_uint64, C.c_uint32, C.c_int, P, P]
    lib.orc_render.restype = C.c_uint64
    lib.orc_bsdf_sample.argtypes = [C.c_int, P, C.c_float, P, C.c_float, C.c_float, P]
    lib.orc_bsdf_sample.restype = C.c_int
    lib.orc_bsdf_eval.argtypes = [C.c_int, P, C.c_float, P, P, P]
    lib.orc_bsdf_eval.restype = None
    lib.orc_cache_levels.argtypes = [P, P]
    lib.orc_cache_levels.restype = None
    lib.orc_cache_loss_grad.argtypes = [P, P, P, P, C.c_uint32, P, C.c_float, C.c_int, C.c_int, P, P, P]
    lib.orc_cache_loss_grad.restype = C.c_double
    lib.orc_adam_step.argtypes = [P, P, P, P, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint32]
    lib.orc_adam_step.restype = None
    lib.orc_camera_rays.argtypes = [P, C.c_uint32, C.c_uint32, P, P, P, C.c_uint32, C.c_uint64, C.c_uint32, P]
    lib.orc_camera_rays.restype = None
    lib.orc_rng_draws.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P]
    lib.orc_rng_draws.restype = None
    lib.orc_sincos_2pi.argtypes = [P, C.c_uint32, P]
    lib.orc_sincos_2pi.restype = None
    lib.orc_wide_max_stack.argtypes = [C.c_int]
    lib.orc_wide_max_stack.restype = C.c_uint32
    _lib = lib
    return lib


def pcg32_seq(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    load().orc_pcg32_seq(initstate, initseq, out.ctypes.data, n)
    return out


def intersect(ray8, tri9):
    r = np.ascontiguousarray(ray8, np.float32)
    t = np.ascontiguousarray(tri9, np.float32)
    out = np.zeros(3, np.float32)
    hit = load().orc_intersect(r.ctypes.data, t.ctypes.data, out.ctypes.data)
    return bool(hit), out


def bsdf_sample(mtype, albedo, rough, wo, u1, u2):
    """Local frame (z = normal).  Returns (ok, wi[3], weight = f cos / pdf [3], pdf)."""
    a = np.ascontiguousarray(albedo, np.float32)
    w = np.ascontiguousarray(wo, np.float32)
    out = np.zeros(7, np.float32)
    ok = load().orc_bsdf_sample(int(mtype), a.ctypes.data, float(rough), w.ctypes.data, float(u1), float(u2),
                                out.ctypes.data)
    return bool(ok), out[0:3].copy(), out[3:6].copy(), float(out[6])


def bsdf_eval(mtype, albedo, rough, wo, wi):
    """Local frame.  Returns (f * cos(theta_i) [3], pdf of bsdf_sample)."""
    a = np.ascontiguousarray(albedo, np.float32)
    o = np.ascontiguousarray(wo, np.float32)
    i = np.ascontiguousarray(wi, np.float32)
    out = np.zeros(4, np.float32)
    load().orc_bsdf_eval(int(mtype), a.ctypes.data, float(rough), o.ctypes.data, i.ctypes.data, out.ctypes.data)
    return out[0:3].copy(), float(out[3])


class OracleScene:
    """Meshes (unindexed, 9 floats per triangle) + optional instances, binary
    SAH trees built by the oracle itself."""

    def __init__(self):
        self.lib = load()
        self.h = self.lib.orc_scene_new()
        self._keep = []

    def close(self):
        if self.h:
            self.lib.orc_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_mesh(self, tri_verts):
        tv = np.ascontiguousarray(tri_verts, np.float32).reshape(-1, 9)
        r = self.lib.orc_scene_add_mesh(self.h, tv.ctypes.data, tv.shape[0])
        assert r >= 0
        return r

    def add_instance(self, mesh, o2w, w2o):
        a = np.ascontiguousarray(o2w, np.float32).reshape(12)
        b = np.ascontiguousarray(w2o, np.float32).reshape(12)
        r = self.lib.orc_scene_add_instance(self.h, mesh, a.ctypes.data, b.ctypes.data)
        assert r >= 0
        return r

    def build(self, leaf_max=4, bins=16, threads=0):
        rc = self.lib.orc_scene_build(self.h, leaf_max, bins, threads)
        assert rc == 0, rc
        return self

    def check(self):
        return self.lib.orc_scene_check(self.h)

    def num_prims(self):
        return self.lib.orc_scene_num_prims(self.h)

    def set_mesh_attrs(self, mesh, material, light):
        m = np.ascontiguousarray(material, np.uint32)
        li = np.ascontiguousarray(light, np.int32)
        assert self.lib.orc_scene_set_mesh_attrs(self.h, mesh, m.ctypes.data, li.ctypes.data) == 0

    def set_materials(self, mats8):
        a = np.ascontiguousarray(mats8, np.float32).reshape(-1, 8)
        assert self.lib.orc_scene_set_materials(self.h, a.ctypes.data, a.shape[0]) == 0

    def set_lights(self, lights12):
        a = np.ascontiguousarray(lights12, np.float32).reshape(-1, 12)
        assert self.lib.orc_scene_set_lights(self.h, a.ctypes.data if a.size else None, a.shape[0]) == 0

    def set_background(self, rgb):
        a = np.ascontiguousarray(rgb, np.float32)
        self.lib.orc_scene_set_background(self.h, a.ctypes.data)

    def render(self, cam10, width, height, spp, max_depth, rr_depth=None, seed=1, sample_offset=0, mode=0,
               threads=0, accum=None, fp64=False):
        """Adds spp samples per pixel; returns (sum image (H,W,3) float32,
        sum image float64 or None, rays traced)."""
        cam = np.ascontiguousarray(cam10, np.float32)
        acc = np.zeros((height, width, 3), np.float32) if accum is None else accum
        acc64 = np.zeros((height, width, 3), np.float64) if fp64 else None
        rays = self.lib.orc_render(self.h, cam.ctypes.data, width, height, spp, sample_offset, max_depth,
                                   max_depth if rr_depth is None else rr_depth, seed, mode, threads,
                                   acc.ctypes.data, acc64.ctypes.data if fp64 else None)
        return acc, acc64, rays

    def trace(self, rays, anyhit=False, brute=False, threads=0, counts=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        cnt = np.zeros((rays.shape[0], 2), np.uint32) if counts else None
        self.lib.orc_trace(self.h, rays.ctypes.data, rays.shape[0], int(anyhit), int(brute), threads,
                           hits.ctypes.data, cnt.ctypes.data if counts else None)
        return (hits, cnt) if counts else hits


def scene_from_generated(gen, leaf_max=4, bins=16, threads=0, shading=False):
    """Oracle scene holding the same triangles and transforms as a
    foundation_amd.pt.GeneratedScene (and its materials and lights if
    shading)."""
    s = OracleScene()
    meshes = gen.meshes()
    for i, m in enumerate(meshes):
        s.add_mesh(m["positions"][m["indices"]].reshape(-1, 9))
        if shading:
            s.set_mesh_attrs(i, m["material"], m["light"])
    for inst in gen.instances():
        s.add_instance(inst["mesh"], inst["o2w"], inst["w2o"])
    if shading:
        d = gen.desc
        mats = np.zeros((max(1, d.num_materials), 8), np.float32)
        for k in range(d.num_materials):
            mt = d.materials[k]
            mats[k, 0] = np.array([mt.type], np.uint32).view(np.float32)[0]
            mats[k, 1:4] = mt.albedo[:]
            mats[k, 4] = mt.roughness
            mats[k, 5:8] = mt.emission[:]
        s.set_materials(mats)
        lights = np.zeros((d.num_lights, 12), np.float32)
        for k in range(d.num_lights):
            L = d.lights[k]
            lights[k] = list(L.corner) + list(L.edge0) + list(L.edge1) + list(L.radiance)
        s.set_lights(lights)
        s.set_background(np.array(d.background[:], np.float32))
    return s.build(leaf_max=leaf_max, bins=bins, threads=threads)


def camera_rays(cam10, width, height, seed=1, sample=0, points=None):
    """Camera rays as orc_render starts its paths, (N, 8) float32.  points =
    (xs, ys, jxy) for explicit image points (x + jx, y + jy); otherwise every
    pixel with the jitter of `sample`."""
    cam = np.ascontiguousarray(cam10, np.float32)
    if points is None:
        out = np.zeros((width * height, 8), np.float32)
        load().orc_camera_rays(cam.ctypes.data, width, height, None, None, None, 0, seed, sample, out.ctypes.data)
        return out
    xs = np.ascontiguousarray(points[0], np.uint32)
    ys = np.ascontiguousarray(points[1], np.uint32)
    jxy = np.ascontiguousarray(points[2], np.float32).reshape(-1, 2)
    out = np.zeros((xs.size, 8), np.float32)
    load().orc_camera_rays(cam.ctypes.data, width, height, xs.ctypes.data, ys.ctypes.data, jxy.ctypes.data, xs.size,
                           seed, sample, out.ctypes.data)
    return out


def rng_draws(seed, sample, pixel0, count, bounce, draws):
    out = np.zeros((count, draws), np.uint32)
    load().orc_rng_draws(seed, sample, pixel0, count, bounce, draws, out.ctypes.data)
    return out


def sincos_2pi(u):
    u = np.ascontiguousarray(u, np.float32).reshape(-1)
    out = np.zeros((u.size, 2), np.float32)
    load().orc_sincos_2pi(u.ctypes.data, u.size, out.ctypes.data)
    return out[:, 0].copy(), out[:, 1].copy()


def camera10(cam):
    return np.array(list(cam.eye) + list(cam.at) + list(cam.up) + [cam.fov_y_deg], np.float32)


def trace_wide(encoded, rays, anyhit=False, threads=0):
    """Host transcription of the GPU walk over fnd_scene_encode output
    (foundation_amd.pt.encode_scene).  Returns (hits, counts (N,3))."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    hits = np.zeros((rays.shape[0], 4), np.float32)
    cnt = np.zeros((rays.shape[0], 3), np.uint32)
    nodes = np.ascontiguousarray(encoded["nodes"])
    tris = np.ascontiguousarray(encoded["tris"])
    insts = np.ascontiguousarray(encoded["instances"])
    load().orc_wide_max_stack(1)
    load().orc_trace_wide(nodes.ctypes.data, tris.ctypes.data, insts.ctypes.data if insts.size else None,
                          encoded["root"], int(encoded["two_level"]), encoded["stack_entries"],
                          rays.ctypes.data, rays.shape[0], int(anyhit), threads, hits.ctypes.data, cnt.ctypes.data)
    max_stack = load().orc_wide_max_stack(0)
    if max_stack > encoded["stack_entries"]:
        raise AssertionError(f"walk needed {max_stack} stack entries, bound is {encoded['stack_entries']}")
    return hits, cnt


CACHE_MLP_PARAMS = 20480
CACHE_GRID_PARAMS = 16 * (1 << 19) * 2


def cache_levels():
    res = np.zeros(16, np.int32)
    mask = np.zeros(1, np.uint32)
    load().orc_cache_levels(res.ctypes.data, mask.ctypes.data)
    return res, int(mask[0])


def cache_loss_grad(mlp, grid, queries, targets, lo, inv_extent, round_fp16=True, want_grid=True, plain_l2=False):
    """FP64 loss, prediction and gradients of the radiance cache."""
    mlp = np.ascontiguousarray(mlp, np.float32)
    grid = np.ascontiguousarray(grid, np.float32)
    q = np.asarray(queries, np.float32)
    if q.shape[-1] == 6:  # diffuse queries: direction 0, roughness 0
        q = np.concatenate([q.reshape(-1, 6), np.zeros((q.size // 6, 4), np.float32)], 1)
    q = np.ascontiguousarray(q, np.float32).reshape(-1, 10)
    t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
    lo = np.ascontiguousarray(lo, np.float32)
    n = q.shape[0]
    pred = np.zeros((n, 3), np.float32)
    dmlp = np.zeros(CACHE_MLP_PARAMS, np.float64)
    dgrid = np.zeros(CACHE_GRID_PARAMS, np.float64) if want_grid else None
    loss = load().orc_cache_loss_grad(mlp.ctypes.data, grid.ctypes.data, q.ctypes.data, t.ctypes.data, n,
                                      lo.ctypes.data, float(inv_extent), int(round_fp16), int(plain_l2),
                                      pred.ctypes.data,
                                      dmlp.ctypes.data, dgrid.ctypes.data if want_grid else None)
    return loss, pred, dmlp, dgrid


def adam_step(p, m, v, g, lr, t, b1=0.9, b2=0.99, eps=1e-15):
    for a in (p, m, v, g):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    load().orc_adam_step(p.ctypes.data, m.ctypes.data, v.ctypes.data, g.ctypes.data, p.size, lr, b1, b2, eps, t)
