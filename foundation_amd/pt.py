"""Python face of the path-tracing pass: thin wrappers over the fnd_pt_* C ABI.

Mirrors the C++ PathTracerPass wrapper (csrc/PathTracerPass.hpp): a
``Context`` owns scenes, device buffers and framebuffers like the reference's
Renderer owns its device objects (src/Renderer/Renderer.hpp:10-44), and every
non-OK status raises ``FndError`` where the reference would CHECK/BugCheck
(src/Core/Core.hpp:17).
"""
import ctypes as C

import numpy as np

from . import _abi as A


class FndError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{A.load().fnd_pt_status_string(status).decode()} ({status}): {message}")
        self.status = status
        self.message = message


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


# ---------------------------------------------------------------- generated scenes


class GeneratedScene:
    """A procedural scene made by the library's generators (fnd_gen_*)."""

    def __init__(self, kind, **overrides):
        lib = A.load()
        p = A.GenParams()
        st = lib.fnd_gen_default_params(kind, C.byref(p))
        if st != A.OK:
            raise FndError(st, f"unknown generator kind {kind}")
        for k, v in overrides.items():
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, v)
        self.params = p
        h = C.c_void_p()
        st = lib.fnd_gen_scene(C.byref(p), C.byref(h))
        if st != A.OK:
            raise FndError(st, "scene generation failed")
        self._h = h
        self.desc = A.SceneDesc()
        self.camera = A.Camera()
        lib.fnd_gen_scene_get(h, C.byref(self.desc), C.byref(self.camera))

    def close(self):
        if self._h:
            A.load().fnd_gen_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def meshes(self):
        """Copies of each mesh: positions (V,3) f32, indices (T,3) u32,
        material (T,) u32, light (T,) i32."""
        out = []
        for i in range(self.desc.num_meshes):
            m = self.desc.meshes[i]
            pos = np.ctypeslib.as_array(m.positions, shape=(m.num_vertices * 3,)).reshape(-1, 3).copy()
            idx = np.ctypeslib.as_array(m.indices, shape=(m.num_triangles * 3,)).reshape(-1, 3).copy()
            mat = (np.ctypeslib.as_array(m.material_ids, shape=(m.num_triangles,)).copy()
                   if m.material_ids else np.zeros(m.num_triangles, np.uint32))
            lig = (np.ctypeslib.as_array(m.light_ids, shape=(m.num_triangles,)).copy()
                   if m.light_ids else np.full(m.num_triangles, -1, np.int32))
            out.append(dict(positions=pos, indices=idx, material=mat, light=lig))
        return out

    def instances(self):
        out = []
        for i in range(self.desc.num_instances):
            d = self.desc.instances[i]
            o2w = np.array(d.object_to_world[:], np.float32).reshape(3, 4)
            w2o = np.array(d.world_to_object[:], np.float32).reshape(3, 4)
            out.append(dict(mesh=d.mesh, o2w=o2w, w2o=w2o))
        return out

    def num_prims(self):
        if self.desc.num_instances == 0:
            return self.desc.meshes[0].num_triangles
        return sum(self.desc.meshes[self.desc.instances[i].mesh].num_triangles
                   for i in range(self.desc.num_instances))

    def triangle_lookup(self):
        """Callable prim ids -> (n, 3, 3) float64 world-space vertices, without
        expanding the whole scene."""
        ms = self.meshes()
        if self.desc.num_instances == 0:
            m = ms[0]
            return lambda prims: m["positions"][m["indices"][prims]].astype(np.float64)
        insts = self.instances()
        counts = np.array([ms[i["mesh"]]["indices"].shape[0] for i in insts], np.int64)
        base = np.concatenate([[0], np.cumsum(counts)[:-1]])

        def lookup(prims):
            prims = np.asarray(prims, np.int64)
            inst = np.searchsorted(base, prims, side="right") - 1
            out = np.zeros((prims.size, 3, 3), np.float64)
            for k in np.unique(inst):
                sel = np.nonzero(inst == k)[0]
                m = ms[insts[k]["mesh"]]
                v = m["positions"][m["indices"][prims[sel] - base[k]]].astype(np.float64)
                o2w = insts[k]["o2w"].astype(np.float64)
                out[sel] = v @ o2w[:, :3].T + o2w[:, 3]
            return out
        return lookup

    def world_triangles(self):
        """(P, 3, 3) float64 world-space vertices indexed by primitive id."""
        ms = self.meshes()
        if self.desc.num_instances == 0:
            m = ms[0]
            return m["positions"][m["indices"]].astype(np.float64)
        parts = []
        for inst in self.instances():
            m = ms[inst["mesh"]]
            v = m["positions"][m["indices"]].astype(np.float64)
            o2w = inst["o2w"].astype(np.float64)
            parts.append(v @ o2w[:, :3].T + o2w[:, 3])
        return np.concatenate(parts, axis=0)


# ---------------------------------------------------------------- host encoding


WIDE_NODE_DTYPE = np.dtype([
    ("p", "<f4", 3), ("e", "i1", 3), ("imask", "u1"), ("entry", "<u4", 4), ("qlo", "<u4", 3), ("qhi", "<u4", 3),
    ("child_base", "<u4"), ("prim_base", "<u4")])
assert WIDE_NODE_DTYPE.itemsize == 64
LEAF_TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("local_id", "<u4"), ("v1", "<f4", 3), ("material", "<u4"),
                           ("v2", "<f4", 3), ("light", "<i4")])
assert LEAF_TRI_DTYPE.itemsize == 48
INSTANCE_DTYPE = np.dtype([("w2o", "<f4", 12), ("blas_root", "<u4"), ("id_base", "<u4"), ("mesh", "<u4"),
                           ("orig_index", "<u4")])
assert INSTANCE_DTYPE.itemsize == 64


# ---------------------------------------------------------------- scene ingestion


def mesh_from_streams(vertex_bytes, stride, position_offset, position_components, num_vertices, indices):
    """fnd_mesh_from_streams over a reference-style interleaved vertex buffer
    (bytes, stride, byte offset of the position, 2 or 3 floats) and a uint16
    or uint32 index array: returns positions (V, 3) f32 and indices (T, 3) u32."""
    lib = A.load()
    vb = np.frombuffer(bytes(vertex_bytes), np.uint8)
    ib = np.ascontiguousarray(indices).reshape(-1)
    if ib.dtype not in (np.uint16, np.uint32):
        raise TypeError("indices must be uint16 or uint32")
    vs = A.VertexStream(vb.ctypes.data, vb.size, stride, position_offset, position_components, num_vertices)
    ist = A.IndexStream(ib.ctypes.data, 8 * ib.dtype.itemsize, ib.size)
    pos = np.zeros((num_vertices, 3), np.float32)
    idx = np.zeros((ib.size // 3, 3), np.uint32)
    err = C.create_string_buffer(256)
    st = lib.fnd_mesh_from_streams(C.byref(vs), C.byref(ist), pos.ctypes.data, idx.ctypes.data, err, 256)
    if st != A.OK:
        raise FndError(st, err.value.decode())
    return pos, idx


# ---------------------------------------------------------------- frame arena


class StackAllocator:
    """Bump allocator over one block of memory (fnd_arena_*), named after
    the reference's Core::StackAllocator (src/Core/Allocator/StackAllocator.hpp:9-63)
    whose semantics it has: Allocate returns None for size 0, without memory,
    or when the block would pass the end; Deallocate does nothing;
    Reallocate raises.  Addresses are integers; the memory is never touched."""

    def __init__(self, memory=0, size=0):
        self.lib = A.load()
        self.arena = A.Arena()
        self.reset(memory, size)

    def reset(self, memory=0, size=0):
        self.lib.fnd_arena_reset(C.byref(self.arena), memory or None, size)

    def allocate(self, size, alignment=None):
        if alignment is None:
            p = self.lib.fnd_arena_alloc(C.byref(self.arena), size)
        else:
            p = self.lib.fnd_arena_alloc_aligned(C.byref(self.arena), size, alignment)
        return p  # ctypes gives None for a null pointer

    def deallocate(self, ptr, size=0):
        pass

    def reallocate(self, ptr, new_size, alignment=None):
        raise RuntimeError("StackAllocator does not support reallocation")

    def used(self):
        return self.lib.fnd_arena_used(C.byref(self.arena))


class DeviceMemory:
    """`nbytes` of device memory from hipMalloc of the HIP runtime that
    libfnd_pt.so itself links: the caller-side allocation a renderer makes for
    memory it lends to the library, such as a frame arena.  (A second copy of
    the HIP runtime loaded into the same process, as some packages ship, is a
    separate runtime; memory is taken from the library's own.)"""
    _hip = None

    def __init__(self, nbytes, device=0):
        hip = DeviceMemory._runtime()
        st = hip.hipSetDevice(device)
        p = C.c_void_p()
        if st == 0:
            st = hip.hipMalloc(C.byref(p), nbytes)
        if st != 0:
            raise FndError(A.E_OOM_DEVICE if st == 2 else A.E_DEVICE, f"hipMalloc({nbytes}) failed with HIP error {st}")
        self.ptr = p.value
        self.nbytes = nbytes

    @staticmethod
    def _runtime():
        if DeviceMemory._hip is None:
            A.load()  # the library's DT_NEEDED runtime is then resolved by soname
            hip = C.CDLL("libamdhip64.so.7")
            hip.hipSetDevice.argtypes = [C.c_int]
            hip.hipSetDevice.restype = C.c_int
            hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            hip.hipMalloc.restype = C.c_int
            hip.hipFree.argtypes = [C.c_void_p]
            hip.hipFree.restype = C.c_int
            hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            hip.hipMemcpy.restype = C.c_int
            DeviceMemory._hip = hip
        return DeviceMemory._hip

    def write(self, data):
        """Copies host bytes (a numpy array or bytes) to the start of the block."""
        buf = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8) if isinstance(data, (bytes, bytearray))
                                   else data).view(np.uint8).reshape(-1)
        if buf.size > self.nbytes:
            raise ValueError("more bytes than the block holds")
        st = DeviceMemory._hip.hipMemcpy(self.ptr, buf.ctypes.data, buf.size, 1)  # hipMemcpyHostToDevice
        if st != 0:
            raise FndError(A.E_DEVICE, f"hipMemcpy failed with HIP error {st}")

    def read(self, nbytes=None):
        """Copies the first nbytes (all by default) of the block to a uint8 array."""
        n = self.nbytes if nbytes is None else nbytes
        out = np.zeros(n, np.uint8)
        st = DeviceMemory._hip.hipMemcpy(out.ctypes.data, self.ptr, n, 2)  # hipMemcpyDeviceToHost
        if st != 0:
            raise FndError(A.E_DEVICE, f"hipMemcpy failed with HIP error {st}")
        return out

    def free(self):
        if self.ptr:
            DeviceMemory._hip.hipFree(self.ptr)
            self.ptr = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def queue_bytes(capacity):
    """Bytes of the per-frame ray queues of a batch of `capacity` paths."""
    out = C.c_uint64()
    st = A.load().fnd_pt_queue_bytes(capacity, C.byref(out))
    if st != A.OK:
        raise FndError(st, "fnd_pt_queue_bytes")
    return out.value


class ObjMesh:
    """A Wavefront OBJ mesh loaded by fnd_obj_load; ``desc`` is its
    fnd_mesh_desc (valid while the object lives)."""

    def __init__(self, path):
        lib = A.load()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        st = lib.fnd_obj_load(str(path).encode(), C.byref(h), err, 512)
        if st != A.OK:
            raise FndError(st, err.value.decode())
        self._h = h
        self.desc = A.MeshDesc()
        lib.fnd_obj_mesh_get(h, C.byref(self.desc))

    def arrays(self):
        d = self.desc
        pos = np.ctypeslib.as_array(d.positions, shape=(d.num_vertices * 3,)).reshape(-1, 3).copy()
        idx = np.ctypeslib.as_array(d.indices, shape=(d.num_triangles * 3,)).reshape(-1, 3).copy()
        return pos, idx

    def close(self):
        if self._h:
            A.load().fnd_obj_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def encode_scene(desc):
    """Host encoding of a scene (fnd_scene_encode): dict of numpy arrays with
    the exact bytes fnd_pt_scene_create uploads."""
    lib = A.load()
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    st = lib.fnd_scene_encode(C.byref(desc), C.byref(h), err, 512)
    if st != A.OK:
        raise FndError(st, err.value.decode())
    try:
        v = A.HostSceneView()
        lib.fnd_host_scene_get(h, C.byref(v))

        def grab(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            buf = (C.c_char * (n * dt.itemsize)).from_address(ptr)
            return np.frombuffer(bytes(buf), dtype=dt).copy()

        out = dict(nodes=grab(v.nodes, v.num_nodes, WIDE_NODE_DTYPE),
                   tris=grab(v.tris, v.num_tris, LEAF_TRI_DTYPE),
                   instances=grab(v.instances, v.num_instances, INSTANCE_DTYPE),
                   root=v.root, two_level=bool(v.two_level),
                   max_depth=v.info.max_depth, stack_entries=v.info.stack_entries,
                   num_prims=v.info.num_prims)
        return out
    finally:
        lib.fnd_host_scene_free(h)


# ---------------------------------------------------------------- context


class Context:
    """One fnd_pt_ctx: one device, one stream, the factory of its objects."""

    def __init__(self, device=0, device_memory_budget=0, allocator=None):
        self.lib = A.load()
        d = A.CreateDesc()
        d.device = device
        d.device_memory_budget = device_memory_budget
        self._alloc = allocator
        h = C.c_void_p()
        st = self.lib.fnd_pt_create(C.byref(d), C.byref(allocator) if allocator else None, C.byref(h))
        if st != A.OK:
            raise FndError(st, "fnd_pt_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.fnd_pt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, st):
        if st != A.OK:
            raise FndError(st, self.lib.fnd_pt_last_error(self.h).decode())

    # scenes
    def scene_create(self, desc):
        h = A.fnd_handle()
        self._ck(self.lib.fnd_pt_scene_create(self.h, C.byref(desc), C.byref(h)))
        return h.value

    def scene_destroy(self, scene):
        self._ck(self.lib.fnd_pt_scene_destroy(self.h, scene))

    def scene_info(self, scene):
        info = A.SceneInfo()
        self._ck(self.lib.fnd_pt_scene_info(self.h, scene, C.byref(info)))
        return info

    # buffers
    def buffer_create(self, nbytes):
        h = A.fnd_handle()
        self._ck(self.lib.fnd_pt_buffer_create(self.h, nbytes, C.byref(h)))
        return h.value

    def buffer_destroy(self, buf):
        self._ck(self.lib.fnd_pt_buffer_destroy(self.h, buf))

    def buffer_write(self, buf, array, offset=0):
        a = np.ascontiguousarray(array)
        self._ck(self.lib.fnd_pt_buffer_write(self.h, buf, offset, a.ctypes.data, a.nbytes))

    def buffer_read(self, buf, out, offset=0):
        assert out.flags["C_CONTIGUOUS"]
        self._ck(self.lib.fnd_pt_buffer_read(self.h, buf, offset, out.ctypes.data, out.nbytes))
        return out

    # ray queries
    def trace(self, scene, rays, anyhit=False, count=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        flags = (A.TRACE_ANYHIT if anyhit else 0) | (A.TRACE_COUNT if count else 0)
        self._ck(self.lib.fnd_pt_trace(self.h, scene, rays.ctypes.data, rays.shape[0], hits.ctypes.data, flags))
        return hits

    def trace_buffers(self, scene, ray_buf, hit_buf, n, flags=0):
        self._ck(self.lib.fnd_pt_trace_buffers(self.h, scene, ray_buf, hit_buf, n, flags))

    def trace_noalloc(self, scene, rays_ptr, hits_ptr, n, flags=0):
        """fnd_pt_trace_noalloc over caller device memory (device pointers)."""
        self._ck(self.lib.fnd_pt_trace_noalloc(self.h, scene, rays_ptr, hits_ptr, n, flags))

    def wait(self):
        self._ck(self.lib.fnd_pt_wait(self.h))

    def stats(self):
        s = A.Stats()
        self._ck(self.lib.fnd_pt_get_stats(self.h, C.byref(s)))
        return s

    def reset_stats(self):
        self._ck(self.lib.fnd_pt_reset_stats(self.h))

    def tuning(self):
        t = A.Tuning()
        self._ck(self.lib.fnd_pt_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        self._ck(self.lib.fnd_pt_set_tuning(self.h, C.byref(t)))

    # rendering
    def framebuffer_create(self, width, height, fp64=False):
        h = A.fnd_handle()
        self._ck(self.lib.fnd_pt_framebuffer_create(self.h, width, height, A.FB_FP64 if fp64 else 0, C.byref(h)))
        return h.value

    def framebuffer_destroy(self, fb):
        self._ck(self.lib.fnd_pt_framebuffer_destroy(self.h, fb))

    def framebuffer_clear(self, fb):
        self._ck(self.lib.fnd_pt_framebuffer_clear(self.h, fb))

    def framebuffer_read(self, fb, width, height, fp64=False):
        out = np.zeros((height, width, 3), np.float64 if fp64 else np.float32)
        n = C.c_uint32()
        fn = self.lib.fnd_pt_framebuffer_read_f64 if fp64 else self.lib.fnd_pt_framebuffer_read
        self._ck(fn(self.h, fb, out.ctypes.data, C.byref(n)))
        return out, n.value

    def framebuffer_rel_mse(self, fb, ref, eps=0.01):
        """relMSE of fb's mean radiance against the FP64 reference framebuffer
        ref, evaluated on the device (fnd_pt_framebuffer_rel_mse)."""
        out = C.c_double()
        self._ck(self.lib.fnd_pt_framebuffer_rel_mse(self.h, fb, ref, eps, C.byref(out)))
        return out.value

    def framebuffer_read_rgba8(self, fb, width, height, exposure=1.0):
        """R8G8B8A8 bytes of the mean radiance, converted on the device."""
        out = np.zeros((height, width, 4), np.uint8)
        n = C.c_uint32()
        self._ck(self.lib.fnd_pt_framebuffer_read_rgba8(self.h, fb, exposure, out.ctypes.data, C.byref(n)))
        return out, n.value

    def framebuffer_present(self, fb, device_ptr, row_pitch, exposure=1.0):
        """R8G8B8A8 bytes of the mean radiance written into device memory
        (rows row_pitch bytes apart); returns the sample count."""
        n = C.c_uint32()
        self._ck(self.lib.fnd_pt_framebuffer_present(self.h, fb, exposure, device_ptr, row_pitch, C.byref(n)))
        return n.value

    def render(self, jobs):
        arr = (A.RenderJob * len(jobs))(*jobs)
        self._ck(self.lib.fnd_pt_render(self.h, arr, len(jobs)))

    def camera_rays(self, camera, width, height, seed=1, sample=0):
        """Every pixel's camera ray for `sample` as the renderer draws it
        (fnd_pt_camera_rays): (H*W, 8) float32."""
        out = np.zeros((height * width, 8), np.float32)
        self._ck(self.lib.fnd_pt_camera_rays(self.h, C.byref(camera), width, height, seed, sample, out.ctypes.data))
        return out

    def rng_draws(self, seed, sample, pixel0, count, bounce, draws):
        """The first `draws` outputs of the renderer's generator of (seed,
        sample, pixel, bounce) for `count` pixels from pixel0: (count, draws) u32."""
        out = np.zeros((count, draws), np.uint32)
        self._ck(self.lib.fnd_pt_rng_draws(self.h, seed, sample, pixel0, count, bounce, draws, out.ctypes.data))
        return out

    def command_list(self):
        return CommandList(self)

    def fence_create(self):
        h = C.c_void_p()
        self._ck(self.lib.fnd_pt_fence_create(self.h, C.byref(h)))
        return h

    def fence_destroy(self, fence):
        self.lib.fnd_pt_fence_destroy(fence)

    def submit(self, lists, fence=None):
        arr = (C.c_void_p * max(1, len(lists)))(*[cl.h for cl in lists])
        self._ck(self.lib.fnd_pt_submit(self.h, arr, len(lists), fence))

    def fence_wait(self, fence):
        self._ck(self.lib.fnd_pt_fence_wait(self.h, fence))

    def set_frame_arena(self, device_ptr, nbytes):
        """Per-frame ray queues from the caller's device memory (0: the
        library's own block); the caller keeps the memory alive."""
        self._ck(self.lib.fnd_pt_set_frame_arena(self.h, device_ptr or None, nbytes))

    def frame_arena_info(self):
        out = A.FrameArenaInfo()
        self._ck(self.lib.fnd_pt_frame_arena_info(self.h, C.byref(out)))
        return out

    def autotune(self, jobs, cache_path=None, force=False):
        """Searches launch configurations for `jobs` (fnd_pt_autotune) and
        applies the winner; returns the fnd_autotune_result."""
        arr = (A.RenderJob * len(jobs))(*jobs)
        res = A.AutotuneResult()
        path = cache_path.encode() if cache_path else None
        self._ck(self.lib.fnd_pt_autotune(self.h, arr, len(jobs), path, A.AUTOTUNE_FORCE if force else 0,
                                          C.byref(res)))
        return res


class CommandList:
    """A compute command list of a context (fnd_pt_cmd_*), chained like the
    reference's RHICommandList (src/Platform/RHI/Command.hpp:32-116):
    ``cl.begin().set_pipeline(p).bind(...).push_constants(c).dispatch(x, y, z).end()``.
    A refused call raises FndError at once."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._ck(ctx.lib.fnd_pt_cmd_create(ctx.h, C.byref(h)))
        self.h = h

    def destroy(self):
        if self.h and self.ctx.h:
            self.ctx.lib.fnd_pt_cmd_destroy(self.h)
        self.h = None

    def begin(self):
        self.ctx._ck(self.ctx.lib.fnd_pt_cmd_begin(self.h))
        return self

    def set_pipeline(self, pipeline):
        self.ctx._ck(self.ctx.lib.fnd_pt_cmd_set_pipeline(self.h, pipeline))
        return self

    def bind(self, *descriptors):
        """Descriptors as (binding, type, resource[, offset, range, device_ptr])."""
        arr = (A.Descriptor * max(1, len(descriptors)))()
        for i, d in enumerate(descriptors):
            b, t, r = d[:3]
            arr[i].binding, arr[i].type, arr[i].resource = b, t, r
            if len(d) > 3:
                arr[i].offset = d[3]
            if len(d) > 4:
                arr[i].range = d[4]
            if len(d) > 5:
                arr[i].device_ptr = d[5]
        self.ctx._ck(self.ctx.lib.fnd_pt_cmd_bind_descriptors(self.h, arr, len(descriptors)))
        return self

    def push_constants(self, data):
        """A ctypes structure or a bytes-like object."""
        buf = bytes(data) if not isinstance(data, (bytes, bytearray)) else bytes(data)
        cbuf = C.create_string_buffer(buf, len(buf))
        self.ctx._ck(self.ctx.lib.fnd_pt_cmd_push_constants(self.h, cbuf, len(buf)))
        return self

    def dispatch(self, x, y=1, z=1):
        self.ctx._ck(self.ctx.lib.fnd_pt_cmd_dispatch(self.h, x, y, z))
        return self

    def end(self):
        self.ctx._ck(self.ctx.lib.fnd_pt_cmd_end(self.h))
        return self

    def size(self):
        return int(self.ctx.lib.fnd_pt_cmd_size(self.h))


def path_trace_constants(camera, max_depth, rr_depth=None, seed=1, sample_offset=0, flags=0):
    k = A.PathTraceConstants()
    k.camera = camera
    k.max_depth = max_depth
    k.rr_depth = max_depth if rr_depth is None else rr_depth
    k.flags = flags
    k.sample_offset = sample_offset
    k.seed = seed
    return k


def camera_basis(camera, width, height):
    """The renderer's camera basis (fnd_camera_basis, host only): dict of
    eye, forward, right, up (float32 arrays), tan_half and aspect.  Raises
    FndError for a camera without a basis."""
    out = np.zeros(14, np.float32)
    st = A.load().fnd_camera_basis(C.byref(camera), width, height, out.ctypes.data)
    if st != A.OK:
        raise FndError(st, "camera has no basis")
    return dict(eye=out[0:3].copy(), forward=out[3:6].copy(), right=out[6:9].copy(), up=out[9:12].copy(),
                tan_half=float(out[12]), aspect=float(out[13]))


def batch_scene(index):
    """Scene `index` of the multi-scene batch workload (fnd_gen_batch_params):
    returns (GeneratedScene, image side)."""
    p = A.GenParams()
    side = C.c_uint32()
    st = A.load().fnd_gen_batch_params(index, C.byref(p), C.byref(side))
    if st != A.OK:
        raise FndError(st, "fnd_gen_batch_params")
    kw = {name: getattr(p, name) for name, _ in A.GenParams._fields_ if name != "kind"}
    return GeneratedScene(p.kind, **kw), side.value


def make_job(scene, framebuffer, camera, spp, max_depth, rr_depth=None, seed=1, sample_offset=0, flags=0):
    j = A.RenderJob()
    j.scene = scene
    j.framebuffer = framebuffer
    j.camera = camera
    j.spp = spp
    j.max_depth = max_depth
    j.rr_depth = max_depth if rr_depth is None else rr_depth
    j.seed = seed
    j.sample_offset = sample_offset
    j.flags = flags
    return j
