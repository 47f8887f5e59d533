"""Host-side ray sets for the traversal benchmark and the parity tests.

Rays are generated on the host in numpy so that the GPU and the CPU oracle
trace identical float32 inputs (the sin/cos of the direction samplers would
otherwise differ by ulps between the host libm and the device).

* PCG32 (O'Neill 2014), vectorised over streams.
* Pinhole camera with glm::lookAt conventions and a Z-up world, the camera of
  the reference's only view (src/Renderer/Renderer.cpp:373-380).
* Cosine-weighted hemisphere by Malley's method with the branchless
  orthonormal basis of Duff et al., "Building an Orthonormal Basis,
  Revisited", JCGT 6(1), 2017.
"""
import numpy as np

_MULT = np.uint64(6364136223846793005)


def pcg32_srandom(initstate, initseq):
    """Vectorised pcg32_srandom_r: returns (state, inc) arrays."""
    initstate = np.asarray(initstate, dtype=np.uint64)
    initseq = np.asarray(initseq, dtype=np.uint64)
    with np.errstate(over="ignore"):
        inc = (initseq << np.uint64(1)) | np.uint64(1)
        state = np.zeros(np.broadcast(initstate, initseq).shape, np.uint64)
        state = state * _MULT + inc
        state = state + initstate
        state = state * _MULT + inc
    return state, inc


def pcg32_next(state, inc):
    """One pcg32_random_r step on arrays: returns (outputs u32, new state)."""
    old = state
    with np.errstate(over="ignore"):
        state = old * _MULT + inc
    xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
    rot = (old >> np.uint64(59)).astype(np.uint32)
    out = (xs >> rot) | (xs << ((np.uint32(32) - rot) & np.uint32(31)))
    return out.astype(np.uint32), state


def u32_to_unit(u):
    return (u >> np.uint32(8)).astype(np.float32) * np.float32(2.0 ** -24)


def camera_basis(camera):
    eye = np.array(camera.eye[:], np.float64)
    at = np.array(camera.at[:], np.float64)
    up = np.array(camera.up[:], np.float64)
    f = at - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return eye, f, r, u


def primary_rays(camera, width, height, tmax=1e30):
    """One ray per pixel centre, row 0 at the top, (H*W, 8) float32."""
    eye, f, r, u = camera_basis(camera)
    th = np.tan(np.radians(camera.fov_y_deg) * 0.5)
    aspect = width / height
    xs = ((np.arange(width) + 0.5) / width * 2.0 - 1.0) * th * aspect
    ys = (1.0 - (np.arange(height) + 0.5) / height * 2.0) * th
    X, Y = np.meshgrid(xs, ys)
    d = f[None, None, :] + X[..., None] * r[None, None, :] + Y[..., None] * u[None, None, :]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.zeros((height * width, 8), np.float32)
    rays[:, 0:3] = eye.astype(np.float32)
    rays[:, 3] = 0.0
    rays[:, 4:7] = d.reshape(-1, 3).astype(np.float32)
    rays[:, 7] = tmax
    return rays


def onb_duff(n):
    """Branchless ONB (Duff et al. 2017) for unit normals n (N,3) float64."""
    sign = np.where(n[:, 2] >= 0.0, 1.0, -1.0)
    a = -1.0 / (sign + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    t = np.stack([1.0 + sign * n[:, 0] ** 2 * a, sign * b, -sign * n[:, 0]], axis=1)
    bt = np.stack([b, sign + n[:, 1] ** 2 * a, -n[:, 1]], axis=1)
    return t, bt


def hit_frames(rays, hits, world_tris):
    """For the rays that hit: indices, hit points, unit geometric normals
    facing the incoming ray (float64).  world_tris: (P,3,3) array or a
    callable prim ids -> (n,3,3)."""
    prim = hits[:, 1].view(np.uint32)
    idx = np.nonzero(prim != 0xFFFFFFFF)[0]
    o = rays[idx, 0:3].astype(np.float64)
    d = rays[idx, 4:7].astype(np.float64)
    t = hits[idx, 0].astype(np.float64)
    p = o + t[:, None] * d
    v = world_tris(prim[idx]) if callable(world_tris) else world_tris[prim[idx]]
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-300)
    flip = np.sum(n * d, axis=1) > 0
    n[flip] *= -1.0
    return idx, p, n


def bounce_rays(rays, hits, world_tris, seed=1, tmax=1e30):
    """Incoherent set: one cosine-weighted hemisphere ray about the geometric
    normal at each primary hit, origin offset 1e-4 * max(|p|, 1) along the
    normal, PCG32 stream = pixel index, state = seed.  Returns (rays, pixel)."""
    idx, p, n = hit_frames(rays, hits, world_tris)
    state, inc = pcg32_srandom(np.full(idx.shape, seed, np.uint64), idx.astype(np.uint64))
    a, state = pcg32_next(state, inc)
    b, state = pcg32_next(state, inc)
    u1 = u32_to_unit(a).astype(np.float64)
    u2 = u32_to_unit(b).astype(np.float64)
    r = np.sqrt(u1)
    phi = 2.0 * np.pi * u2
    lx, ly, lz = r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - u1))
    t1, t2 = onb_duff(n)
    d = lx[:, None] * t1 + ly[:, None] * t2 + lz[:, None] * n
    off = 1e-4 * np.maximum(np.linalg.norm(p, axis=1), 1.0)
    org = p + off[:, None] * n
    out = np.zeros((idx.size, 8), np.float32)
    out[:, 0:3] = org.astype(np.float32)
    out[:, 3] = 0.0
    out[:, 4:7] = d.astype(np.float32)
    out[:, 7] = tmax
    return out, idx


def shadow_rays(rays, hits, world_tris, light_pos):
    """Any-hit set: from each primary hit (offset as in bounce_rays) toward a
    point light; dir = light - origin, tmax = 1."""
    idx, p, n = hit_frames(rays, hits, world_tris)
    off = 1e-4 * np.maximum(np.linalg.norm(p, axis=1), 1.0)
    org = p + off[:, None] * n
    d = np.asarray(light_pos, np.float64)[None, :] - org
    out = np.zeros((idx.size, 8), np.float32)
    out[:, 0:3] = org.astype(np.float32)
    out[:, 3] = 0.0
    out[:, 4:7] = d.astype(np.float32)
    out[:, 7] = 1.0
    return out, idx
