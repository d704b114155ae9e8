"""Seeded ray sets for closest-hit parity (test data generation, numpy float32).

The extend set on the triangle room: half primary camera rays (uniform jitter
over the image), half secondary rays leaving the closest hits of an independent
primary batch with cosine-distributed directions about the hit normal (the
reference Lambertian sampler, cpu/material.cpp:23-33).  Rays are 8 floats:
org.xyz, tmin, dir.xyz, tmax; directions normalized as v / sqrt(v.v) in float32.
"""
import numpy as np

EPS = np.float32(1e-3)
INF = np.float32(1e20)


def normalize32(v):
    v = v.astype(np.float32)
    n = v[:, 0] * v[:, 0] + (v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    s = np.sqrt(n).astype(np.float32)
    return v / s[:, None]


def camera_rays(cam_loc, cam_rot, width, height, fov, n, rng):
    """Pinhole rays following cpu/camera.cpp (no aperture) for uniform random pixel positions."""
    R = np.asarray(cam_rot, dtype=np.float32)
    u, v, w = R[:, 0], R[:, 1], R[:, 2]
    if width > height:
        vw = np.float32(2.0 * np.tan(np.float32(fov) * np.float32(0.5)))
        vh = np.float32(vw * np.float32(height) / np.float32(width))
    else:
        vh = np.float32(2.0 * np.tan(np.float32(fov) * np.float32(0.5)))
        vw = np.float32(vh * np.float32(width) / np.float32(height))
    px = (rng.random(n) * width).astype(np.float32)
    py = (rng.random(n) * height).astype(np.float32)
    a = (px / np.float32(width) - np.float32(0.5))[:, None]
    b = (py / np.float32(height) - np.float32(0.5))[:, None]
    d = (a * (vw * u)[None] + b * (vh * v)[None]) + (-w)[None]
    d = normalize32(d)
    o = np.broadcast_to(np.asarray(cam_loc, dtype=np.float32), d.shape)
    return pack(o, d)


def pack(o, d, tmin=EPS, tmax=INF):
    r = np.empty((o.shape[0], 8), dtype=np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def cosine_dirs(normals, rng):
    n = normals.astype(np.float64)
    phi = rng.random(len(n)) * 2 * np.pi
    s2 = rng.random(len(n))
    st, ct = np.sqrt(s2), np.sqrt(1 - s2)
    up = np.where(np.abs(n[:, 1:2]) < 0.9, np.array([[0, 1.0, 0]]), np.array([[1.0, 0, 0]]))
    tu = np.cross(up, n)
    tu /= np.linalg.norm(tu, axis=1, keepdims=True)
    tv = np.cross(n, tu)
    d = tu * (np.cos(phi) * st)[:, None] + tv * (np.sin(phi) * st)[:, None] + n * ct[:, None]
    return normalize32(d.astype(np.float32))


def extend_set(oracle_scene, cam, n_total, seed):
    """cam = (loc, rot, width, height, fov).  Returns (N, 8) float32 rays."""
    rng = np.random.default_rng(seed)
    half = n_total // 2
    prim = camera_rays(*cam, half, rng)
    aux = camera_rays(*cam, n_total - half, rng)
    pid, surf = oracle_scene.intersect_surface_batch(aux)
    hit = pid >= 0
    # missed rays stay primaries from the camera
    o = aux[:, 0:3].copy()
    t = surf[:, 0]
    o[hit] = (aux[hit, 0:3] + aux[hit, 4:7] * t[hit, None]).astype(np.float32)
    nrm = surf[:, 1:4]
    d = aux[:, 4:7].copy()
    d[hit] = cosine_dirs(nrm[hit], rng)
    sec = pack(o, d)
    return np.concatenate([prim, sec], axis=0)
