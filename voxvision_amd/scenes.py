"""Seeded synthetic workloads for the ray-casting path.

Volumes become voxel sets the way src/voxtrees/datareader.c:50-77 does it:
voxel (i, j, k) of an occupancy volume becomes the dot (i, j, k) * V, the
dots enumerated with i outermost and k innermost.  Everything here is
deterministic for a given seed (numpy PCG64).

Workloads:
  blobs64     64^3 sum of 8 Gaussian blobs, ~15% occupancy (CPU-runnable case)
  ct256       256^3 4-octave value noise shaped as a shell plus interior,
              ~8% occupancy (~1.3 M voxels)
  random_rays rays with origins uniform in the box grown by 10% of its
              extent per side and directions normalised N(0, I)
  random_balls ball queries, centres uniform in the box
"""
import math

import numpy as np

from . import camera as _camera

BLOB_AMPLITUDE = 1.4  # gives 15-16% occupancy on 64^3 for seeds 1..3


def volume_to_dots(occ, voxel=(1.0, 1.0, 1.0)):
    """Occupancy volume [i, j, k] -> (n, 3) float32 dots, i-major order."""
    idx = np.nonzero(occ)
    v = np.asarray(voxel, np.float32)
    dots = np.empty((len(idx[0]), 3), np.float32)
    for a in range(3):
        # i*vox_voxel[0] in float (datareader.c:69-71)
        dots[:, a] = idx[a].astype(np.float32) * v[a]
    return dots


def blobs64(seed=1, n=64, nblobs=8, amplitude=BLOB_AMPLITUDE):
    """density = min(255, round(255*A*sum_b exp(-|x-c_b|^2 / 2 s_b^2))),
    c_b ~ U[0,n)^3, s_b ~ U[4,10]; voxel present where density > 127."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, n, (nblobs, 3))
    s = rng.uniform(4, 10, nblobs)
    x = np.arange(n, dtype=np.float64)
    dens = np.zeros((n, n, n))
    for b in range(nblobs):
        g = [np.exp(-(x - c[b, k]) ** 2 / (2 * s[b] ** 2)) for k in range(3)]
        dens += g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    d8 = np.minimum(255, np.rint(255 * amplitude * dens)).astype(np.uint8)
    return d8 > 127


def _upsample_axis(a, axis, n):
    """Linear interpolation of lattice samples along one axis to n points
    (lattice spans the whole axis)."""
    g = a.shape[axis]
    pos = np.arange(n, dtype=np.float32) * ((g - 1) / float(n))
    i0 = np.floor(pos).astype(np.int64)
    i0 = np.minimum(i0, g - 2)
    w = (pos - i0).astype(np.float32)
    a0 = np.take(a, i0, axis=axis)
    a1 = np.take(a, i0 + 1, axis=axis)
    shape = [1] * a.ndim
    shape[axis] = n
    w = w.reshape(shape)
    return a0 * (1 - w) + a1 * w


def value_noise(n, seed, octaves=4, base=4):
    """Sum over octaves k of 0.5^(k+1) * trilinear value noise with
    base*2^k cells per axis; result in [0, 1)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, n, n), np.float32)
    total = 0.0
    for k in range(octaves):
        cells = base << k
        lat = rng.random((cells + 1,) * 3, dtype=np.float32)
        v = _upsample_axis(lat, 0, n)
        v = _upsample_axis(v, 1, n)
        v = _upsample_axis(v, 2, n)
        amp = 0.5 ** (k + 1)
        out += amp * v
        total += amp
    out /= total
    return out


def ct_volume(n=256, seed=7, target=0.08):
    """CT-like occupancy: a porous spherical shell (radius 0.82..0.88 of the
    half edge) plus interior structures where the noise exceeds the quantile
    that brings total occupancy to `target`."""
    noise = value_noise(n, seed)
    x = (np.arange(n, dtype=np.float32) + 0.5) / (n / 2.0) - 1.0
    r2 = (x[:, None, None] ** 2 + x[None, :, None] ** 2 +
          x[None, None, :] ** 2)
    shell = (r2 >= 0.82 ** 2) & (r2 <= 0.88 ** 2) & (noise > 0.35)
    inside = r2 < 0.82 ** 2
    want = int(target * n ** 3) - int(shell.sum())
    vals = noise[inside]
    if want <= 0:
        thr = np.inf
    else:
        thr = np.partition(vals, len(vals) - want)[len(vals) - want]
    interior = inside & (noise >= thr)
    return shell | interior


def bbox_of(dots, voxel=(1.0, 1.0, 1.0)):
    d = np.asarray(dots, np.float32)
    return d.min(axis=0), d.max(axis=0) + np.asarray(voxel, np.float32)


def random_rays(n, bmin, bmax, seed=11, grow=0.1):
    """Origins uniform in the box grown by `grow` x extent on every side,
    directions = normalised N(0, I)."""
    rng = np.random.default_rng(seed)
    bmin = np.asarray(bmin, np.float64)
    bmax = np.asarray(bmax, np.float64)
    ext = bmax - bmin
    lo, hi = bmin - grow * ext, bmax + grow * ext
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def random_balls(n, bmin, bmax, seed=13, rmin=0.5, rmax=8.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(np.asarray(bmin, np.float64), np.asarray(bmax, np.float64),
                    (n, 3)).astype(np.float32)
    if rmin == rmax:
        r = np.full(n, rmin, np.float32)
    else:
        r = rng.uniform(rmin, rmax, n).astype(np.float32)
    return c, r


def orbit_cameras(center, distance, count, W, H, fov_deg=60.0,
                  elevation_deg=20.0, phase=0.0):
    """`count` poses on a circle of radius `distance` around `center`,
    raised by `elevation_deg`, each looking at the centre."""
    fov = math.tan(math.radians(fov_deg) / 2.0)
    cams = []
    el = math.radians(elevation_deg)
    for k in range(count):
        a = phase + 2.0 * math.pi * k / count
        pos = (center[0] + distance * math.cos(el) * math.cos(a),
               center[1] + distance * math.cos(el) * math.sin(a),
               center[2] + distance * math.sin(el))
        cams.append(_camera.make_camera(pos, center, fov, W, H))
    return cams


def config_c1(seed=1):
    """64^3 blobs, 640x480 pinhole camera at a seeded pose outside the
    volume, 10^4 balls of radius 2 uniform in the box."""
    occ = blobs64(seed)
    dots = volume_to_dots(occ)
    bmin, bmax = bbox_of(dots)
    rng = np.random.default_rng(seed + 100)
    center = (bmin + bmax) / 2.0
    a = rng.uniform(0, 2 * math.pi)
    el = rng.uniform(-0.5, 0.5)
    dist = 1.5 * float(np.linalg.norm(bmax - bmin))
    pos = (center[0] + dist * math.cos(el) * math.cos(a),
           center[1] + dist * math.cos(el) * math.sin(a),
           center[2] + dist * math.sin(el))
    cam = _camera.make_camera(pos, center, math.tan(math.radians(30)), 640,
                              480)
    c, r = random_balls(10000, bmin, bmax, seed=seed + 200, rmin=2.0,
                        rmax=2.0)
    return dict(dots=dots, occupancy=float(occ.mean()), camera=cam, W=640,
                H=480, balls_c=c, balls_r=r, bbox=(bmin, bmax))


def config_c2(n=256, seed=7, frames=64, W=1920, H=1080):
    occ = ct_volume(n, seed)
    dots = volume_to_dots(occ)
    bmin, bmax = bbox_of(dots)
    center = tuple(float(v) for v in (np.zeros(3) + n / 2.0))
    diag = math.sqrt(3.0) * n
    cams = orbit_cameras(center, 1.5 * diag, frames, W, H)
    return dict(dots=dots, occupancy=float(occ.mean()), cameras=cams, W=W,
                H=H, bbox=(bmin, bmax))


# ---------------------------------------------------------------------------
# uneven multi-scene batch (32 scenes of 32^3..512^3)
# ---------------------------------------------------------------------------
def menger_volume(edge, level=4):
    """Level-`level` Menger sponge on a 3^level grid, resampled to edge^3:
    a cell is empty when, at some level, two or more of its base-3 digits
    are 1."""
    base = 3 ** level
    idx = (np.arange(edge) * base) // edge
    holes = np.zeros((edge,) * 3, bool)
    for lv in range(level):
        dig = (idx // (3 ** lv)) % 3 == 1
        ones = (dig[:, None, None].astype(np.int8) +
                dig[None, :, None].astype(np.int8) +
                dig[None, None, :].astype(np.int8))
        holes |= ones >= 2
    return ~holes


def shell_volume(edge, occupancy):
    """Hollow sphere: outer radius 0.48*edge, thickness chosen so the shell
    fills `occupancy` of the cube (capped at a full ball)."""
    c = (edge - 1) / 2.0
    R = 0.48 * edge
    t = min(R, occupancy * edge ** 3 / (4.0 * np.pi * R * R))
    x = np.arange(edge, dtype=np.float32) - c
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    return (r2 <= R * R) & (r2 >= (R - t) ** 2)


def noise_volume(edge, occupancy, seed):
    v = value_noise(edge, seed)
    k = max(1, int(occupancy * v.size))
    flat = v.reshape(-1)
    thr = np.partition(flat, flat.size - k)[flat.size - k]
    return v >= thr


def camera_rays(cam, W, H):
    """Host-side pinhole rays of a simple camera, (W*H, 3) origins and
    directions in float32 (row-major pixels)."""
    q = np.array(cam["rotation"], np.float64)
    xs = np.arange(W, dtype=np.float64) * cam["mul"] - cam["xsub"]
    ys = cam["ysub"] - np.arange(H, dtype=np.float64) * cam["mul"]
    vx = np.broadcast_to(xs[None, :], (H, W)).reshape(-1)
    vz = np.broadcast_to(ys[:, None], (H, W)).reshape(-1)
    v = np.stack([vx, np.ones_like(vx), vz], 1)
    w, u = q[0], q[1:]
    t = 2.0 * np.cross(u, v)
    d = v + w * t + np.cross(u, t)
    o = np.broadcast_to(np.asarray(cam["position"], np.float32), d.shape)
    return np.ascontiguousarray(o, np.float32), d.astype(np.float32)


def _c4_scene(kind, edge, occ, nrays, a, seed, s):
    if kind == "noise":
        vol = noise_volume(edge, occ, seed * 1000 + s)
    elif kind == "shell":
        vol = shell_volume(edge, occ)
    else:
        vol = menger_volume(edge)
    dots = volume_to_dots(vol)
    occ_real = float(vol.mean())
    del vol
    bmin, bmax = bbox_of(dots) if len(dots) else (np.zeros(3),
                                                  np.ones(3) * edge)
    ncam = nrays // 2
    W = max(16, int(np.sqrt(ncam * 16 / 9)))
    H = max(1, ncam // W)
    center = (bmin + bmax) / 2.0
    dist = 1.5 * np.sqrt(3.0) * edge
    pos = center + dist * np.array([np.cos(a) * 0.94, np.sin(a) * 0.94,
                                    0.34])
    cam = _camera.make_camera(pos, center, np.tan(np.radians(30)), W, H)
    oc, dc = camera_rays(cam, W, H)
    orr, drr = random_rays(nrays - len(oc), bmin, bmax,
                           seed=seed * 1000 + 500 + s)
    return dict(kind=kind, edge=edge, occupancy=occ_real, dots=dots,
                o=np.concatenate([oc, orr]), d=np.concatenate([dc, drr]),
                camera_pixels=len(oc))


def config_c4(seed=23, n_scenes=32, edge_range=(32, 512),
              ray_range=(1 << 16, 1 << 23), max_edge=None, ray_scale=1.0,
              workers=None):
    """32 scenes of uneven size: edges log-uniform in edge_range, a mix of
    value-noise volumes (1-40% occupancy), hollow sphere shells (1-40%)
    and level-4 Menger sponges; rays per scene log-uniform in ray_range,
    half from a pinhole camera looking at the scene, half random.  The
    scenes' parameters are drawn in order from one generator; the volumes
    are then built on `workers` threads (default: the usable cores, at most
    16), which does not change them."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    params = []
    for s in range(n_scenes):
        edge = int(round(np.exp(rng.uniform(np.log(edge_range[0]),
                                            np.log(edge_range[1])))))
        if max_edge:
            edge = min(edge, max_edge)
        kind = ("noise", "shell", "menger")[s % 3]
        occ = float(rng.uniform(0.01, 0.40))
        nrays = int(np.exp(rng.uniform(np.log(ray_range[0]),
                                       np.log(ray_range[1]))) * ray_scale)
        nrays = max(64, nrays)
        a = rng.uniform(0, 2 * np.pi)
        params.append((kind, edge, occ, nrays, a, seed, s))
    if workers is None:
        workers = min(16, len(os.sched_getaffinity(0)))
    # the largest volumes first, so that the pool finishes together
    order = sorted(range(n_scenes), key=lambda i: -params[i][1])
    out = [None] * n_scenes
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        futs = {i: ex.submit(_c4_scene, *params[i]) for i in order}
        for i, f in futs.items():
            out[i] = f.result()
    return out
