"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and
by the benchmark's cpu_baseline leg, always as the checker or the CPU
baseline, never as the thing measured or shipped.  The oracle restates the
reference's C code (see vox_oracle.h for the file:line map and for what
pins each function).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MISS = 0xFFFFFFFF

QUALITY_BEST = 0
QUALITY_FAST = 1
QUALITY_ADAPTIVE = 2
QUALITY_RAY_MERGE = 4
QUALITY_RAY_MERGE_ACCURATE = 8


def build(force=False):
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE],
                       check=True)
    return LIB_PATH


class _Counters(C.Structure):
    _fields_ = [("node_visits", C.c_uint64), ("voxel_tests", C.c_uint64)]


class _Stats(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("inner", C.c_uint64),
                ("leaves", C.c_uint64), ("dense", C.c_uint64),
                ("overflow", C.c_uint64), ("dots", C.c_uint64),
                ("voxel_ids", C.c_uint64), ("depth", C.c_uint32)]


class _Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("rotation", C.c_float * 4),
                ("fov", C.c_float), ("mul", C.c_float), ("xsub", C.c_float),
                ("ysub", C.c_float)]


class _Shading(C.Structure):
    _fields_ = [("texture", C.c_void_p), ("rshift", C.c_uint32),
                ("gshift", C.c_uint32), ("bshift", C.c_uint32),
                ("amask", C.c_uint32), ("lights", C.c_void_p)]


class _Sphere(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("color", C.c_float * 3),
                ("radius", C.c_float), ("sqr_radius", C.c_float)]


class _RenderStats(C.Structure):
    _fields_ = [("blocks", C.c_uint64), ("fast_blocks", C.c_uint64),
                ("best_blocks", C.c_uint64), ("merge_blocks", C.c_uint64),
                ("root_traversals", C.c_uint64),
                ("leaf_traversals", C.c_uint64), ("c", _Counters)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        F = C.POINTER(C.c_float)
        L.orc_set_voxel.argtypes = [C.c_float] * 3
        L.orc_make_tree_from.argtypes = [F, C.c_size_t]
        L.orc_make_tree_from.restype = P
        L.orc_destroy_tree.argtypes = [P]
        L.orc_tree_stats_get.argtypes = [P, C.POINTER(_Stats)]
        L.orc_tree_voxel_table.argtypes = [P, F]
        L.orc_tree_hash.argtypes = [P]
        L.orc_tree_hash.restype = C.c_uint64
        L.orc_ray_tree_intersection.argtypes = [P, F, F, F]
        L.orc_ray_tree_intersection.restype = P
        L.orc_tree_ball_collidep.argtypes = [P, F, C.c_float]
        L.orc_ray_batch.argtypes = [P, C.c_size_t] + [F] * 6 + [
            C.POINTER(C.c_uint32), F, F, F, C.c_int, C.POINTER(_Counters)]
        L.orc_ball_batch.argtypes = [P, C.c_size_t] + [F] * 4 + [
            C.POINTER(C.c_uint8), C.c_int, C.POINTER(_Counters)]
        L.orc_hit_box.argtypes = [F, F, F, F]
        L.orc_hit_plane_within_box.argtypes = [F, F, F, C.c_int, F, F]
        L.orc_box_ball_interp.argtypes = [F, F, C.c_float]
        L.orc_dense_set_p.argtypes = [F, C.c_size_t]
        L.orc_abs_metric.argtypes = [F, F]
        L.orc_abs_metric.restype = C.c_float
        L.orc_get_corrected_subspace_idx.argtypes = [F, F, F]
        L.orc_get_subspace_idx.argtypes = [F, F]
        L.orc_camera_set_window_size.argtypes = [C.POINTER(_Camera), C.c_int,
                                                 C.c_int]
        L.orc_rotate_vector.argtypes = [F, F, F]
        L.orc_quat_mul.argtypes = [F, F, F]
        L.orc_screen2world.argtypes = [C.POINTER(_Camera), C.c_int, C.c_int,
                                       F]
        L.orc_render.argtypes = [P, C.POINTER(_Camera), C.c_int, C.c_int,
                                 C.c_uint, C.POINTER(_Shading),
                                 C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32), F, C.c_int,
                                 C.POINTER(_RenderStats)]
        L.orc_render_trace.argtypes = [P, C.POINTER(_Camera), C.c_int,
                                       C.c_int, C.c_uint,
                                       C.POINTER(_Shading),
                                       C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32), F, C.c_int,
                                       C.POINTER(_RenderStats),
                                       C.POINTER(C.c_uint8),
                                       C.POINTER(C.c_uint8), F]
        L.orc_ray_max_pending.argtypes = [P, C.c_size_t] + [F] * 6 + [
            C.POINTER(C.c_uint32)]
        L.orc_lights_create.restype = P
        L.orc_lights_destroy.argtypes = [P]
        L.orc_lights_set_ambient.argtypes = [P, F]
        L.orc_lights_insert.argtypes = [P, F, C.c_float, F]
        L.orc_lights_delete.argtypes = [P, F, C.c_float]
        L.orc_lights_clear.argtypes = [P]
        L.orc_lights_count.argtypes = [P]
        L.orc_lights_count.restype = C.c_size_t
        L.orc_get_light.argtypes = [P, F, F]
        L.orc_lights_spheres.argtypes = [P, C.POINTER(_Sphere), C.c_size_t]
        L.orc_lights_spheres.restype = C.c_size_t
        L.orc_lights_height.argtypes = [P]
        L.orc_lights_height.restype = C.c_uint
        L.orc_tree_inner_nodes.argtypes = [P, F, C.POINTER(C.c_uint64),
                                           C.c_size_t]
        L.orc_tree_inner_nodes.restype = C.c_size_t
        L.orc_brute64_rays.argtypes = [F, C.c_size_t, F, C.c_size_t] + [
            F] * 6 + [C.c_double, C.POINTER(C.c_int32),
                      C.POINTER(C.c_double), C.POINTER(C.c_uint8), C.c_int]
        L.orc_camera_rays.argtypes = [C.POINTER(_Camera), C.c_int, C.c_int,
                                      F]
        L.orc_dda64_rays.argtypes = L.orc_brute64_rays.argtypes
        L.orc_dda64_rays.restype = C.c_int
        L.orc_brute_balls.argtypes = [F, C.c_size_t, F, C.c_size_t] + [
            F] * 4 + [C.POINTER(C.c_uint8), C.c_int]
        L.orc_cpu_count.restype = C.c_int
        L.orc_check_tree.argtypes = [P]
        U32 = C.POINTER(C.c_uint32)
        L.orc_copy_squares.argtypes = [U32, U32, C.c_uint, C.c_uint]
        L.orc_cd_collide.argtypes = [P, C.c_size_t, F, F, F,
                                     C.POINTER(C.c_uint8)]
        L.orc_read_raw.argtypes = [C.POINTER(C.c_uint8),
                                   C.POINTER(C.c_uint), C.c_uint,
                                   C.c_uint64, C.c_uint64, F]
        L.orc_read_raw.restype = C.c_size_t
        L.orc_insert_voxel.argtypes = [C.POINTER(P), F]
        L.orc_delete_voxel.argtypes = [C.POINTER(P), F]
        L.orc_rebuild_tree.argtypes = [P]
        L.orc_rebuild_tree.restype = P
        L.orc_make_dense_leaf_box.argtypes = [F, F]
        L.orc_make_dense_leaf_box.restype = P
        L.orc_tree_flags.argtypes = [P]
        L.orc_voxels_in_tree.argtypes = [P]
        L.orc_voxels_in_tree.restype = C.c_size_t
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def set_voxel(v=(1.0, 1.0, 1.0)):
    lib().orc_set_voxel(*[float(x) for x in v])


def cpu_count():
    return int(lib().orc_cpu_count())


class Tree:
    """A tree built by the oracle's restatement of vox_make_tree."""

    def __init__(self, dots_xyz, voxel=(1.0, 1.0, 1.0)):
        L = lib()
        set_voxel(voxel)
        self.voxel = tuple(float(v) for v in voxel)
        d = _f32(dots_xyz).reshape(-1, 3)
        buf = np.zeros((len(d), 4), np.float32)
        buf[:, :3] = d
        self.n = len(d)
        self.ptr = L.orc_make_tree_from(_fp(buf), len(d)) if len(d) else None
        self._restat()

    def _restat(self):
        st = _Stats()
        lib().orc_tree_stats_get(self.ptr, C.byref(st))
        self.stats = {k: getattr(st, k) for k, _ in _Stats._fields_}

    @classmethod
    def empty(cls, voxel=(1.0, 1.0, 1.0)):
        return cls(np.zeros((0, 3), np.float32), voxel)

    @classmethod
    def dense_box(cls, mn, mx, voxel=(1.0, 1.0, 1.0)):
        """A tree that is one dense leaf (vox_make_dense_leaf)."""
        t = cls.empty(voxel)
        t.ptr = lib().orc_make_dense_leaf_box(_fp(_f32(mn)), _fp(_f32(mx)))
        t._restat()
        return t

    def _edit(self, fn, xyz):
        set_voxel(self.voxel)
        p = C.c_void_p(self.ptr)
        r = fn(C.byref(p), _fp(_f32(xyz).reshape(3)))
        self.ptr = p.value
        self._restat()
        return int(r)

    def insert(self, xyz):
        """vox_insert_voxel (tree.c:608-616): 1 when added."""
        return self._edit(lib().orc_insert_voxel, xyz)

    def delete(self, xyz):
        """vox_delete_voxel (tree.c:820-828): 1 when removed."""
        return self._edit(lib().orc_delete_voxel, xyz)

    def rebuild(self):
        """vox_rebuild_tree (tree.c:422-432) as a new Tree."""
        set_voxel(self.voxel)
        t = Tree.empty(self.voxel)
        t.ptr = lib().orc_rebuild_tree(self.ptr)
        t._restat()
        return t

    @property
    def flags(self):
        return int(lib().orc_tree_flags(self.ptr))

    @property
    def voxels(self):
        return int(lib().orc_voxels_in_tree(self.ptr)) if self.ptr else 0

    def close(self):
        if self.ptr:
            lib().orc_destroy_tree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def hash(self):
        return int(lib().orc_tree_hash(self.ptr))

    def check(self):
        """check_tree invariants (tests/tests.c:239-300); 0 = all hold."""
        set_voxel(self.voxel)
        return int(lib().orc_check_tree(self.ptr))

    def voxel_table(self):
        out = np.zeros((max(self.stats["voxel_ids"], 1), 3), np.float32)
        if self.ptr:
            lib().orc_tree_voxel_table(self.ptr, _fp(out))
        return out[: self.stats["voxel_ids"]]

    def rays(self, o, d, nthreads=0, with_counts=False):
        set_voxel(self.voxel)
        o = _f32(o).reshape(-1, 3)
        d = _f32(d).reshape(-1, 3)
        n = len(o)
        ox, oy, oz = (_f32(o[:, k]) for k in range(3))
        dx, dy, dz = (_f32(d[:, k]) for k in range(3))
        vox = np.full(n, MISS, np.uint32)
        hit = np.full((3, n), np.nan, np.float32)
        cnt = _Counters()
        lib().orc_ray_batch(self.ptr, n, _fp(ox), _fp(oy), _fp(oz), _fp(dx),
                            _fp(dy), _fp(dz),
                            vox.ctypes.data_as(C.POINTER(C.c_uint32)),
                            _fp(hit[0]), _fp(hit[1]), _fp(hit[2]), nthreads,
                            C.byref(cnt))
        if with_counts:
            return vox, hit, (cnt.node_visits, cnt.voxel_tests)
        return vox, hit

    def cd_collide(self, pos, last, radius):
        """vox_cd_collide for n bodies (cd.c:38-55): returns (new
        positions, new last positions, collided flags)."""
        set_voxel(self.voxel)
        pos = _f32(pos).reshape(-1, 3).copy()
        last = _f32(last).reshape(-1, 3).copy()
        radius = _f32(radius).reshape(-1)
        out = np.zeros(len(pos), np.uint8)
        lib().orc_cd_collide(self.ptr, len(pos), _fp(pos), _fp(last),
                             _fp(radius),
                             out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return pos, last, out

    def balls(self, c, r, nthreads=0, with_counts=False):
        set_voxel(self.voxel)
        c = _f32(c).reshape(-1, 3)
        r = _f32(r).reshape(-1)
        n = len(c)
        cx, cy, cz = (_f32(c[:, k]) for k in range(3))
        out = np.zeros(n, np.uint8)
        cnt = _Counters()
        lib().orc_ball_batch(self.ptr, n, _fp(cx), _fp(cy), _fp(cz), _fp(r),
                             out.ctypes.data_as(C.POINTER(C.c_uint8)),
                             nthreads, C.byref(cnt))
        if with_counts:
            return out, (cnt.node_visits, cnt.voxel_tests)
        return out

    def render(self, cam, W, H, quality=QUALITY_BEST, texture=None,
               fmt=(16, 8, 0, 0xFF000000), nthreads=0, squares=None,
               lights=None):
        """Returns (squares, voxel_px, hit_px, stats) in the square layout;
        lights: a Lights object (the context's light manager) or None."""
        set_voxel(self.voxel)
        c = camera_struct(cam)
        sh = _Shading()
        sh.lights = lights.ptr if lights is not None else None
        tex = None
        if texture is not None:
            tex = np.ascontiguousarray(texture, dtype=np.uint8)
            sh.texture = tex.ctypes.data
        sh.rshift, sh.gshift, sh.bshift, sh.amask = fmt
        if squares is None:
            squares = np.zeros(W * H, np.uint32)
        vox = np.full(W * H, MISS, np.uint32)
        hit = np.full((3, W * H), np.nan, np.float32)
        st = _RenderStats()
        rc = lib().orc_render(self.ptr, C.byref(c), W, H, quality,
                              C.byref(sh),
                              squares.ctypes.data_as(C.POINTER(C.c_uint32)),
                              vox.ctypes.data_as(C.POINTER(C.c_uint32)),
                              _fp(hit), nthreads, C.byref(st))
        if rc != 0:
            raise ValueError("orc_render rejected the arguments (%d)" % rc)
        stats = {k: getattr(st, k) for k, _ in _RenderStats._fields_
                 if k != "c"}
        stats["node_visits"] = st.c.node_visits
        stats["voxel_tests"] = st.c.voxel_tests
        return squares, vox, hit, stats


class Lights:
    """The oracle's light manager (src/voxrnd/lights.c): an M-tree of
    shadowless light spheres and an ambient light."""

    def __init__(self):
        self.ptr = lib().orc_lights_create()

    def set_ambient(self, color):
        return int(lib().orc_lights_set_ambient(self.ptr, _fp(_f32(color))))

    def insert(self, center, radius, color):
        return int(lib().orc_lights_insert(self.ptr, _fp(_f32(center)),
                                           float(radius), _fp(_f32(color))))

    def delete(self, center, radius):
        return int(lib().orc_lights_delete(self.ptr, _fp(_f32(center)),
                                           float(radius)))

    def clear(self):
        lib().orc_lights_clear(self.ptr)

    def __len__(self):
        return int(lib().orc_lights_count(self.ptr))

    def height(self):
        return int(lib().orc_lights_height(self.ptr))

    def light(self, p):
        """vox_get_light at each of the (n, 3) points."""
        p = _f32(p).reshape(-1, 3)
        out = np.zeros((len(p), 3), np.float32)
        for i in range(len(p)):
            lib().orc_get_light(self.ptr, _fp(np.ascontiguousarray(p[i])),
                                out[i].ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def spheres(self):
        """(centers (k, 3), colors (k, 3), radii (k,)) in walk order."""
        k = lib().orc_lights_spheres(self.ptr, None, 0)
        arr = (_Sphere * max(1, k))()
        lib().orc_lights_spheres(self.ptr, arr, k)
        c = np.array([list(a.center) for a in arr[:k]], np.float32)
        col = np.array([list(a.color) for a in arr[:k]], np.float32)
        r = np.array([a.radius for a in arr[:k]], np.float32)
        return c.reshape(-1, 3), col.reshape(-1, 3), r

    def close(self):
        if self.ptr:
            lib().orc_lights_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_trace(tree, cam, W, H, quality, nthreads=0):
    """orc_render_trace: (voxel_px, hit_px, block_flags, px_src, try_box)
    in the square layout (see vox_oracle.h)."""
    set_voxel(tree.voxel)
    c = camera_struct(cam)
    sh = _Shading()
    vox = np.full(W * H, MISS, np.uint32)
    hit = np.full((3, W * H), np.nan, np.float32)
    flags = np.full((W // 4) * (H // 4), 255, np.uint8)
    src = np.full(W * H, 255, np.uint8)
    box = np.full((6, W * H), np.nan, np.float32)
    st = _RenderStats()
    rc = lib().orc_render_trace(
        tree.ptr, C.byref(c), W, H, quality, C.byref(sh), None,
        vox.ctypes.data_as(C.POINTER(C.c_uint32)), _fp(hit), nthreads,
        C.byref(st), flags.ctypes.data_as(C.POINTER(C.c_uint8)),
        src.ctypes.data_as(C.POINTER(C.c_uint8)), _fp(box))
    if rc != 0:
        raise ValueError("orc_render_trace rejected the arguments (%d)" % rc)
    return vox, hit, flags, src, box


def camera_rays(cam, W, H):
    """(W*H, 3) fp32 ray directions of a frame in the square layout."""
    c = camera_struct(cam)
    out = np.zeros((3, W * H), np.float32)
    lib().orc_camera_rays(C.byref(c), W, H, _fp(out))
    return np.ascontiguousarray(out.T)


def ray_max_pending(tree, o, d):
    """orc_ray_max_pending over (n, 3) rays."""
    set_voxel(tree.voxel)
    o = _f32(o).reshape(-1, 3)
    d = _f32(d).reshape(-1, 3)
    cols = [_f32(o[:, k]) for k in range(3)] + [_f32(d[:, k])
                                                for k in range(3)]
    out = np.zeros(len(o), np.uint32)
    lib().orc_ray_max_pending(tree.ptr, len(o), *[_fp(c) for c in cols],
                              out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def inner_nodes(tree):
    """(boxes_min (k,3), boxes_max (k,3), centres (k,3), counts (k,)) of
    the tree's inner nodes in depth-first preorder."""
    if not tree.ptr:
        z = np.zeros((0, 3), np.float32)
        return z, z, z, np.zeros(0, np.uint64)
    k = lib().orc_tree_inner_nodes(tree.ptr, None, None, 0)
    out = np.zeros((max(k, 1), 9), np.float32)
    cnt = np.zeros(max(k, 1), np.uint64)
    lib().orc_tree_inner_nodes(tree.ptr, _fp(out),
                               cnt.ctypes.data_as(C.POINTER(C.c_uint64)), k)
    return out[:k, 0:3], out[:k, 3:6], out[:k, 6:9], cnt[:k]


def camera_struct(cam):
    """cam: dict with position, rotation (w,i,j,k), mul, xsub, ysub."""
    c = _Camera()
    for k in range(3):
        c.position[k] = float(cam["position"][k])
    for k in range(4):
        c.rotation[k] = float(cam["rotation"][k])
    c.fov = float(cam.get("fov", 1.0))
    c.mul = float(cam["mul"])
    c.xsub = float(cam["xsub"])
    c.ysub = float(cam["ysub"])
    return c


def copy_squares(squares, W, H):
    """copy_squares (copy-helper.c:5-40) of one W x H frame."""
    sq = np.ascontiguousarray(squares, np.uint32).reshape(-1)
    px = np.zeros(W * H, np.uint32)
    U32 = C.POINTER(C.c_uint32)
    lib().orc_copy_squares(sq.ctypes.data_as(U32), px.ctypes.data_as(U32),
                           W // 4, H // 4)
    return px.reshape(H, W)


def read_raw(data, dim, samplesize, lo, hi, voxel=(1.0, 1.0, 1.0)):
    """Sample loop of vox_read_raw_data (datareader.c:49-75) with the
    test lo <= sample < hi; returns the selected voxel corners (m, 3)."""
    set_voxel(voxel)
    data = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8)
                                if not isinstance(data, np.ndarray)
                                else data.view(np.uint8).reshape(-1))
    d = (C.c_uint * 3)(*[int(x) for x in dim])
    n = int(dim[0]) * int(dim[1]) * int(dim[2])
    out = np.zeros((max(n, 1), 3), np.float32)
    m = lib().orc_read_raw(data.ctypes.data_as(C.POINTER(C.c_uint8)), d,
                           int(samplesize), int(lo), int(hi), _fp(out))
    return out[:m].copy()


def set_window_size(fov, w, h):
    c = _Camera()
    c.fov = float(fov)
    lib().orc_camera_set_window_size(C.byref(c), int(w), int(h))
    return float(c.mul), float(c.xsub), float(c.ysub)


def screen2world(cam, sx, sy):
    c = camera_struct(cam)
    out = np.zeros(4, np.float32)
    lib().orc_screen2world(C.byref(c), int(sx), int(sy), _fp(out))
    return out[:3].copy()


def rotate_vector(q, v):
    q = _f32(q)
    v = _f32(v)
    out = np.zeros(3, np.float32)
    lib().orc_rotate_vector(_fp(q), _fp(v), _fp(out))
    return out


def quat_mul(q1, q2):
    q1 = _f32(q1)
    q2 = _f32(q2)
    out = np.zeros(4, np.float32)
    lib().orc_quat_mul(_fp(q1), _fp(q2), _fp(out))
    return out


def _dot4(v):
    a = np.zeros(4, np.float32)
    a[:3] = v[:3]
    return a


def _box(mn, mx):
    b = np.zeros(8, np.float32)
    b[:3] = mn
    b[4:7] = mx
    return b


def hit_box(bmin, bmax, o, d):
    b = _box(bmin, bmax)
    o4, d4 = _dot4(o), _dot4(d)
    res = np.zeros(4, np.float32)
    h = lib().orc_hit_box(_fp(b), _fp(o4), _fp(d4), _fp(res))
    return bool(h), res[:3].copy()


def hit_plane_within_box(o, d, center, plane, bmin, bmax):
    b = _box(bmin, bmax)
    res = np.zeros(4, np.float32)
    h = lib().orc_hit_plane_within_box(_fp(_dot4(o)), _fp(_dot4(d)),
                                       _fp(_dot4(center)), int(plane),
                                       _fp(res), _fp(b))
    return bool(h), res[:3].copy()


def box_ball_interp(bmin, bmax, c, r):
    b = _box(bmin, bmax)
    return bool(lib().orc_box_ball_interp(_fp(b), _fp(_dot4(c)), float(r)))


def dense_set_p(bmin, bmax, n):
    return bool(lib().orc_dense_set_p(_fp(_box(bmin, bmax)), int(n)))


def corrected_subspace(center, dot, direction):
    return int(lib().orc_get_corrected_subspace_idx(
        _fp(_dot4(center)), _fp(_dot4(dot)), _fp(_dot4(direction))))


def subspace(center, dot):
    return int(lib().orc_get_subspace_idx(_fp(_dot4(center)),
                                          _fp(_dot4(dot))))


def abs_metric(a, b):
    return float(lib().orc_abs_metric(_fp(_dot4(a)), _fp(_dot4(b))))


def brute64_rays(dots_xyz, o, d, voxel=(1.0, 1.0, 1.0), eps=1e-6,
                 nthreads=0):
    """FP64 brute force: (voxel index into dots or -1, hit xyz f64, tie)."""
    dots = _f32(dots_xyz).reshape(-1, 3)
    o = _f32(o).reshape(-1, 3)
    d = _f32(d).reshape(-1, 3)
    n = len(o)
    V = _f32(voxel)
    cols = [_f32(o[:, k]) for k in range(3)] + [_f32(d[:, k])
                                                for k in range(3)]
    idx = np.full(n, -1, np.int32)
    hit = np.full((n, 3), np.nan, np.float64)
    tie = np.zeros(n, np.uint8)
    lib().orc_brute64_rays(_fp(dots), len(dots), _fp(V), n,
                           *[_fp(c) for c in cols], float(eps),
                           idx.ctypes.data_as(C.POINTER(C.c_int32)),
                           hit.ctypes.data_as(C.POINTER(C.c_double)),
                           tie.ctypes.data_as(C.POINTER(C.c_uint8)),
                           nthreads)
    return idx, hit, tie.astype(bool)


def dda64_rays(dots_xyz, o, d, voxel=(1.0, 1.0, 1.0), eps=1e-6,
               nthreads=0):
    """FP64 grid walk with the brute force's answers (see vox_oracle.h):
    (voxel index into dots or -1, hit xyz f64, tie)."""
    dots = _f32(dots_xyz).reshape(-1, 3)
    o = _f32(o).reshape(-1, 3)
    d = _f32(d).reshape(-1, 3)
    n = len(o)
    V = _f32(voxel)
    cols = [_f32(o[:, k]) for k in range(3)] + [_f32(d[:, k])
                                                for k in range(3)]
    idx = np.full(n, -1, np.int32)
    hit = np.full((n, 3), np.nan, np.float64)
    tie = np.zeros(n, np.uint8)
    rc = lib().orc_dda64_rays(_fp(dots), len(dots), _fp(V), n,
                              *[_fp(c) for c in cols], float(eps),
                              idx.ctypes.data_as(C.POINTER(C.c_int32)),
                              hit.ctypes.data_as(C.POINTER(C.c_double)),
                              tie.ctypes.data_as(C.POINTER(C.c_uint8)),
                              nthreads)
    if rc != 0:
        raise ValueError("orc_dda64_rays failed (%d)" % rc)
    return idx, hit, tie.astype(bool)


def brute_balls(dots_xyz, c, r, voxel=(1.0, 1.0, 1.0), nthreads=0):
    dots = _f32(dots_xyz).reshape(-1, 3)
    c = _f32(c).reshape(-1, 3)
    r = _f32(r).reshape(-1)
    V = _f32(voxel)
    out = np.zeros(len(c), np.uint8)
    lib().orc_brute_balls(_fp(dots), len(dots), _fp(V), len(c),
                          *[_fp(_f32(c[:, k])) for k in range(3)], _fp(r),
                          out.ctypes.data_as(C.POINTER(C.c_uint8)), nthreads)
    return out
