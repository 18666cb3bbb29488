"""Python front end of the serving loop (voxg_server_* in include/voxgpu.h).

A Server keeps trees resident on the device and runs render and ball
requests submitted from any thread; requests admitted together (by free
request slots and the device-memory budget) run as one launch group, and
groups overlap on several streams.  Result buffers are pinned host tensors
so the device-to-host copies stay asynchronous.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_inflight", C.c_int),
                ("streams", C.c_int), ("texture", C.c_void_p),
                ("rshift", C.c_uint32), ("gshift", C.c_uint32),
                ("bshift", C.c_uint32), ("amask", C.c_uint32)]


class RenderRequest(C.Structure):
    _fields_ = [("tree", C.c_int), ("cam", _lib.Camera), ("W", C.c_int),
                ("H", C.c_int), ("quality", C.c_uint),
                ("voxel_out", C.c_void_p), ("hit_out", C.c_void_p),
                ("color_out", C.c_void_p)]


class BallRequest(C.Structure):
    _fields_ = [("tree", C.c_int), ("n", C.c_size_t), ("cx", C.c_void_p),
                ("cy", C.c_void_p), ("cz", C.c_void_p), ("r", C.c_void_p),
                ("out", C.c_void_p)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double),
                ("group", C.c_uint64), ("group_size", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64),
                ("rejected", C.c_uint64), ("groups", C.c_uint64),
                ("peak_inflight", C.c_uint64),
                ("peak_groups_inflight", C.c_uint64),
                ("peak_bytes", C.c_uint64), ("budget", C.c_uint64)]


def _bind(L):
    if getattr(L, "_server_bound", False):
        return
    P = C.c_void_p
    L.voxg_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(P)]
    L.voxg_server_add_tree.argtypes = [P, P, C.POINTER(C.c_int)]
    L.voxg_server_replace_tree.argtypes = [P, C.c_int, P]
    L.voxg_server_submit_render.argtypes = [P, C.POINTER(RenderRequest),
                                            C.POINTER(C.c_uint64)]
    L.voxg_server_submit_balls.argtypes = [P, C.POINTER(BallRequest),
                                           C.POINTER(C.c_uint64)]
    L.voxg_server_wait.argtypes = [P, C.c_uint64, C.c_double,
                                   C.POINTER(RequestInfo)]
    L.voxg_server_get_stats.argtypes = [P, C.POINTER(ServerStats)]
    L.voxg_server_now.argtypes = [P]
    L.voxg_server_now.restype = C.c_double
    L.voxg_server_destroy.argtypes = [P]
    L._server_bound = True


def camera_struct(cam):
    c = _lib.Camera()
    for k in range(3):
        c.position[k] = float(cam["position"][k])
    for k in range(4):
        c.rotation[k] = float(cam["rotation"][k])
    c.mul, c.xsub, c.ysub = cam["mul"], cam["xsub"], cam["ysub"]
    c.fov = cam.get("fov", 1.0)
    return c


class Server:
    def __init__(self, hbm_budget=0, max_inflight=0, streams=0, texture=None,
                 fmt=(16, 8, 0, 0xFF000000)):
        L = _lib.load()
        _bind(L)
        cfg = ServerConfig()
        cfg.hbm_budget = int(hbm_budget)
        cfg.max_inflight = int(max_inflight)
        cfg.streams = int(streams)
        self._tex = None
        if texture is not None:
            self._tex = np.ascontiguousarray(texture, np.uint8)
            cfg.texture = self._tex.ctypes.data
        cfg.rshift, cfg.gshift, cfg.bshift, cfg.amask = fmt
        self._h = C.c_void_p()
        check(L.voxg_server_create(C.byref(cfg), C.byref(self._h)),
              "voxg_server_create")
        self._trees = []
        self._keep = {}  # ticket -> buffers that must outlive the request

    def add_tree(self, tree):
        i = C.c_int()
        check(_lib.load().voxg_server_add_tree(self._h, tree.handle,
                                               C.byref(i)),
              "voxg_server_add_tree")
        self._trees.append(tree)
        return i.value

    def replace_tree(self, tree_id, tree):
        """New version of a resident tree (e.g. EditTree.commit()): requests
        submitted from now on use it; earlier ones keep theirs.  Replaced
        versions are kept alive as long as this server object."""
        check(_lib.load().voxg_server_replace_tree(self._h, int(tree_id),
                                                   tree.handle),
              "voxg_server_replace_tree")
        self._trees.append(tree)

    def submit_render(self, tree_id, cam, W, H, quality=0, colors=False,
                      hits=True, out=None):
        """Returns (ticket, outputs dict of pinned host tensors); `out`
        may supply the pinned buffers (keys voxel, hit, colors)."""
        n = W * H
        if out is None:
            out = {"voxel": torch.empty(n, dtype=torch.int32,
                                        pin_memory=True)}
            if hits:
                out["hit"] = torch.empty(3 * n, dtype=torch.float32,
                                         pin_memory=True)
        if colors and "colors" not in out:
            out["colors"] = torch.zeros(n, dtype=torch.int32,
                                        pin_memory=True)
        r = RenderRequest()
        r.tree = tree_id
        r.cam = camera_struct(cam)
        r.W, r.H, r.quality = W, H, quality
        r.voxel_out = out["voxel"].data_ptr()
        r.hit_out = out["hit"].data_ptr() if "hit" in out else None
        r.color_out = out["colors"].data_ptr() if colors else None
        t = C.c_uint64()
        check(_lib.load().voxg_server_submit_render(self._h, C.byref(r),
                                                    C.byref(t)),
              "voxg_server_submit_render")
        self._keep[t.value] = out
        return t.value, out

    def submit_balls(self, tree_id, c, r):
        c = np.ascontiguousarray(np.asarray(c, np.float32).T)
        r = np.ascontiguousarray(r, np.float32)
        n = len(r)
        # every byte is written by the server (an empty tensor: a zero fill
        # would run on torch's CPU thread pool on the submitting thread)
        out = torch.empty(n, dtype=torch.uint8, pin_memory=True)
        q = BallRequest()
        q.tree = tree_id
        q.n = n
        q.cx, q.cy, q.cz = (c[k].ctypes.data for k in range(3))
        q.r = r.ctypes.data
        q.out = out.data_ptr()
        t = C.c_uint64()
        check(_lib.load().voxg_server_submit_balls(self._h, C.byref(q),
                                                   C.byref(t)),
              "voxg_server_submit_balls")
        self._keep[t.value] = out
        return t.value, out

    def wait(self, ticket, timeout=0.0):
        info = RequestInfo()
        check(_lib.load().voxg_server_wait(self._h, ticket, timeout,
                                           C.byref(info)),
              "voxg_server_wait")
        self._keep.pop(ticket, None)
        return {k: getattr(info, k) for k, _ in info._fields_}

    def stats(self):
        st = ServerStats()
        check(_lib.load().voxg_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def now(self):
        return _lib.load().voxg_server_now(self._h)

    def close(self):
        if self._h is not None and self._h.value:
            _lib.load().voxg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
