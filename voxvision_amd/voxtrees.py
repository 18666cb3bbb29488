"""Python mirror of the voxtrees interface, backed by libvoxgpu.

Reference interface (src/voxtrees/tree.h, search.h):
  vox_make_tree(set, n)                       -> make_tree(dots, voxel)
  vox_destroy_tree(tree)                      -> Tree.release()
  vox_voxels_in_tree / vox_bounding_box       -> Tree.info
  vox_ray_tree_intersection(tree, o, d, res)  -> ray_tree_intersection(tree, o, d)
  vox_tree_ball_collidep(tree, c, r)          -> tree_ball_collidep(tree, c, r)
The searches are batched: they take (n, 3) device tensors and return
device tensors; a ray that misses gets voxel id VOXG_MISS and an untouched
(NaN-initialised) hit point, as res is untouched in the reference.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import VOXG_MISS, check


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _dev_f32(x, n=None, cols=None):
    t = torch.as_tensor(x)
    if t.device.type != "cuda":
        t = t.to("cuda")
    t = t.to(torch.float32).contiguous()
    return t


class Tree:
    """A flattened tree resident in device memory (immutable)."""

    def __init__(self, handle, voxel):
        self._h = handle
        self.voxel = tuple(float(v) for v in voxel)
        info = _lib.TreeInfo()
        check(_lib.load().voxg_tree_get_info(self._h, C.byref(info)),
              "voxg_tree_get_info")
        self.info = {k: (list(getattr(info, k)) if k in
                         ("bbox_min", "bbox_max", "voxel")
                         else getattr(info, k)) for k, _ in info._fields_}

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("tree was released")
        return self._h

    @classmethod
    def build(cls, dots, voxel=(1.0, 1.0, 1.0)):
        """vox_make_tree over an (n, 3) or (n, 4) float array (host)."""
        L = _lib.load()
        d = np.ascontiguousarray(np.asarray(dots, dtype=np.float32))
        if d.ndim != 2 or d.shape[1] not in (3, 4):
            raise ValueError("dots must be (n, 3) or (n, 4)")
        v = (C.c_float * 3)(*voxel)
        h = C.c_void_p()
        check(L.voxg_tree_build(d.ctypes.data if len(d) else None, len(d),
                                d.shape[1], v, C.byref(h)),
              "voxg_tree_build")
        return cls(h, voxel)

    @classmethod
    def upload(cls, root_ptr, voxel=(1.0, 1.0, 1.0)):
        """Flatten a tree in the reference's struct vox_node layout."""
        L = _lib.load()
        v = (C.c_float * 3)(*voxel)
        h = C.c_void_p()
        check(L.voxg_tree_upload(root_ptr, v, C.byref(h)), "voxg_tree_upload")
        return cls(h, voxel)

    def release(self):
        if self._h is not None:
            _lib.load().voxg_tree_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def hash(self):
        return int(_lib.load().voxg_tree_hash(self.handle))

    def voxel_coords(self, ids):
        ids = np.ascontiguousarray(np.asarray(ids, dtype=np.uint32))
        out = np.zeros((len(ids), 3), np.float32)
        check(_lib.load().voxg_tree_voxel_coords(
            self.handle, len(ids), ids.ctypes.data, out.ctypes.data),
            "voxg_tree_voxel_coords")
        return out

    def node_ptr(self, idx):
        return _lib.load().voxg_tree_node_ptr(self.handle, int(idx))


def _info_dict(info):
    return {k: (list(getattr(info, k)) if k in ("bbox_min", "bbox_max",
                                                "voxel")
                else getattr(info, k)) for k, _ in info._fields_}


def plan(dots, voxel=(1.0, 1.0, 1.0)):
    """Host-only build: (info dict, structure hash); no device needed."""
    L = _lib.load()
    d = np.ascontiguousarray(np.asarray(dots, dtype=np.float32))
    v = (C.c_float * 3)(*voxel)
    info = _lib.TreeInfo()
    h = C.c_uint64()
    check(L.voxg_tree_plan(d.ctypes.data if len(d) else None, len(d),
                           d.shape[1] if d.ndim == 2 else 3, v,
                           C.byref(info), C.byref(h)), "voxg_tree_plan")
    return _info_dict(info), int(h.value)


def plan_reference(root_ptr, voxel=(1.0, 1.0, 1.0)):
    """Host-only flattening of a reference-layout tree."""
    L = _lib.load()
    v = (C.c_float * 3)(*voxel)
    info = _lib.TreeInfo()
    h = C.c_uint64()
    check(L.voxg_tree_plan_reference(root_ptr, v, C.byref(info),
                                     C.byref(h)), "voxg_tree_plan_reference")
    return _info_dict(info), int(h.value)


def make_tree(dots, voxel=(1.0, 1.0, 1.0)):
    return Tree.build(dots, voxel)


def ray_tree_intersection(tree, o, d, want_leaf=False, hit_init=None,
                          stream=None):
    """Batched vox_ray_tree_intersection.  o, d: (n, 3) arrays or tensors.
    Returns (voxel_id uint32-as-int64 tensor, hit (3, n) float32 tensor
    [, leaf node index tensor])."""
    L = _lib.load()
    o = _dev_f32(o).reshape(-1, 3)
    d = _dev_f32(d).reshape(-1, 3)
    n = o.shape[0]
    oc = o.t().contiguous()
    dc = d.t().contiguous()
    # every traced ray overwrites its entry (an id or VOXG_MISS); -2 marks
    # a ray that was not traced
    vox = torch.full((n,), -2, dtype=torch.int32, device="cuda")
    leaf = torch.empty(n, dtype=torch.int32, device="cuda") if want_leaf \
        else None
    if hit_init is None:
        hit = torch.full((3, n), float("nan"), dtype=torch.float32,
                         device="cuda")
    else:
        hit = hit_init
    check(L.voxg_ray_batch(tree.handle, n, oc[0].data_ptr(),
                           oc[1].data_ptr(), oc[2].data_ptr(),
                           dc[0].data_ptr(), dc[1].data_ptr(),
                           dc[2].data_ptr(), vox.data_ptr(),
                           leaf.data_ptr() if leaf is not None else None,
                           hit[0].data_ptr(), hit[1].data_ptr(),
                           hit[2].data_ptr(), _stream(stream)),
          "voxg_ray_batch")
    vox = vox.to(torch.int64) & 0xFFFFFFFF
    if want_leaf:
        return vox, hit, leaf.to(torch.int64) & 0xFFFFFFFF
    return vox, hit


def ray_counts(tree, o, d, stream=None):
    """Node visits and voxel tests of the traversal of these rays."""
    L = _lib.load()
    o = _dev_f32(o).reshape(-1, 3).t().contiguous()
    d = _dev_f32(d).reshape(-1, 3).t().contiguous()
    c = _lib.Counts()
    check(L.voxg_ray_batch_count(tree.handle, o.shape[1], o[0].data_ptr(),
                                 o[1].data_ptr(), o[2].data_ptr(),
                                 d[0].data_ptr(), d[1].data_ptr(),
                                 d[2].data_ptr(), C.byref(c),
                                 _stream(stream)), "voxg_ray_batch_count")
    return dict(node_visits=c.node_visits, voxel_tests=c.voxel_tests,
                rays=c.rays, stack_spills=c.stack_spills)


def tree_ball_collidep(tree, c, r, stream=None):
    """Batched vox_tree_ball_collidep: (n,3) centres, (n,) radii -> uint8."""
    L = _lib.load()
    c = _dev_f32(c).reshape(-1, 3).t().contiguous()
    r = _dev_f32(r).reshape(-1).contiguous()
    n = c.shape[1]
    out = torch.empty(n, dtype=torch.uint8, device="cuda")
    check(L.voxg_ball_batch(tree.handle, n, c[0].data_ptr(), c[1].data_ptr(),
                            c[2].data_ptr(), r.data_ptr(), out.data_ptr(),
                            _stream(stream)), "voxg_ball_batch")
    return out


def ball_counts(tree, c, r, stream=None):
    L = _lib.load()
    c = _dev_f32(c).reshape(-1, 3).t().contiguous()
    r = _dev_f32(r).reshape(-1).contiguous()
    cnt = _lib.Counts()
    check(L.voxg_ball_batch_count(tree.handle, c.shape[1], c[0].data_ptr(),
                                  c[1].data_ptr(), c[2].data_ptr(),
                                  r.data_ptr(), C.byref(cnt),
                                  _stream(stream)), "voxg_ball_batch_count")
    return dict(node_visits=cnt.node_visits, voxel_tests=cnt.voxel_tests,
                rays=cnt.rays)


def _raw_range(samplesize, lo, hi, threshold):
    top = 1 << (8 * samplesize)
    if threshold is not None:  # the demo's test: sample > threshold
        return int(threshold) + 1, top
    # the Lua ranged reader's defaults (voxengine/modules/voxtrees.c:313-316)
    lo = (1 << (8 * samplesize - 1)) if lo is None else int(lo)
    hi = top if hi is None else int(hi)
    return lo, hi


def raw_select(samples, dim, samplesize, lo=None, hi=None, threshold=None,
               voxel=(1.0, 1.0, 1.0), stream=None):
    """Sample loop of vox_read_raw_data (datareader.c:49-75) on the device:
    samples is a device uint8 tensor of dim[0]*dim[1]*dim[2]*samplesize
    bytes; the test is lo <= sample < hi (or sample > threshold).  Returns
    the selected voxel corners as a device (m, 3) float tensor, in the
    reader's order: one pass over the samples (voxg_raw_count), the count
    to the host, then the corners from the selection it recorded
    (voxg_raw_write)."""
    L = _lib.load()
    lo, hi = _raw_range(samplesize, lo, hi, threshold)
    d = (C.c_uint * 3)(*[int(x) for x in dim])
    v = (C.c_float * 3)(*voxel)
    nb = L.voxg_raw_workspace_bytes(d)
    if nb == 0:
        raise _lib.VoxgError("voxg_raw_workspace_bytes: the volume must hold "
                             "fewer than 2^32 samples")
    work = torch.empty(nb, dtype=torch.uint8, device="cuda")
    cnt = torch.empty(1, dtype=torch.int64, device="cuda")
    st = _stream(stream)
    check(L.voxg_raw_count(samples.data_ptr(), d, samplesize, lo, hi,
                           work.data_ptr(), cnt.data_ptr(), st),
          "voxg_raw_count")
    m = int(cnt.item())
    xyz = torch.empty((m, 3), dtype=torch.float32, device="cuda")
    if m:
        check(L.voxg_raw_write(d, v, work.data_ptr(), xyz.data_ptr(), m, st),
              "voxg_raw_write")
    return xyz


def read_raw_data(filename, dim, samplesize, lo=None, hi=None, threshold=None,
                  voxel=(1.0, 1.0, 1.0)):
    """vox_read_raw_data (datareader.c:11-85): file -> device tree."""
    L = _lib.load()
    lo, hi = _raw_range(samplesize, lo, hi, threshold)
    d = (C.c_uint * 3)(*[int(x) for x in dim])
    v = (C.c_float * 3)(*voxel)
    h = C.c_void_p()
    check(L.voxg_read_raw_data(str(filename).encode(), d, samplesize, lo, hi,
                               v, C.byref(h)), "voxg_read_raw_data")
    return Tree(h, voxel)


# kinds of traversal launch (voxg_set_kind_launch_config)
LAUNCH_KINDS = {"rays": 0, "render": 1, "render_best": 2, "batch": 3}


def _kind(kind):
    if kind not in LAUNCH_KINDS:
        raise ValueError("unknown kind of launch %r (one of %s)"
                         % (kind, ", ".join(LAUNCH_KINDS)))
    return LAUNCH_KINDS[kind]


def set_launch_config(kind=None, **fields):
    """Change fields of the traversal launch configuration
    (voxg_launch_config: stack_lds, block, waves_per_cu, refill, chunk,
    render_cull, leaf_batch, finish_batch, split_best) of one kind of launch
    ("rays", "render", "render_best", "batch"), or of every kind when kind
    is None; the other fields keep each kind's values.  Returns the new
    configuration of the kind (of "rays" when kind is None)."""
    L = _lib.load()
    kinds = list(LAUNCH_KINDS) if kind is None else [kind]
    for k in kinds:
        cfg = _lib.LaunchConfig()
        check(L.voxg_get_kind_launch_config(_kind(k), C.byref(cfg)))
        names = [n for n, _ in cfg._fields_]
        for n, v in fields.items():
            if n not in names:
                raise TypeError("unknown launch configuration field %r" % n)
            if v is not None:
                setattr(cfg, n, int(v))
        check(L.voxg_set_kind_launch_config(_kind(k), C.byref(cfg)),
              "voxg_set_kind_launch_config")
    return get_launch_config(kind or "rays")


def get_launch_config(kind="rays"):
    cfg = _lib.LaunchConfig()
    check(_lib.load().voxg_get_kind_launch_config(_kind(kind),
                                                  C.byref(cfg)))
    return {k: getattr(cfg, k) for k, _ in cfg._fields_}


def get_launch_configs():
    """Every kind's configuration, {kind: fields}."""
    return {k: get_launch_config(k) for k in LAUNCH_KINDS}


def set_launch_configs(cfgs):
    """Restore configurations returned by get_launch_configs()."""
    for k, f in cfgs.items():
        set_launch_config(k, **f)


class saved_launch_configs:
    """Context manager that restores every kind's launch configuration."""

    def __enter__(self):
        self._saved = get_launch_configs()
        return self._saved

    def __exit__(self, *exc):
        set_launch_configs(self._saved)
        return False
