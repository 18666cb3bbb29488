"""Editable trees: vox_insert_voxel / vox_delete_voxel / vox_rebuild_tree
(src/voxtrees/tree.c:422-835) on a host tree in the reference's node
layout, committed to the device as new tree versions.

    E = EditTree(dots)            # vox_make_tree
    E.insert((x, y, z))           # True when added
    E.delete((x, y, z))           # True when removed
    G = E.commit()                # a new device Tree; older ones unchanged
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .voxtrees import Tree


class EditTree:
    def __init__(self, dots=None, voxel=(1.0, 1.0, 1.0)):
        L = _lib.load()
        d = np.zeros((0, 3), np.float32) if dots is None else \
            np.ascontiguousarray(np.asarray(dots, np.float32))
        if d.ndim != 2 or d.shape[1] not in (3, 4):
            raise ValueError("dots must be (n, 3) or (n, 4)")
        self.voxel = tuple(float(v) for v in voxel)
        self._v = (C.c_float * 3)(*self.voxel)
        self._h = C.c_void_p()
        check(L.voxg_edit_create(d.ctypes.data if len(d) else None, len(d),
                                 d.shape[1], self._v, C.byref(self._h)),
              "voxg_edit_create")

    def _op(self, fn, xyz, what):
        p = (C.c_float * 3)(*[float(x) for x in xyz])
        r = C.c_int(0)
        check(fn(self._h, p, C.byref(r)), what)
        return bool(r.value)

    def insert(self, xyz):
        """vox_insert_voxel (tree.c:608-616)."""
        return self._op(_lib.load().voxg_edit_insert, xyz, "voxg_edit_insert")

    def delete(self, xyz):
        """vox_delete_voxel (tree.c:820-828)."""
        return self._op(_lib.load().voxg_edit_delete, xyz, "voxg_edit_delete")

    def rebuild(self):
        """vox_rebuild_tree (tree.c:422-432), in place."""
        check(_lib.load().voxg_edit_rebuild(self._h), "voxg_edit_rebuild")

    @property
    def voxels(self):
        return int(_lib.load().voxg_edit_voxels(self._h))

    def root(self):
        return _lib.load().voxg_edit_root(self._h)

    def plan(self):
        """(info dict, structure hash) of the current state, host only."""
        L = _lib.load()
        info = _lib.TreeInfo()
        h = C.c_uint64(0)
        check(L.voxg_tree_plan_reference(self.root(), self._v, C.byref(info),
                                         C.byref(h)),
              "voxg_tree_plan_reference")
        return ({k: getattr(info, k) for k, _ in info._fields_
                 if k not in ("bbox_min", "bbox_max", "voxel")}, int(h.value))

    def hash(self):
        return self.plan()[1]

    def commit(self):
        """Upload the current state as a new device Tree."""
        h = C.c_void_p()
        check(_lib.load().voxg_edit_commit(self._h, C.byref(h)),
              "voxg_edit_commit")
        return Tree(h, self.voxel)

    def release(self):
        if self._h is not None and self._h.value:
            _lib.load().voxg_edit_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
