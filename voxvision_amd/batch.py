"""Batched form: many trees of uneven size in one persistent launch.

The batch is a device work queue of (tree, ray range) items uploaded once
(voxg_batch_create); voxg_batch_ray then traces every item's rays in one
kernel launch that needs no memset, so it can be captured in a HIP graph
and replayed.  Results equal voxg_ray_batch on each item's own tree.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check
from .voxtrees import _stream


class Batch:
    def __init__(self, items):
        """items: list of (Tree, first_ray, ray_count)."""
        L = _lib.load()
        arr = (_lib.BatchItem * len(items))()
        for i, (t, first, count) in enumerate(items):
            arr[i].tree = t.handle
            arr[i].first = int(first)
            arr[i].count = int(count)
        self._trees = [t for t, _, _ in items]  # keep the trees alive
        self._h = C.c_void_p()
        check(L.voxg_batch_create(len(items), arr, C.byref(self._h)),
              "voxg_batch_create")
        self.n_rays = max(int(f) + int(c) for _, f, c in items)

    def release(self):
        if self._h is not None and self._h.value:
            _lib.load().voxg_batch_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def ray(self, soa_o, soa_d, voxel=None, leaf=None, hit=None, stream=None):
        """soa_o, soa_d: sequences of three device float tensors; outputs
        are device tensors (voxel int32, leaf int32, hit (3, n) float32)."""
        L = _lib.load()
        hp = [None, None, None] if hit is None else [h.data_ptr()
                                                     for h in hit]
        check(L.voxg_batch_ray(
            self._h, *[t.data_ptr() for t in list(soa_o) + list(soa_d)],
            voxel.data_ptr() if voxel is not None else None,
            leaf.data_ptr() if leaf is not None else None, *hp,
            _stream(stream)), "voxg_batch_ray")


def soa(x):
    """(n, 3) host or device array -> three contiguous device tensors."""
    t = torch.as_tensor(np.asarray(x, np.float32) if not torch.is_tensor(x)
                        else x)
    t = t.to("cuda", torch.float32)
    return [t[:, k].contiguous() for k in range(3)]
