"""Python mirror of the reference's light manager, backed by libvoxgpu.

Reference interface (src/voxrnd/lights.h):
  vox_make_light_manager()                   -> Lights()
  vox_destroy_light_manager(lm)              -> Lights.release()
  vox_set_ambient_light(lm, color)           -> Lights.set_ambient(color)
  vox_insert_shadowless_light(lm, c, r, col) -> Lights.insert(c, r, col)
  vox_delete_shadowless_light(lm, c, r)      -> Lights.delete(c, r)
  vox_delete_shadowless_lights(lm)           -> Lights.clear()
  vox_shadowless_lights_number(lm)           -> len(Lights)
  vox_get_light(lm, point, light)            -> Lights.light(points)
The lights are kept in the reference's M-tree, so a render that uses them
(voxrnd.render(..., lights=L), vox_context_set_light_manager) sums the
lights containing a hit point in the reference's order.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


class Lights:
    def __init__(self):
        self._h = C.c_void_p()
        check(_lib.load().voxg_lights_create(C.byref(self._h)),
              "voxg_lights_create")

    @property
    def handle(self):
        if not self._h:
            raise ValueError("light manager was released")
        return self._h

    def set_ambient(self, color):
        """1 when every component is in [0, 1] and the light was set."""
        ok = C.c_int()
        check(_lib.load().voxg_lights_set_ambient(self.handle, _f3(color),
                                                  C.byref(ok)),
              "voxg_lights_set_ambient")
        return ok.value

    def insert(self, center, radius, color):
        """1 when the light was added, 0 when refused as the reference
        refuses it; raises for a centre already present."""
        ok = C.c_int()
        check(_lib.load().voxg_lights_insert(self.handle, _f3(center),
                                             float(radius), _f3(color),
                                             C.byref(ok)),
              "voxg_lights_insert")
        return ok.value

    def delete(self, center, radius):
        ok = C.c_int()
        check(_lib.load().voxg_lights_delete(self.handle, _f3(center),
                                             float(radius), C.byref(ok)),
              "voxg_lights_delete")
        return ok.value

    def clear(self):
        check(_lib.load().voxg_lights_clear(self.handle), "voxg_lights_clear")

    def __len__(self):
        return int(_lib.load().voxg_lights_count(self.handle))

    def light(self, points):
        """vox_get_light at (n, 3) points -> (n, 3) float32 (host)."""
        p = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
        out = np.zeros_like(p)
        check(_lib.load().voxg_lights_get(self.handle, len(p),
                                          p.ctypes.data if len(p) else None,
                                          out.ctypes.data if len(p) else None),
              "voxg_lights_get")
        return out

    def release(self):
        if self._h:
            _lib.load().voxg_lights_release(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
