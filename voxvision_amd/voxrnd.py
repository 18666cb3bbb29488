"""Python mirror of the vox_render per-pixel loop, backed by libvoxgpu.

Reference: vox_render (src/voxrnd/renderer.c:200-319) over 4x4-pixel
squares (renderer.h:14) with the simple camera's screen2world
(src/voxrnd/simple-camera.c:21-31).  Cameras are data (see camera.py);
colours use a caller-supplied 32^3 texture (the reference generates it
with vn3d) and a 32-bit pixel format (SDL_MapRGB).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, QUALITY_BEST
from .camera import camera_array
from .voxtrees import _stream

ARGB8888 = (16, 8, 0, 0xFF000000)


def cameras_tensor(cams):
    """List of camera dicts -> device tensor in the voxg_camera layout."""
    return torch.from_numpy(camera_array(cams)).to("cuda")


def _params(quality, texture, fmt, lights=None):
    p = _lib.RenderParams()
    p.quality = quality
    p.texture = texture.data_ptr() if texture is not None else None
    p.rshift, p.gshift, p.bshift, p.amask = fmt
    p.lights = lights.handle if lights is not None else None
    return p


def render(tree, cams, W, H, quality=QUALITY_BEST, texture=None,
           fmt=ARGB8888, want_colors=True, want_voxels=True, want_hits=True,
           out=None, stream=None, lights=None):
    """Render len(cams) frames (voxg_render).  Returns dict of device
    tensors in the square layout (index f*W*H + cs*16 + p): squares (uint32
    colours, 0 where nothing was hit), voxel (int32, -1 = miss), hit
    (3, f*W*H; NaN where nothing was hit when allocated here).

    Every quality gives the reference's results.  BEST traces one pixel per
    work unit (launch kind "render_best"); FAST and ADAPTIVE (with or
    without ray merging) root-search the pixels the same way, then replay
    each 4x4 block in the reference's order (leaf reuse, corners, merging),
    using a per-stream scratch buffer of 20 bytes per pixel.  With plain
    ray merging the odd pixels of a block, copies while it merges, are
    traced only for the blocks whose merging stops.  Counting launches
    trace block by block (launch kind "render").  `lights` (a
    voxvision_amd.lights.Lights, vox_context_set_light_manager) scales the
    colours by the light at each hit point."""
    L = _lib.load()
    if not torch.is_tensor(cams):
        cams = cameras_tensor(cams)
    nf = cams.shape[0]
    n = nf * W * H
    if out is None:
        out = {}
        if want_colors:
            out["squares"] = torch.zeros(n, dtype=torch.int32, device="cuda")
        if want_voxels:
            out["voxel"] = torch.empty(n, dtype=torch.int32, device="cuda")
        if want_hits:
            out["hit"] = torch.full((3, n), float("nan"), dtype=torch.float32,
                                    device="cuda")
    tex = None
    if texture is not None:
        tex = torch.as_tensor(texture, dtype=torch.uint8).to("cuda")
    p = _params(quality, tex, fmt, lights)
    sq = out.get("squares")
    vx = out.get("voxel")
    hv = out.get("hit")
    check(L.voxg_render(tree.handle, nf, cams.data_ptr(), W, H, C.byref(p),
                        sq.data_ptr() if sq is not None else None,
                        vx.data_ptr() if vx is not None else None,
                        hv.data_ptr() if hv is not None else None,
                        _stream(stream)), "voxg_render")
    return out


def render_counts(tree, cams, W, H, quality=QUALITY_BEST, stream=None):
    L = _lib.load()
    if not torch.is_tensor(cams):
        cams = cameras_tensor(cams)
    p = _params(quality, None, ARGB8888)
    c = _lib.Counts()
    check(L.voxg_render_count(tree.handle, cams.shape[0], cams.data_ptr(), W,
                              H, C.byref(p), C.byref(c), _stream(stream)),
          "voxg_render_count")
    return {k: getattr(c, k) for k, _ in c._fields_}


def redraw(squares, W, H, pixels=None, clear=True, stream=None):
    """vox_redraw (renderer.c:321-327): copy the square-layout frames of
    `squares` (device uint32/int32, n_frames*W*H) to row-major pixels
    (returned as an (n_frames, H, W) device tensor) and, when clear, zero
    the squares for the next frame."""
    L = _lib.load()
    n = squares.numel()
    nf = n // (W * H)
    if pixels is None:
        pixels = torch.empty((nf, H, W), dtype=squares.dtype, device="cuda")
    check(L.voxg_redraw(squares.data_ptr(), pixels.data_ptr(), nf, W, H,
                        1 if clear else 0, _stream(stream)), "voxg_redraw")
    return pixels


def cd_collide(tree, pos, last, radius, collided=None, stream=None):
    """vox_cd_collide (cd.c:38-55) for many bodies: pos and last are lists
    of three device float tensors (x, y, z), updated in place; radius a
    device float tensor.  Returns the device uint8 tensor of bodies moved
    back to their last position."""
    L = _lib.load()
    n = radius.numel()
    if collided is None:
        collided = torch.empty(n, dtype=torch.uint8, device="cuda")
    check(L.voxg_cd_collide_batch(
        tree.handle, n, *[t.data_ptr() for t in list(pos) + list(last)],
        radius.data_ptr(), collided.data_ptr(), _stream(stream)),
        "voxg_cd_collide_batch")
    return collided


def squares_to_image(squares, W, H):
    """Host helper: square layout -> (H, W) image (copy_squares,
    src/voxrnd/copy-helper.c:5-40)."""
    s = np.asarray(squares).reshape(H // 4, W // 4, 4, 4)
    return s.transpose(0, 2, 1, 3).reshape(H, W)
