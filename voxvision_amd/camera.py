"""Host-side simple camera (src/voxrnd/simple-camera.c).

The GPU renderer takes the camera as data (position, rotation quaternion,
mul/xsub/ysub), because the reference's cameras are host vtables loaded
with dlopen (src/voxrnd/modules.c:43-69) and cannot run on the device.
"""
import math

import numpy as np


def set_window_size(fov, w, h):
    """simple-camera.c:120-136 in fp32: returns (mul, xsub, ysub)."""
    f = np.float32(fov)
    w32, h32 = np.float32(w), np.float32(h)
    if w > h:
        xsub = f
        ysub = (f * h32) / w32
        mul = (f * np.float32(2.0)) / w32
    else:
        ysub = f
        xsub = (f * w32) / h32
        mul = (f * np.float32(2.0)) / h32
    return float(mul), float(xsub), float(ysub)


def _qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return (w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
            w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
            w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
            w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2)


def look_at_quaternion(position, target):
    """Rotation taking the camera's forward axis (+y) to target - position
    with +z up, as simple_look_at (simple-camera.c:106-118) composes it:
    yaw -atan2(x, y) about z, then pitch atan2(z, hypot(x, y)) about the
    camera's x axis.  Computed in float64 and normalised exactly (the
    reference normalises with the approximate _mm_rsqrt_ps), then rounded
    to fp32; the quaternion is handed to the renderer as data."""
    sub = [float(target[k]) - float(position[k]) for k in range(3)]
    yaw = -math.atan2(sub[0], sub[1])
    pitch = math.atan2(sub[2], math.hypot(sub[0], sub[1]))
    qy = (math.cos(yaw / 2), 0.0, 0.0, math.sin(yaw / 2))
    qp = (math.cos(pitch / 2), math.sin(pitch / 2), 0.0, 0.0)
    q = _qmul(qy, qp)
    n = math.sqrt(sum(c * c for c in q))
    return tuple(float(np.float32(c / n)) for c in q)


def make_camera(position, target, fov, W, H):
    mul, xsub, ysub = set_window_size(fov, W, H)
    return dict(position=tuple(float(np.float32(p)) for p in position),
                rotation=look_at_quaternion(position, target),
                fov=float(np.float32(fov)), mul=mul, xsub=xsub, ysub=ysub)


def camera_array(cams):
    """Pack camera dicts into the voxg_camera layout: (n, 11) float32 rows
    position[3], rotation[4], mul, xsub, ysub, fov."""
    out = np.zeros((len(cams), 11), np.float32)
    for i, c in enumerate(cams):
        out[i, 0:3] = c["position"]
        out[i, 3:7] = c["rotation"]
        out[i, 7] = c["mul"]
        out[i, 8] = c["xsub"]
        out[i, 9] = c["ysub"]
        out[i, 10] = c.get("fov", 1.0)
    return out
