"""ctypes binding of libvoxgpu.so (the C ABI in include/voxgpu.h).

There is no fallback: if the library is missing or cannot be loaded this
module raises, so nothing can silently run another implementation.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libvoxgpu.so")

VOXG_MISS = 0xFFFFFFFF
QUALITY_BEST = 0x0
QUALITY_FAST = 0x1
QUALITY_ADAPTIVE = 0x2
QUALITY_RAY_MERGE = 0x4
QUALITY_RAY_MERGE_ACCURATE = 0x8

# Every symbol include/voxgpu.h and include/voxgpu_testing.h declare.
EXPORTS = [
    "voxg_init", "voxg_tree_build", "voxg_tree_upload", "voxg_tree_plan",
    "voxg_tree_plan_reference", "voxg_tree_release",
    "voxg_tree_get_info", "voxg_tree_hash", "voxg_tree_voxel_coords",
    "voxg_tree_node_ptr", "voxg_ray_batch", "voxg_ball_batch",
    "voxg_camera_set_window_size", "voxg_render", "voxg_ray_batch_count",
    "voxg_ball_batch_count", "voxg_render_count", "voxg_set_launch_config",
    "voxg_get_launch_config", "voxg_set_kind_launch_config",
    "voxg_get_kind_launch_config", "voxg_last_error", "voxg_batch_create",
    "voxg_batch_ray", "voxg_batch_release", "voxg_server_create",
    "voxg_server_add_tree", "voxg_server_submit_render",
    "voxg_server_submit_balls", "voxg_server_wait", "voxg_server_get_stats",
    "voxg_server_now", "voxg_server_destroy", "voxg_cd_collide_batch",
    "voxg_redraw", "voxg_raw_select", "voxg_raw_workspace_bytes",
    "voxg_raw_count", "voxg_raw_write", "voxg_read_raw_data",
    "voxg_edit_create", "voxg_edit_insert", "voxg_edit_delete",
    "voxg_edit_rebuild", "voxg_edit_root", "voxg_edit_voxels",
    "voxg_edit_commit", "voxg_edit_release", "voxg_server_replace_tree",
    "voxg_synchronize", "voxg_release_stream", "voxg_test_fail_at",
    "voxg_test_stack_bound", "voxg_test_server_hold", "voxg_test_plan_cut",
    "voxg_ray_batch_noalloc",
    "voxg_lights_create", "voxg_lights_release", "voxg_lights_set_ambient",
    "voxg_lights_insert", "voxg_lights_delete", "voxg_lights_clear",
    "voxg_lights_count", "voxg_lights_get",
]


class VoxgError(RuntimeError):
    pass


class TreeInfo(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("inner", C.c_uint64),
                ("leaves", C.c_uint64), ("dense", C.c_uint64),
                ("overflow", C.c_uint64), ("dots", C.c_uint64),
                ("voxel_ids", C.c_uint64), ("depth", C.c_uint32),
                ("stack_entries", C.c_uint32), ("device_bytes", C.c_uint64),
                ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
                ("voxel", C.c_float * 3)]


class Counts(C.Structure):
    _fields_ = [("node_visits", C.c_uint64), ("voxel_tests", C.c_uint64),
                ("rays", C.c_uint64), ("leaf_traversals", C.c_uint64),
                ("fast_blocks", C.c_uint64), ("best_blocks", C.c_uint64),
                ("merge_blocks", C.c_uint64), ("stack_spills", C.c_uint64)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("rotation", C.c_float * 4),
                ("mul", C.c_float), ("xsub", C.c_float), ("ysub", C.c_float),
                ("fov", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("quality", C.c_uint), ("texture", C.c_void_p),
                ("rshift", C.c_uint32), ("gshift", C.c_uint32),
                ("bshift", C.c_uint32), ("amask", C.c_uint32),
                ("lights", C.c_void_p)]


class BatchItem(C.Structure):
    _fields_ = [("tree", C.c_void_p), ("first", C.c_size_t),
                ("count", C.c_size_t)]


class LaunchConfig(C.Structure):
    _fields_ = [("stack_lds", C.c_int), ("block", C.c_int),
                ("waves_per_cu", C.c_int), ("refill", C.c_int),
                ("chunk", C.c_int), ("render_cull", C.c_int),
                ("leaf_batch", C.c_int), ("finish_batch", C.c_int),
                ("split_best", C.c_int)]


_lib = None


def load():
    """Load libvoxgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libvoxgpu.so is not built (%s); run "
            "`python -m voxvision_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    I = C.c_int
    S = C.c_size_t
    L.voxg_init.argtypes = [I]
    L.voxg_tree_build.argtypes = [P, S, S, P, C.POINTER(P)]
    L.voxg_tree_upload.argtypes = [P, P, C.POINTER(P)]
    L.voxg_tree_plan.argtypes = [P, S, S, P, C.POINTER(TreeInfo),
                                 C.POINTER(C.c_uint64)]
    L.voxg_tree_plan_reference.argtypes = [P, P, C.POINTER(TreeInfo),
                                           C.POINTER(C.c_uint64)]
    L.voxg_tree_release.argtypes = [P]
    L.voxg_tree_get_info.argtypes = [P, C.POINTER(TreeInfo)]
    L.voxg_tree_hash.argtypes = [P]
    L.voxg_tree_hash.restype = C.c_uint64
    L.voxg_tree_voxel_coords.argtypes = [P, S, P, P]
    L.voxg_tree_node_ptr.argtypes = [P, C.c_uint32]
    L.voxg_tree_node_ptr.restype = P
    L.voxg_ray_batch.argtypes = [P, S] + [P] * 6 + [P] * 5 + [P]
    L.voxg_ray_batch_noalloc.argtypes = [P, S] + [P] * 6 + [P] * 5 + [P]
    L.voxg_ray_batch_count.argtypes = [P, S] + [P] * 6 + [
        C.POINTER(Counts), P]
    L.voxg_ball_batch.argtypes = [P, S, P, P, P, P, P, P]
    L.voxg_ball_batch_count.argtypes = [P, S, P, P, P, P,
                                        C.POINTER(Counts), P]
    L.voxg_camera_set_window_size.argtypes = [C.POINTER(Camera), I, I]
    L.voxg_camera_set_window_size.restype = None
    L.voxg_render.argtypes = [P, S, P, I, I, C.POINTER(RenderParams), P, P,
                              P, P]
    L.voxg_render_count.argtypes = [P, S, P, I, I, C.POINTER(RenderParams),
                                    C.POINTER(Counts), P]
    L.voxg_set_launch_config.argtypes = [C.POINTER(LaunchConfig)]
    L.voxg_get_launch_config.argtypes = [C.POINTER(LaunchConfig)]
    L.voxg_set_kind_launch_config.argtypes = [C.c_int,
                                              C.POINTER(LaunchConfig)]
    L.voxg_get_kind_launch_config.argtypes = [C.c_int,
                                              C.POINTER(LaunchConfig)]
    L.voxg_last_error.restype = C.c_char_p
    L.voxg_batch_create.argtypes = [S, C.POINTER(BatchItem), C.POINTER(P)]
    L.voxg_batch_ray.argtypes = [P] + [P] * 6 + [P] * 5 + [P]
    L.voxg_batch_release.argtypes = [P]
    L.voxg_cd_collide_batch.argtypes = [P, S] + [P] * 8 + [P]
    L.voxg_redraw.argtypes = [P, P, S, I, I, I, P]
    U64 = C.c_uint64
    L.voxg_raw_select.argtypes = [P, P, C.c_uint, U64, U64, P, P, S, P, P]
    L.voxg_raw_workspace_bytes.argtypes = [P]
    L.voxg_raw_workspace_bytes.restype = S
    L.voxg_raw_count.argtypes = [P, P, C.c_uint, U64, U64, P, P, P]
    L.voxg_raw_write.argtypes = [P, P, P, P, S, P]
    L.voxg_read_raw_data.argtypes = [C.c_char_p, P, C.c_uint, U64, U64, P,
                                     C.POINTER(P)]
    L.voxg_edit_create.argtypes = [P, S, S, P, C.POINTER(P)]
    L.voxg_edit_insert.argtypes = [P, P, C.POINTER(C.c_int)]
    L.voxg_edit_delete.argtypes = [P, P, C.POINTER(C.c_int)]
    L.voxg_edit_rebuild.argtypes = [P]
    L.voxg_edit_root.argtypes = [P]
    L.voxg_edit_root.restype = P
    L.voxg_edit_voxels.argtypes = [P]
    L.voxg_edit_voxels.restype = S
    L.voxg_edit_commit.argtypes = [P, C.POINTER(P)]
    L.voxg_edit_release.argtypes = [P]
    L.voxg_synchronize.argtypes = [P]
    L.voxg_release_stream.argtypes = [P]
    L.voxg_lights_create.argtypes = [C.POINTER(P)]
    L.voxg_lights_release.argtypes = [P]
    L.voxg_lights_set_ambient.argtypes = [P, P, C.POINTER(C.c_int)]
    L.voxg_lights_insert.argtypes = [P, P, C.c_float, P, C.POINTER(C.c_int)]
    L.voxg_lights_delete.argtypes = [P, P, C.c_float, C.POINTER(C.c_int)]
    L.voxg_lights_clear.argtypes = [P]
    L.voxg_lights_count.argtypes = [P]
    L.voxg_lights_count.restype = S
    L.voxg_lights_get.argtypes = [P, S, P, P]
    L.voxg_test_fail_at.argtypes = [I]
    L.voxg_test_stack_bound.argtypes = [C.c_uint]
    L.voxg_test_server_hold.argtypes = [P, C.c_int]
    L.voxg_test_plan_cut.argtypes = [P, S, S, P, P, S, C.POINTER(S)]
    _lib = L
    return L


def check(rc, what="voxg call"):
    if rc != 0:
        msg = load().voxg_last_error()
        raise VoxgError("%s failed (%d): %s" %
                        (what, rc, msg.decode() if msg else ""))
    return rc
