"""libvoxgpu host-side checks that need no GPU: the C ABI library loads and
exports every symbol include/voxgpu.h declares, the reference tree layout
matches src/voxtrees/tree.h:28-38, and the product's tree builder and
flattener reproduce the oracle's vox_make_tree bit for bit."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(voxg_[a-z_]+)\s*\(", text))
    return sorted(names)


def test_exports_every_declared_symbol(voxlib):
    from voxvision_amd import _lib
    names = declared_functions()
    assert len(names) >= 15
    assert sorted(_lib.EXPORTS) == names
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(voxlib, n)


def test_reference_layout_offsets(tmp_path):
    """struct vox_node as declared in include/voxgpu.h must have the layout
    of the reference's SSE build (vox_dot = float[4], 16-byte aligned)."""
    src = tmp_path / "l.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "voxgpu.h"
#include "vox_oracle.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(struct vox_node),
    offsetof(struct vox_node, flags), offsetof(struct vox_node, dots_num),
    offsetof(struct vox_node, data),
    offsetof(struct vox_node, data.inner.center),
    offsetof(struct vox_node, data.inner.children));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(struct orc_node),
    offsetof(struct orc_node, flags), offsetof(struct orc_node, dots_num),
    offsetof(struct orc_node, data),
    offsetof(struct orc_node, data.inner.center),
    offsetof(struct orc_node, data.inner.children));
  return 0;
}''')
    exe = tmp_path / "l"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "oracle"), str(src), "-o", str(exe)],
                   check=True)
    lines = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True,
                           check=True).stdout.split("\n")
    # box 32 B, two uints, the union 16-byte aligned, 8 child pointers
    assert lines[0].split() == ["128", "32", "36", "48", "48", "64"]
    assert lines[0] == lines[1]


def _scenes():
    from voxvision_amd import scenes
    rng = np.random.default_rng(21)
    ii = np.arange(-50, 50)
    I, J, K = np.meshgrid(ii, ii, ii, indexing="ij")
    m = (I * I + J * J + K * K) < 2500
    yield "ball50", np.stack([I[m], J[m], K[m]], 1).astype(np.float32), \
        (1.0, 1.0, 1.0)
    yield "c1", scenes.config_c1()["dots"], (1.0, 1.0, 1.0)
    yield "random", scenes.volume_to_dots(rng.random((30, 20, 25)) < 0.2), \
        (1.0, 1.0, 1.0)
    V = (0.5, 2.0, 1.25)
    yield "nonunit", scenes.volume_to_dots(rng.random((20, 20, 20)) < 0.3,
                                           V), V
    # duplicates -> overflow leaves; off-grid floats -> overlapping boxes
    dup = np.concatenate([np.zeros((12, 3), np.float32) + 3,
                          rng.integers(0, 9, (200, 3)).astype(np.float32)])
    yield "duplicates", dup, (1.0, 1.0, 1.0)
    yield "offgrid", (rng.random((3000, 3)) * 40 - 7).astype(np.float32), \
        (1.0, 1.0, 1.0)


@pytest.mark.parametrize("case", list(range(6)))
def test_builder_matches_oracle_tree(oracle, voxlib, case):
    from voxvision_amd import voxtrees as VT
    name, dots, V = list(_scenes())[case]
    T = oracle.Tree(dots, V)
    info, h = VT.plan(dots, V)
    assert h == T.hash(), name
    for k in ("nodes", "inner", "leaves", "dense", "overflow", "dots",
              "voxel_ids", "depth"):
        assert info[k] == T.stats[k], (name, k)
    # flattening the oracle's reference-layout tree gives the same tree
    info2, h2 = VT.plan_reference(T.ptr, V)
    assert h2 == h and info2 == info, name
    oracle.set_voxel((1.0, 1.0, 1.0))


def test_plan_rejects_bad_arguments(voxlib):
    import ctypes as C
    from voxvision_amd import _lib
    v = (C.c_float * 3)(1, 1, 1)
    assert voxlib.voxg_tree_plan(None, 5, 3, v, None, None) == -22
    d = np.zeros((4, 3), np.float32)
    assert voxlib.voxg_tree_plan(d.ctypes.data, 4, 2, v, None, None) == -22
    bad = (C.c_float * 3)(1, 0, 1)
    assert voxlib.voxg_tree_plan(d.ctypes.data, 4, 3, bad, None, None) == -22
    assert b"voxel" in voxlib.voxg_last_error()
    # empty set is a valid (empty) tree
    info = _lib.TreeInfo()
    assert voxlib.voxg_tree_plan(None, 0, 3, v, C.byref(info), None) == 0
    assert info.nodes == 0


def test_volume_sizes_that_would_wrap_are_refused(voxlib):
    """voxg_raw_select and voxg_read_raw_data refuse volumes of 2^32
    samples or more before touching any memory, file or GPU, also where
    dim[0] * dim[1] * dim[2] wraps around 2^64 (2^21 * 2^21 * 2^22 would
    wrap to an empty volume)."""
    cnt = C.c_ulonglong(0)
    v = (C.c_float * 3)(1, 1, 1)
    for dims in ((1 << 21, 1 << 21, 1 << 22), (2, 1 << 31, 1),
                 (1 << 16, 1 << 16, 1), (4294967295, 4294967295, 2)):
        d = (C.c_uint * 3)(*dims)
        assert voxlib.voxg_raw_select(None, d, 1, 0, 255, v, None, 0,
                                      C.byref(cnt), None) == -34, dims
        out = C.c_void_p()
        assert voxlib.voxg_read_raw_data(b"/nonexistent/volume.raw", d, 1,
                                         0, 255, v, C.byref(out)) == -34
        assert not out.value


def test_edit_create_rejects_bad_voxel_sizes(voxlib):
    """Like voxg_tree_build, voxg_edit_create takes only positive voxel
    sizes (a zero or NaN size has no grid to align points to)."""
    pts = np.zeros((2, 3), np.float32)
    pts[1] = 3
    for bad in ((1, 0, 1), (1, 1, -2), (float("nan"), 1, 1)):
        v = (C.c_float * 3)(*bad)
        out = C.c_void_p()
        assert voxlib.voxg_edit_create(pts.ctypes.data, 2, 3, v,
                                       C.byref(out)) == -22
        assert not out.value
    out = C.c_void_p()
    assert voxlib.voxg_edit_create(pts.ctypes.data, 2, 3,
                                   (C.c_float * 3)(1, 2, 1),
                                   C.byref(out)) == 0
    assert voxlib.voxg_edit_voxels(out) == 2
    voxlib.voxg_edit_release(out)


def test_camera_setter_matches_reference_formula(voxlib):
    from voxvision_amd import _lib, camera
    for (w, h) in ((640, 480), (1920, 1080), (100, 100), (240, 320)):
        c = _lib.Camera()
        c.fov = np.float32(np.tan(np.radians(30)))
        voxlib.voxg_camera_set_window_size(c, w, h)
        mul, xs, ys = camera.set_window_size(c.fov, w, h)
        assert (c.mul, c.xsub, c.ysub) == (mul, xs, ys)


# ---- editable trees (host code: no device needed) ----
def _edit_pair(O, dots=None):
    from voxvision_amd.edit import EditTree
    if dots is None:
        return EditTree(), O.Tree.empty()
    return EditTree(dots), O.Tree(dots)


def _run_script(E, T, script):
    for op, p in script:
        a = E.insert(p) if op == "ins" else E.delete(p)
        b = T.insert(p) if op == "ins" else T.delete(p)
        assert a == bool(b), (op, p)
        assert E.voxels == T.voxels
        assert E.hash() == T.hash(), (op, p)


def test_edit_reference_scripts_match_oracle(oracle):
    """The insertion/deletion sequences of tests.c:367-659 give the same
    tree, node for node, as the oracle after every step."""
    scripts = []
    s = [("ins", (0, 0, 0)), ("ins", (0.5, 0.1, 0.8)), ("ins", (6, 6, 6))]
    scripts.append(s)
    for step in (2, 1):
        scripts.append([("ins", (step * i, step * j, step * k))
                        for i in range(3) for j in range(3) for k in range(3)])
    for axis, start, d in ((1, 0, 1), (2, 16, -1), (0, 1, 1)):
        s = []
        p = [0.0, 0.0, 0.0]
        for _ in range(17):
            s.append(("ins", tuple(p)))
            p[axis] += 1
        p = [0.0, 0.0, 0.0]
        p[axis] = start
        for _ in range(10):
            s.append(("del", tuple(p)))
            p[axis] += d
        scripts.append(s)
    cube = [("ins", (i, j, k)) for i in range(3) for j in range(3)
            for k in range(3)]
    scripts.append(cube + [("del", (0, 0, 0))])
    scripts.append(cube + [("del", (1, 1, 1))])
    scripts.append([("ins", (a, 0, 0)) for a in range(1, 12)] +
                   [("del", (a, 0, 0)) for a in range(1, 12)])
    cross = [(0, 0, 0), (0, 0, -1), (0, 0, -2), (0, 0, 1), (0, 0, 2)]
    cross += [(-a, 0, 0) for a in (1, 2, 3, 4)] + [(a, 0, 0) for a in (1, 2, 3, 4)]
    cross += [(0, -a, 0) for a in (1, 2, 3, 4)] + [(0, a, 0) for a in (1, 2, 3, 4)]
    scripts.append([("ins", p) for p in cross])
    for s in scripts:
        E, T = _edit_pair(oracle)
        _run_script(E, T, s)
    # a dense-leaf root (tests.c:390-402), built as the 5x5x5 block
    blk = np.array([(i, j, k) for i in range(5, 10) for j in range(5, 10)
                    for k in range(5, 10)], np.float32)
    E, T = _edit_pair(oracle, blk)
    assert E.hash() == T.hash()
    _run_script(E, T, [("ins", (6, 6, 6)), ("ins", (0, 0, 0))])


@pytest.mark.parametrize("scene", ["sparse", "dense_cube", "blobs",
                                   "anisotropic", "tiny_voxel"])
def test_edit_random_sequences_match_oracle(oracle, scene):
    """Random inserts and deletes (misses, hits, aligned and unaligned
    points, negative coordinates, non-unit voxels: 0.5, (1, 0.5, 3) and
    2^-10) on sparse, dense and blob scenes: product and oracle trees are
    identical after every edit, and vox_rebuild_tree agrees too."""
    from voxvision_amd import scenes
    from voxvision_amd.edit import EditTree
    O = oracle
    rng = np.random.default_rng({"sparse": 1, "dense_cube": 2, "blobs": 3,
                                 "anisotropic": 4, "tiny_voxel": 5}[scene])
    V = (1.0, 1.0, 1.0)
    if scene in ("anisotropic", "tiny_voxel"):
        V = (1.0, 0.5, 3.0) if scene == "anisotropic" else (2.0 ** -10,) * 3
        idx = np.unique(rng.integers(-6, 10, (300, 3)), axis=0)
        dots = (idx.astype(np.float32) * np.asarray(V, np.float32))
        lo = -8 * np.asarray(V)
        hi = 12 * np.asarray(V)
        n_ops = 500
    elif scene == "sparse":
        dots = np.unique(rng.integers(0, 16, (400, 3)), axis=0).astype(np.float32)
        lo, hi, n_ops = -2, 18, 800
    elif scene == "dense_cube":
        dots = np.array([(i, j, k) for i in range(10) for j in range(10)
                         for k in range(10)], np.float32)
        lo, hi, n_ops = -1, 11, 600
    else:
        occ = scenes.blobs64(2)[:32, :32, :32]
        dots = np.argwhere(occ).astype(np.float32) * 0.5
        V = (0.5, 0.5, 0.5)
        lo, hi, n_ops = 0, 16, 300
    E = EditTree(dots, voxel=V)
    T = O.Tree(dots, voxel=V)
    assert E.hash() == T.hash()
    for _ in range(n_ops):
        p = rng.uniform(lo, hi, 3)
        if rng.random() < 0.3:
            p = np.floor(p / np.asarray(V)) * np.asarray(V)
        op = "ins" if rng.random() < 0.5 else "del"
        a = E.insert(p) if op == "ins" else E.delete(p)
        b = T.insert(p) if op == "ins" else T.delete(p)
        assert a == bool(b)
        assert E.hash() == T.hash(), (op, p)
    assert E.voxels == T.voxels and T.check() == 0
    R = T.rebuild()
    E.rebuild()
    assert E.hash() == R.hash()


def test_edit_with_an_inexact_voxel_size_stops_at_a_bad_dense_leaf(oracle):
    """With a voxel size that is not exact in binary (0.001), a dense leaf's
    box divided by it can truncate one short of its voxel count
    (get_dimensions, geom.c:184-196); listing that leaf's voxels is then
    undefined and the reference stops on an assertion (tree.c:389).  The
    edit that meets such a leaf fails with VOXG_ERANGE and the tree refuses
    every later edit, rebuild and commit; until then every step equals the
    oracle's."""
    from voxvision_amd import _lib
    from voxvision_amd.edit import EditTree
    rng = np.random.default_rng(5)
    V = (1e-3,) * 3
    idx = np.unique(rng.integers(-6, 10, (300, 3)), axis=0)
    dots = idx.astype(np.float32) * np.asarray(V, np.float32)
    E = EditTree(dots, voxel=V)
    T = oracle.Tree(dots, voxel=V)
    lo, hi = -8 * np.asarray(V), 12 * np.asarray(V)
    stopped = None
    for step in range(500):
        p = rng.uniform(lo, hi, 3)
        if rng.random() < 0.3:
            p = np.floor(p / np.asarray(V)) * np.asarray(V)
        ins = rng.random() < 0.5
        try:
            a = E.insert(p) if ins else E.delete(p)
        except _lib.VoxgError as e:
            assert "dense leaf" in str(e)
            stopped = step
            break
        b = T.insert(p) if ins else T.delete(p)
        assert a == bool(b)
        assert E.hash() == T.hash(), step
    assert stopped is not None
    for call in (lambda: E.insert((0.0, 0.0, 0.0)),
                 lambda: E.delete((0.0, 0.0, 0.0)), E.rebuild):
        with pytest.raises(_lib.VoxgError):
            call()


@pytest.mark.parametrize("n", [2, 3])
def test_edit_delete_off_the_dense_grid_is_refused(n):
    """Dots that are not multiples of the voxel size (a cube of n^3 at
    offset 0.5 makes one dense leaf): deleting an aligned point inside its
    box, which lies on none of its voxels, is refused with VOXG_EINVAL and
    leaves the tree unchanged (the reference asserts, tree.c:808, or
    divides the box off its grid).  Deleting one of its own voxels works."""
    from voxvision_amd import _lib
    from voxvision_amd.edit import EditTree
    g = np.arange(n, dtype=np.float32) + np.float32(0.5)
    dots = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    E = EditTree(dots)
    before = E.hash()
    with pytest.raises(_lib.VoxgError, match="off its voxel grid"):
        E.delete((1.2, 1.2, 1.2))
    assert E.hash() == before and E.voxels == n ** 3
    E2 = EditTree(dots)
    assert not E2.delete((0.5, 0.5, 0.5))   # aligns to (0, 0, 0): outside
    assert E2.voxels == n ** 3


def _edit_fuzz_case(seed):
    """A random scene and 1500 random edits with a rebuild every 300: voxel
    sizes exact and inexact in binary, dense blocks, planes, offsets up to
    -6e4 (where 2^-10 is below the coordinates' precision)."""
    rng = np.random.default_rng(seed)
    choices = [(1.0,) * 3, (0.5,) * 3, (1.0, 0.5, 3.0), (1e-3,) * 3,
               (0.1, 0.3, 0.7), (2.0 ** -10,) * 3]
    V = np.asarray(choices[seed % 6], np.float32)
    kind = (seed // 6) % 4
    n = int(rng.integers(1, 400))
    if kind == 1:
        a = rng.integers(0, 4, 3)
        b = a + rng.integers(1, 6, 3)
        idx = np.stack(np.meshgrid(*[np.arange(a[i], b[i]) for i in range(3)],
                                   indexing="ij"), -1).reshape(-1, 3)
    else:
        r = (-3, 6) if kind == 3 else (-5, 12)
        idx = np.unique(rng.integers(r[0], r[1], (n, 3)), axis=0)
        if kind == 2:
            idx[:, rng.integers(0, 3)] = 2
    off = (rng.choice([0.0, 0.5, 0.25, 1000.0, -6e4], 3)
           if rng.random() < 0.5 else np.zeros(3))
    dots = (idx.astype(np.float32) * V + off.astype(np.float32))
    lo, hi = dots.min(0) - 3 * V, dots.max(0) + 4 * V
    ops = []
    for i in range(1500):
        if rng.random() < 0.5:
            p = dots[rng.integers(0, len(dots))] + V * rng.integers(-2, 3, 3)
        else:
            p = rng.uniform(lo, hi)
        ops.append(("ins" if rng.random() < 0.5 else "del",
                    p.astype(np.float32)))
        if i % 300 == 299:
            ops.append(("rebuild", None))
    return tuple(float(x) for x in V), dots.astype(np.float32), ops


@pytest.mark.parametrize("seed", [11, 35, 425, 473, 0, 7, 100])
def test_edit_sequences_never_corrupt_memory(seed):
    """Edit sequences that once overran a flattening buffer (a dense leaf
    whose truncated dimensions multiply to more than its count: seeds 11,
    35) or recursed without end (a new voxel that does not separate from a
    dense leaf: 425, 473) now end in VOXG_ERANGE, after which the tree
    refuses every call; the others run through.  A crash here is the
    failure."""
    from voxvision_amd import _lib
    from voxvision_amd.edit import EditTree
    V, dots, ops = _edit_fuzz_case(seed)
    E = EditTree(dots, voxel=V)
    failed = False
    for op, p in ops:
        try:
            if op == "ins":
                E.insert(p)
            elif op == "del":
                E.delete(p)
            else:
                E.rebuild()
        except _lib.VoxgError as e:
            if "off its voxel grid" in str(e):
                continue   # refused, tree unchanged
            failed = True
            with pytest.raises(_lib.VoxgError):
                E.insert(dots[0])
            break
    if seed in (11, 35, 425, 473):
        assert failed


def test_launch_configs_per_kind():
    """Each kind of launch keeps a configuration of its own; setting all
    kinds at once, one kind, and the restore context manager; invalid
    kinds and fields are refused.  Host state only, no GPU needed."""
    from voxvision_amd import _lib, voxtrees as VT
    with VT.saved_launch_configs() as saved:
        assert set(saved) == {"rays", "render", "render_best", "batch"}
        # BEST renders default to one pixel per unit, the others to blocks
        assert saved["render_best"]["split_best"] == 4
        VT.set_launch_config("render", refill=5)
        assert VT.get_launch_config("render")["refill"] == 5
        assert VT.get_launch_config("rays") == saved["rays"]
        assert VT.get_launch_config("render_best") == saved["render_best"]
        VT.set_launch_config(leaf_batch=7)  # every kind, other fields kept
        for k in saved:
            c = VT.get_launch_config(k)
            assert c["leaf_batch"] == 7
            assert c["split_best"] == saved[k]["split_best"]
        with pytest.raises(ValueError):
            VT.get_launch_config("nope")
        with pytest.raises(_lib.VoxgError):
            VT.set_launch_config("render_best", split_best=5)
        L = _lib.load()
        cfg = _lib.LaunchConfig()
        assert L.voxg_get_kind_launch_config(4, C.byref(cfg)) != 0
        assert L.voxg_get_kind_launch_config(-1, C.byref(cfg)) != 0
        # the plain getter reports the ray-batch kind
        assert L.voxg_get_launch_config(C.byref(cfg)) == 0
        assert cfg.leaf_batch == 7 and cfg.refill == saved["rays"]["refill"]
    assert VT.get_launch_configs() == saved


def _deep_scenes():
    from voxvision_amd import scenes
    rng = np.random.default_rng(17)
    ii = np.arange(-50, 50)
    I, J, K = np.meshgrid(ii, ii, ii, indexing="ij")
    m = (I * I + J * J + K * K) < 2500
    yield "ball50", np.stack([I[m], J[m], K[m]], 1).astype(np.float32)
    yield "menger81", scenes.volume_to_dots(scenes.menger_volume(81))
    # sparse points in a large box: many inner levels
    yield "sparse", np.unique(rng.integers(0, 4096, (3000, 3)),
                              axis=0).astype(np.float32)
    # a diagonal chain of voxels with clusters along it
    t = np.arange(0, 2000, 3)
    chain = np.stack([t, t // 2, t // 3], 1)
    blobs = (chain[::25, None, :] + rng.integers(-2, 3, (len(chain[::25]),
                                                         20, 3))).reshape(-1, 3)
    yield "chain", np.unique(np.concatenate([chain, blobs]),
                             axis=0).astype(np.float32)
    yield "ct256", scenes.config_c2(frames=1)["dots"]


@pytest.mark.parametrize("name", ["ball50", "menger81", "sparse", "chain",
                                  "ct256"])
def test_stack_bound_covers_the_search(voxlib, oracle, name):
    """The traversal stack holds one entry per inner node on the current
    path that still has a child behind a dividing plane to visit; the bound
    the library sizes its stack for (voxg_tree_plan's stack_entries,
    depth - 1) must cover the most such nodes of any search, measured by
    the oracle's instrumented vox_ray_tree_intersection (search.c:84-148)
    on random rays and rays from outside the box on the deepest scenes."""
    from voxvision_amd import scenes, voxtrees as VT
    dots = dict(_deep_scenes())[name]
    info, _ = VT.plan(dots)
    T = oracle.Tree(dots)
    assert info["depth"] == T.stats["depth"]
    assert info["stack_entries"] == max(1, info["depth"] - 1)
    bmin, bmax = scenes.bbox_of(dots)
    o, d = scenes.random_rays(60000, bmin, bmax, seed=3)
    # rays from far outside, aimed at random voxels (long searches)
    rng = np.random.default_rng(4)
    tgt = dots[rng.integers(0, len(dots), 20000)] + 0.5
    src = (bmin + bmax) / 2 + rng.standard_normal((20000, 3)) * \
        float(np.max(bmax - bmin))
    o = np.concatenate([o, src.astype(np.float32)])
    d = np.concatenate([d, (tgt - src).astype(np.float32)])
    mp = oracle.ray_max_pending(T, o, d)
    print(name, "depth", info["depth"], "bound", info["stack_entries"],
          "max pending", int(mp.max()), "mean", float(mp.mean()))
    assert int(mp.max()) <= info["stack_entries"]
    assert int(mp.max()) >= 2
    T.close()


def _light_ops(seed, n_ins=300):
    rng = np.random.default_rng(seed)
    ops = []
    live = []
    for k in range(n_ins):
        c = rng.uniform(0, 64, 3).astype(np.float32)
        if k % 7 == 3 and live:       # a duplicate sphere: refused (0)
            c, r = live[int(rng.integers(0, len(live)))]
            ops.append(("ins", c, r, rng.uniform(0, 1, 3)))
            continue
        r = np.float32(rng.uniform(1, 20))
        col = rng.uniform(0, 1, 3)
        if k % 11 == 5:
            col[1] = 1.5               # colour outside [0, 1]: refused
        ops.append(("ins", c, r, col))
        if col[1] <= 1:
            live.append((c, r))
        if k % 3 == 2 and live:        # delete one present, one absent
            i = int(rng.integers(0, len(live)))
            ops.append(("del", live[i][0], live[i][1]))
            del live[i]
            ops.append(("del", c + 100, r))
    return ops


@pytest.mark.parametrize("seed", [1, 2])
def test_lights_match_the_oracle(voxlib, oracle, seed):
    """The library's light manager keeps the reference's M-tree
    (src/voxtrees/mtree.c) node for node: after every step of a long
    sequence of inserts (with node splits) and deletes (with merges of
    parents left with one child), the light at random points
    (vox_get_light, src/voxrnd/lights.c:103-116) equals the oracle's bit
    for bit, and both refuse the same inserts and deletes."""
    from voxvision_amd.lights import Lights
    G, O = Lights(), oracle.Lights()
    assert G.set_ambient((0.2, 0.3, 0.4)) == O.set_ambient((0.2, 0.3, 0.4))
    assert G.set_ambient((0.2, 1.3, 0.4)) == O.set_ambient((0.2, 1.3, 0.4))
    rng = np.random.default_rng(seed + 10)
    pts = rng.uniform(-5, 70, (400, 3)).astype(np.float32)
    for k, op in enumerate(_light_ops(seed)):
        if op[0] == "ins":
            assert G.insert(op[1], op[2], op[3]) == O.insert(op[1], op[2],
                                                             op[3])
        else:
            assert G.delete(op[1], op[2]) == O.delete(op[1], op[2])
        assert len(G) == len(O)
        if k % 25 == 0:
            assert np.array_equal(G.light(pts).view(np.uint32),
                                  O.light(pts).view(np.uint32)), k
    assert O.height() >= 3
    assert np.array_equal(G.light(pts).view(np.uint32),
                          O.light(pts).view(np.uint32))
    # what the reference's M-tree cannot hold is refused loudly
    c, _, r = O.spheres()
    with pytest.raises(Exception):
        G.insert(c[0], float(r[0]) + 1.0, (1, 1, 1))
    assert O.insert(c[0], float(r[0]) + 1.0, (1, 1, 1)) == -1
    G.clear()
    O.clear()
    assert len(G) == 0 and len(O) == 0
    assert np.array_equal(G.light(pts), O.light(pts))


def _cut_scenes():
    yield from _scenes()
    from voxvision_amd import scenes
    # a sparse plate of voxels at z = 0 and three above it: the root's
    # children above the plate hold one voxel each and are leaves, which
    # the cut must keep from the level above its own
    rng = np.random.default_rng(5)
    ii = np.arange(40, dtype=np.float32)
    X, Y = np.meshgrid(ii, ii, indexing="ij")
    keep = rng.random(X.size) < 0.6
    plate = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size, np.float32)],
                     1)[keep]
    above = np.array([[3, 4, 50], [30, 5, 60], [20, 33, 40]], np.float32)
    yield "plate_above", np.concatenate([plate, above]), (1.0, 1.0, 1.0)
    yield "ct64", scenes.volume_to_dots(scenes.ct_volume(64, seed=3)), \
        (1.0, 1.0, 1.0)


@pytest.mark.parametrize("case", list(range(8)))
def test_tree_cut_covers_every_voxel(voxlib, case):
    """The render pre-pass finishes a block when its rays miss every box
    of the tree's cut, so every voxel's box must lie in one of those boxes
    (a cut that left out the leaves of the levels above it would drop the
    voxels below them).  At most 64 boxes, and more than one, or none."""
    import ctypes as C
    name, dots, V = list(_cut_scenes())[case]
    d = np.ascontiguousarray(dots, np.float32)
    v = (C.c_float * 3)(*V)
    out = np.zeros((64, 6), np.float32)
    cnt = C.c_size_t()
    assert voxlib.voxg_test_plan_cut(d.ctypes.data, len(d), 3, v,
                                     out.ctypes.data, 64, C.byref(cnt)) == 0
    n = cnt.value
    assert n == 0 or 2 <= n <= 64, (name, n)
    if n == 0:
        return
    box = out[:n]
    lo = d[:, None, :]
    hi = (d + np.asarray(V, np.float32))[:, None, :]
    inside = ((box[None, :, 0:3] <= lo) & (hi <= box[None, :, 3:6])).all(2)
    assert inside.any(1).all(), name
    # a cap below the cut's size is refused
    if n > 1:
        assert voxlib.voxg_test_plan_cut(d.ctypes.data, len(d), 3, v,
                                         out.ctypes.data, n - 1,
                                         C.byref(cnt)) == -22
