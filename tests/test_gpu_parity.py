"""GPU parity: libvoxgpu against the CPU oracle on identical seeded inputs.

Bar: bit-exact voxel ids, leaf identity and hit-point bits (the kernels
perform the reference's fp32 operations in the reference's order), exact
ball booleans, exact traversal counters.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def _c1():
    from voxvision_amd import scenes
    return scenes.config_c1()


def _np_u32(t):
    return (t.cpu().numpy().astype(np.int64) & 0xFFFFFFFF).astype(np.uint32)


def _assert_rays_equal(vox_g, hit_g, vox_o, hit_o):
    vg = _np_u32(vox_g)
    assert np.array_equal(vg, vox_o), \
        "voxel ids differ on %d rays" % int((vg != vox_o).sum())
    m = vox_o != MISS
    hg = hit_g.cpu().numpy()
    assert np.array_equal(hg[:, m].view(np.uint32), hit_o[:, m].view(np.uint32))
    assert np.isnan(hg[:, ~m]).all()  # untouched on a miss


def _edge_rays(bmin, bmax):
    """Axis-parallel rays, zero direction components, origins on faces, on
    voxel corners, inside voxels, rays along dividing planes."""
    o, d = [], []
    lo, hi = np.floor(bmin), np.ceil(bmax)
    c = np.floor((lo + hi) / 2)
    for ax in range(3):
        for sgn in (-1, 1):
            for off in (0.0, 0.5, 1.0, 3.0):
                oo = c.copy() + off
                oo[ax] = lo[ax] - 5 if sgn > 0 else hi[ax] + 5
                dd = np.zeros(3)
                dd[ax] = sgn
                o.append(oo)
                d.append(dd)
    rng = np.random.default_rng(99)
    for _ in range(400):
        oo = rng.integers(lo - 3, hi + 3).astype(np.float64)
        dd = rng.integers(-2, 3, 3).astype(np.float64)
        if not dd.any():
            dd[0] = 1
        o.append(oo)
        d.append(dd)
    for _ in range(400):  # origins on a face, inside, half-integers
        oo = rng.integers(lo, hi).astype(np.float64) + \
            rng.choice([0.0, 0.5, 1.0], 3)
        dd = rng.standard_normal(3)
        dd[rng.integers(0, 3)] = 0.0
        o.append(oo)
        d.append(dd)
    return np.array(o, np.float32), np.array(d, np.float32)


def test_ray_batch_random_and_edge_rays(gpu, oracle):
    from voxvision_amd import scenes, voxtrees as VT
    c1 = _c1()
    dots = c1["dots"]
    T = oracle.Tree(dots)
    G = VT.make_tree(dots)
    assert G.hash() == T.hash()
    bmin, bmax = c1["bbox"]
    o, d = scenes.random_rays(200000, bmin, bmax, seed=3)
    eo, ed = _edge_rays(bmin, bmax)
    o = np.concatenate([o, eo])
    d = np.concatenate([d, ed])
    vox_o, hit_o = T.rays(o, d)
    vox_g, hit_g = VT.ray_tree_intersection(G, o, d)
    torch.cuda.synchronize()
    _assert_rays_equal(vox_g, hit_g, vox_o, hit_o)
    assert (vox_o != MISS).mean() > 0.05


def test_ray_counts_match_oracle(gpu, oracle):
    from voxvision_amd import scenes, voxtrees as VT
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    o, d = scenes.random_rays(50000, *c1["bbox"], seed=4)
    _, _, (nv, nt) = T.rays(o, d, with_counts=True)
    cnt = VT.ray_counts(G, o, d)
    assert cnt["node_visits"] == nv and cnt["voxel_tests"] == nt
    assert cnt["rays"] == len(o)


def test_upload_reference_layout_tree(gpu, oracle):
    """A tree in the reference's struct vox_node layout (built by the
    oracle's vox_make_tree) flattens to the same results, and leaf indices
    map back to the reference nodes the CPU search returns."""
    from voxvision_amd import scenes, voxtrees as VT
    rng = np.random.default_rng(8)
    dots = scenes.volume_to_dots(rng.random((24, 24, 24)) < 0.2)
    T = oracle.Tree(dots)
    G = VT.Tree.upload(T.ptr)
    assert G.hash() == T.hash()
    o, d = scenes.random_rays(5000, *scenes.bbox_of(dots), seed=6)
    vox_o, hit_o = T.rays(o, d)
    vox_g, hit_g, leaf_g = VT.ray_tree_intersection(G, o, d, want_leaf=True)
    _assert_rays_equal(vox_g, hit_g, vox_o, hit_o)
    L = oracle.lib()
    leaf_g = _np_u32(leaf_g)
    for i in range(0, len(o), 97):
        od = np.zeros(4, np.float32)
        dd = np.zeros(4, np.float32)
        res = np.zeros(4, np.float32)
        od[:3] = o[i]
        dd[:3] = d[i]
        p = L.orc_ray_tree_intersection(T.ptr, oracle._fp(od), oracle._fp(dd),
                                        oracle._fp(res))
        if p is None:
            assert leaf_g[i] == MISS
        else:
            assert G.node_ptr(int(leaf_g[i])) == p


def test_voxel_ids_map_to_the_oracle_voxels(gpu, oracle):
    from voxvision_amd import voxtrees as VT
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    n = T.stats["voxel_ids"]
    ids = np.arange(n, dtype=np.uint32)
    assert np.array_equal(G.voxel_coords(ids), T.voxel_table())


def test_ball_batch(gpu, oracle):
    from voxvision_amd import scenes, voxtrees as VT
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    c, r = c1["balls_c"], c1["balls_r"]
    out_g = VT.tree_ball_collidep(G, c, r).cpu().numpy()
    assert np.array_equal(out_g, T.balls(c, r))
    assert np.array_equal(out_g, oracle.brute_balls(c1["dots"], c, r))
    c, r = scenes.random_balls(100000, *c1["bbox"], seed=17)
    out_o, (nv, nt) = T.balls(c, r, with_counts=True)
    assert np.array_equal(VT.tree_ball_collidep(G, c, r).cpu().numpy(), out_o)
    cnt = VT.ball_counts(G, c, r)
    assert cnt["rays"] == len(c)
    assert 0 < out_o.mean() < 1


def test_render_best_matches_oracle(gpu, oracle):
    from voxvision_amd import voxrnd
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    from voxvision_amd import voxtrees as VT
    G = VT.make_tree(c1["dots"])
    tex = np.random.default_rng(2).integers(0, 256, 32 ** 3).astype(np.uint8)
    W, H = c1["W"], c1["H"]
    sq_o, vox_o, hit_o, _ = T.render(c1["camera"], W, H, oracle.QUALITY_BEST,
                                     texture=tex)
    out = voxrnd.render(G, [c1["camera"]], W, H, texture=tex)
    torch.cuda.synchronize()
    _assert_rays_equal(out["voxel"], out["hit"], vox_o, hit_o)
    sq_g = _np_u32(out["squares"])
    assert np.array_equal(sq_g, sq_o)


def test_render_many_poses_ct(gpu, oracle):
    """Two of the 64 orbit poses of the 256^3 scene at 1920x1080, compared
    pixel by pixel with the oracle."""
    from voxvision_amd import scenes, voxrnd, voxtrees as VT
    c2 = scenes.config_c2(frames=64)
    T = oracle.Tree(c2["dots"])
    G = VT.make_tree(c2["dots"])
    W, H = c2["W"], c2["H"]
    picks = [0, 37]
    cams = [c2["cameras"][k] for k in picks]
    out = voxrnd.render(G, cams, W, H, want_colors=False)
    torch.cuda.synchronize()
    vg = _np_u32(out["voxel"]).reshape(len(picks), -1)
    hg = out["hit"].cpu().numpy().reshape(3, len(picks), -1)
    for i, cam in enumerate(cams):
        _, vox_o, hit_o, _ = T.render(cam, W, H, oracle.QUALITY_BEST)
        assert np.array_equal(vg[i], vox_o)
        m = vox_o != MISS
        assert np.array_equal(hg[:, i, m].view(np.uint32),
                              hit_o[:, m].view(np.uint32))


def test_stack_spill_is_exact(gpu, oracle):
    """A 2-entry LDS ring forces the HBM spill path on a deep tree; the
    results and traversal counts must not change, only the number of
    entries moved to HBM (none with a 32-entry ring)."""
    from voxvision_amd import scenes, voxtrees as VT
    ii = np.arange(-50, 50)
    I, J, K = np.meshgrid(ii, ii, ii, indexing="ij")
    m = (I * I + J * J + K * K) < 2500
    dots = np.stack([I[m], J[m], K[m]], 1).astype(np.float32)
    G = VT.make_tree(dots)
    T = oracle.Tree(dots)
    o, d = scenes.random_rays(60000, *scenes.bbox_of(dots), seed=12)
    vox_o, hit_o = T.rays(o, d)
    with VT.saved_launch_configs():
        spills, visits = {}, set()
        for k in (2, 8, 32):
            VT.set_launch_config(stack_lds=k, block=64 if k == 32 else 256)
            vox_g, hit_g = VT.ray_tree_intersection(G, o, d)
            torch.cuda.synchronize()
            _assert_rays_equal(vox_g, hit_g, vox_o, hit_o)
            c = VT.ray_counts(G, o, d)
            spills[k] = c["stack_spills"]
            visits.add((c["node_visits"], c["voxel_tests"]))
        assert len(visits) == 1
        assert spills[2] > spills[8] >= spills[32] == 0


def test_launch_kinds_reach_their_launches(gpu, oracle):
    """Each kind of launch runs with its own configuration: a 2-entry
    stack ring set for BEST renders only makes those spill to HBM, while
    ray batches and ADAPTIVE renders (32 entries) do not; results stay the
    oracle's."""
    from voxvision_amd import scenes, voxrnd, voxtrees as VT
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    W, H = c1["W"], c1["H"]
    cam = [c1["camera"]]
    o, d = scenes.random_rays(20000, *c1["bbox"], seed=5)
    with VT.saved_launch_configs():
        VT.set_launch_config(stack_lds=32, block=64)
        VT.set_launch_config("render_best", stack_lds=2, block=256)
        best = voxrnd.render_counts(G, cam, W, H, quality=0)
        adap = voxrnd.render_counts(G, cam, W, H, quality=2)
        rays = VT.ray_counts(G, o, d)
        out = voxrnd.render(G, cam, W, H, quality=0, want_colors=False)
        torch.cuda.synchronize()
    assert best["stack_spills"] > 0
    assert adap["stack_spills"] == 0 and rays["stack_spills"] == 0
    _, vox_o, hit_o, st = T.render(c1["camera"], W, H, 0)
    _assert_rays_equal(out["voxel"], out["hit"], vox_o, hit_o)
    assert best["node_visits"] == st["node_visits"]


def test_empty_tree_and_empty_batches(gpu, oracle):
    from voxvision_amd import voxtrees as VT
    G = VT.make_tree(np.zeros((0, 3), np.float32))
    o = np.zeros((100, 3), np.float32)
    d = np.ones((100, 3), np.float32)
    vox, hit = VT.ray_tree_intersection(G, o, d)
    assert (_np_u32(vox) == MISS).all()
    assert VT.tree_ball_collidep(G, o, np.ones(100)).sum().item() == 0
    G2 = VT.make_tree(np.ones((5, 3), np.float32))
    vox, _ = VT.ray_tree_intersection(G2, o[:0], d[:0])
    assert vox.numel() == 0


def test_single_voxel_known_answers(gpu, oracle):
    """benchmarks/voxtrees/hit-box.c and ball-intersection.c closed forms."""
    from voxvision_amd import voxtrees as VT
    G = VT.make_tree(np.array([[0, 10, 0]], np.float32))
    j = np.arange(10)
    o = np.stack([j - 5, np.zeros(10), np.zeros(10)], 1).astype(np.float32)
    d = np.stack([(j - 5) / 10.0, np.ones(10), np.zeros(10)],
                 1).astype(np.float32)
    vox, hit = VT.ray_tree_intersection(G, o, d)
    v = _np_u32(vox)
    assert int((v != MISS).sum()) == 1
    k = int(np.nonzero(v != MISS)[0][0])
    assert tuple(hit[:, k].cpu().numpy()) == (0.0, 10.0, 0.0)
    G = VT.make_tree(np.zeros((1, 3), np.float32))
    jj = np.arange(-100, 100)
    c = np.stack([np.zeros(200), np.full(200, 10.0), jj], 1)
    assert VT.tree_ball_collidep(G, c, np.full(200, 1000.0)).sum().item() == 200


def test_invalid_arguments_are_rejected(gpu):
    from voxvision_amd import _lib, voxtrees as VT
    L = _lib.load()
    G = VT.make_tree(np.ones((3, 3), np.float32))
    x = torch.zeros(4, device="cuda")
    rc = L.voxg_ray_batch(G.handle, 4, None, None, None, None, None, None,
                          None, None, None, None, None, None)
    assert rc == -22
    cam = torch.zeros(11, device="cuda")
    rc = L.voxg_render(G.handle, 1, cam.data_ptr(), 20, 8, None, None,
                       x.data_ptr(), None, None)
    assert rc == -22  # W % 16 != 0
    assert L.voxg_ray_batch(None, 4, *([x.data_ptr()] * 6), None, None, None,
                            None, None, None) == -22


@pytest.mark.parametrize("quality", [1, 2, 2 | 4, 2 | 8])
def test_render_fast_adaptive_merge_match_oracle(gpu, oracle, quality):
    """FAST, ADAPTIVE and both ray-merge modes (renderer.c:200-319): the
    per-block decisions, the leaf-reuse chain, merged pixels and colours
    all match the oracle; so do the mode and traversal counters."""
    from voxvision_amd import voxrnd, voxtrees as VT
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    tex = np.random.default_rng(4).integers(0, 256, 32 ** 3).astype(np.uint8)
    W, H = c1["W"], c1["H"]
    sq_o, vox_o, hit_o, st = T.render(c1["camera"], W, H, quality,
                                      texture=tex)
    out = voxrnd.render(G, [c1["camera"]], W, H, quality=quality,
                        texture=tex)
    torch.cuda.synchronize()
    _assert_rays_equal(out["voxel"], out["hit"], vox_o, hit_o)
    assert np.array_equal(_np_u32(out["squares"]), sq_o)
    cnt = voxrnd.render_counts(G, [c1["camera"]], W, H, quality=quality)
    assert cnt["fast_blocks"] == st["fast_blocks"]
    assert cnt["best_blocks"] == st["best_blocks"]
    assert cnt["merge_blocks"] == st["merge_blocks"]
    assert cnt["rays"] == st["root_traversals"]
    assert cnt["leaf_traversals"] == st["leaf_traversals"]
    assert cnt["node_visits"] == st["node_visits"]
    assert cnt["voxel_tests"] == st["voxel_tests"]


@pytest.mark.parametrize("q", [2, 2 | 4, 2 | 8, 1])
def test_render_adaptive_ct_frames(gpu, oracle, q):
    """ADAPTIVE (the reference's default quality, renderer.c:78), ADAPTIVE
    with plain or accurate ray merging and FAST on two 1080p poses of the
    256^3 scene: many blocks of both modes, long leaf-reuse chains."""
    from voxvision_amd import scenes, voxrnd, voxtrees as VT
    c2 = scenes.config_c2(frames=64)
    T = oracle.Tree(c2["dots"])
    G = VT.make_tree(c2["dots"])
    W, H = c2["W"], c2["H"]
    cams = [c2["cameras"][k] for k in (5, 50)]
    out = voxrnd.render(G, cams, W, H, quality=q, want_colors=False)
    torch.cuda.synchronize()
    vg = _np_u32(out["voxel"]).reshape(2, -1)
    hg = out["hit"].cpu().numpy().reshape(3, 2, -1)
    for i, cam in enumerate(cams):
        _, vox_o, hit_o, _ = T.render(cam, W, H, q)
        assert np.array_equal(vg[i], vox_o)
        m = vox_o != MISS
        assert np.array_equal(hg[:, i, m].view(np.uint32),
                              hit_o[:, m].view(np.uint32))


@pytest.mark.parametrize("quality", [0, 1, 2, 6, 10])
def test_render_with_and_without_block_culling(gpu, oracle, quality):
    """The pre-pass that finishes blocks missing the scene box, BEST
    mode's units of 16 >> split_best pixels, and the three-pass path of
    FAST and ADAPTIVE renders (taken only with the pre-pass; without it
    they trace block by block) change no output and no counter, in every
    quality mode."""
    from voxvision_amd import voxrnd, voxtrees as VT
    c1 = _c1()
    G = VT.make_tree(c1["dots"])
    tex = np.random.default_rng(4).integers(0, 256, 32 ** 3).astype(np.uint8)
    W, H = c1["W"], c1["H"]
    res = []
    with VT.saved_launch_configs():
        for cull, split in ((0, 0), (1, 0), (1, 1), (0, 2), (1, 3), (1, 4)):
            VT.set_launch_config(render_cull=cull, split_best=split)
            out = voxrnd.render(G, [c1["camera"]], W, H, quality=quality,
                                texture=tex)
            cnt = voxrnd.render_counts(G, [c1["camera"]], W, H,
                                       quality=quality)
            torch.cuda.synchronize()
            res.append((out, cnt))
    (a, ca) = res[0]
    for b, cb in res[1:]:
        assert torch.equal(a["voxel"], b["voxel"])
        assert torch.equal(a["squares"], b["squares"])
        m = a["voxel"] != -1
        assert torch.equal(a["hit"][:, m], b["hit"][:, m])
        assert ca == cb
    T = oracle.Tree(c1["dots"])
    _, vox_o, _, st = T.render(c1["camera"], W, H, quality, texture=tex)
    assert np.array_equal(_np_u32(b["voxel"]), vox_o)
    assert cb["rays"] == st["root_traversals"]
    assert cb["node_visits"] == st["node_visits"]


@pytest.mark.parametrize("leaf_batch,finish_batch", [(1, 1), (64, 64),
                                                     (8, 32), (32, 4)])
def test_lane_batching_changes_nothing(gpu, oracle, leaf_batch, finish_batch):
    """Waves gather lanes at leaves and at ended traversals before they
    serve them; any thresholds give the oracle's rays, pixels and counts."""
    from voxvision_amd import scenes, voxrnd, voxtrees as VT
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    o, d = scenes.random_rays(20000, *c1["bbox"], seed=23)
    vox_o, hit_o = T.rays(o, d)
    W, H = c1["W"], c1["H"]
    q = 2 | 8
    _, pvox_o, phit_o, st = T.render(c1["camera"], W, H, q)
    with VT.saved_launch_configs():
        VT.set_launch_config(leaf_batch=leaf_batch, finish_batch=finish_batch)
        vox_g, hit_g = VT.ray_tree_intersection(G, o, d)
        out = voxrnd.render(G, [c1["camera"]], W, H, quality=q,
                            want_colors=False)
        cnt = voxrnd.render_counts(G, [c1["camera"]], W, H, quality=q)
        outb = voxrnd.render(G, [c1["camera"]], W, H, quality=0,
                             want_colors=False)
        torch.cuda.synchronize()
    _assert_rays_equal(vox_g, hit_g, vox_o, hit_o)
    _assert_rays_equal(out["voxel"], out["hit"], pvox_o, phit_o)
    _, bvox_o, bhit_o, _ = T.render(c1["camera"], W, H, 0)
    _assert_rays_equal(outb["voxel"], outb["hit"], bvox_o, bhit_o)
    assert cnt["node_visits"] == st["node_visits"]
    assert cnt["voxel_tests"] == st["voxel_tests"]
    assert cnt["fast_blocks"] == st["fast_blocks"]


@pytest.mark.parametrize("chunk", [1, 4])
def test_ray_batch_every_size_is_complete(gpu, oracle, chunk):
    """Units are handed out from per-XCD ranges with stealing; ray counts
    around the range and chunk boundaries must still trace every ray once
    (outputs start as a sentinel that no traced ray writes)."""
    from voxvision_amd import scenes, voxtrees as VT
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    o, d = scenes.random_rays(5000, *c1["bbox"], seed=31)
    vox_o, hit_o = T.rays(o, d)
    with VT.saved_launch_configs():
        VT.set_launch_config(chunk=chunk)
        for n in (1, 2, 7, 8, 9, 63, 64, 65, 511, 1000, 4097, 5000):
            vox_g, hit_g = VT.ray_tree_intersection(G, o[:n], d[:n])
            torch.cuda.synchronize()
            _assert_rays_equal(vox_g, hit_g, vox_o[:n], hit_o[:, :n])


def test_block_pre_pass_never_drops_a_hit(gpu, oracle):
    """The pre-pass finishes a block only when none of its 16 rays enters
    the scene box.  Cameras close to the box, at its silhouette, inside it
    and far away, with narrow and wide fields of view: the render with the
    pre-pass equals the one without, pixel for pixel, and the oracle's."""
    from voxvision_amd import camera, scenes, voxrnd, voxtrees as VT
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    bmin, bmax = c1["bbox"]
    ctr = (bmin + bmax) / 2
    diag = float(np.linalg.norm(bmax - bmin))
    rng = np.random.default_rng(41)
    W, H = 160, 96
    cams = []
    for k in range(24):
        dist = [0.3, 0.55, 0.75, 1.0, 3.0, 20.0][k % 6] * diag
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pos = ctr + dist * u
        # look at the centre, or graze the box edge
        tgt = ctr if k % 2 == 0 else bmin + (bmax - bmin) * rng.random(3)
        fov = [np.tan(np.radians(30)), np.tan(np.radians(5)),
               np.tan(np.radians(60))][k % 3]
        cams.append(camera.make_camera(pos, tgt, fov, W, H))
    outs = []
    with VT.saved_launch_configs():
        for cull in (0, 1):
            VT.set_launch_config(render_cull=cull)
            o = voxrnd.render(G, cams, W, H, want_colors=False)
            torch.cuda.synchronize()
            outs.append(o)
    assert torch.equal(outs[0]["voxel"], outs[1]["voxel"])
    vg = _np_u32(outs[1]["voxel"]).reshape(len(cams), -1)
    for i in (0, 3, 7, 12, 17, 22):
        _, vox_o, _, _ = T.render(cams[i], W, H, oracle.QUALITY_BEST)
        assert np.array_equal(vg[i], vox_o)


def test_render_rejects_oversized_frames(gpu):
    """Block corners are 16-bit: a side of 65536 pixels or more is refused
    before any launch, as is a frame count that makes 2^31 pixels."""
    from voxvision_amd import _lib, scenes, voxrnd, voxtrees as VT
    c1 = _c1()
    G = VT.make_tree(c1["dots"])
    with pytest.raises(_lib.VoxgError, match="65536"):
        voxrnd.render(G, [c1["camera"]], 65536, 4, want_colors=False,
                      want_hits=False)
    with pytest.raises(_lib.VoxgError, match="65536"):
        voxrnd.render(G, [c1["camera"]], 16, 65536, want_colors=False,
                      want_hits=False)
    # frame counts whose pixel count reaches 2^31, also where the product
    # n_frames * W * H would wrap around 2^64
    L = _lib.load()
    cams = voxrnd.cameras_tensor([c1["camera"]])
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n_frames, W, H in ((1 << 25, 16, 4), (1 << 60, 16, 4),
                           ((1 << 62) + 1, 64, 64)):
        rc = L.voxg_render(G.handle, n_frames, cams.data_ptr(), W, H, None,
                           None, None, None, stream)
        assert rc == -34, L.voxg_last_error()  # VOXG_ERANGE
    torch.cuda.synchronize()


def test_large_ball_batch(gpu, oracle):
    """300 K balls on the 256^3 scene: every answer is the oracle's,
    including centres outside the scene box and centres that are not
    finite."""
    from voxvision_amd import scenes, voxtrees as VT
    c2 = scenes.config_c2(frames=1)
    T = oracle.Tree(c2["dots"])
    G = VT.make_tree(c2["dots"])
    bmin, bmax = c2["bbox"]
    ext = bmax - bmin
    rng = np.random.default_rng(17)
    n = 300000
    c = rng.uniform(bmin - 0.2 * ext, bmax + 0.2 * ext, (n, 3)).astype(np.float32)
    r = rng.uniform(0.25, 8.0, n).astype(np.float32)
    c[::997, 0] = np.nan
    c[1::991, 1] = np.inf
    c[2::983, 2] = -np.inf
    got = VT.tree_ball_collidep(G, c, r)
    torch.cuda.synchronize()
    want = T.balls(c, r)
    assert np.array_equal(got.cpu().numpy(), want)
    assert 0.05 < want.mean() < 0.95


def test_incoherent_rays_and_balls_at_full_size(gpu, oracle):
    """The stress workload on the 256^3 scene at its full size: 16 M rays
    with origins in the box grown by 10% per side and random directions,
    and 4 M ball queries with radii in [0.5, 8].  2 M of the rays are
    compared with the oracle bit for bit; of the full launches every 61st
    ray and every 17th ball is, and each full launch gives the same
    answers as separate launches of its own slices."""
    from voxvision_amd import scenes, voxtrees as VT
    c2 = scenes.config_c2(frames=1)
    T = oracle.Tree(c2["dots"])
    G = VT.make_tree(c2["dots"])
    bmin, bmax = c2["bbox"]
    o2, d2 = scenes.random_rays(1 << 21, bmin, bmax, seed=41)
    vox_g, hit_g = VT.ray_tree_intersection(G, o2, d2)
    torch.cuda.synchronize()
    vox_o, hit_o = T.rays(o2, d2)
    _assert_rays_equal(vox_g, hit_g, vox_o, hit_o)
    assert 0.02 < (vox_o != MISS).mean() < 0.9

    n = 1 << 24
    o, d = scenes.random_rays(n, bmin, bmax, seed=11)
    vox_f, hit_f = VT.ray_tree_intersection(G, o, d)
    torch.cuda.synchronize()
    pick = np.arange(0, n, 61)
    vox_o, hit_o = T.rays(o[pick], d[pick])
    _assert_rays_equal(vox_f[pick], hit_f[:, pick], vox_o, hit_o)
    for a, b in ((0, 1000), (n // 2 - 777, n // 2 + 4096), (n - 65536, n)):
        vs, hs = VT.ray_tree_intersection(G, o[a:b], d[a:b])
        torch.cuda.synchronize()
        assert torch.equal(vs, vox_f[a:b])
        m = vs != MISS
        assert torch.equal(hs[:, m], hit_f[:, a:b][:, m])

    nb = 1 << 22
    c, r = scenes.random_balls(nb, bmin, bmax, seed=13)
    got = VT.tree_ball_collidep(G, c, r)
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    bpick = np.arange(0, nb, 17)
    want = T.balls(c[bpick], r[bpick])
    assert np.array_equal(got[bpick], want)
    assert 0.05 < want.mean() < 0.95
    for a, b in ((0, 999), (nb - 70000, nb)):
        part = VT.tree_ball_collidep(G, c[a:b], r[a:b])
        torch.cuda.synchronize()
        assert np.array_equal(part.cpu().numpy(), got[a:b])


@pytest.mark.parametrize("quality", [0, 2])
def test_render_into_unaligned_voxel_plane(gpu, quality):
    """The pre-pass writes the miss ids of the blocks it finishes in 16-byte
    stores from the whole workgroup when the voxel plane is 16-byte
    aligned, and one by one from the testing thread when it is not: a
    plane offset by 4 bytes gets the same ids, and the words around it
    stay untouched."""
    from voxvision_amd import voxrnd, voxtrees as VT
    c1 = _c1()
    G = VT.make_tree(c1["dots"])
    W, H, cam = c1["W"], c1["H"], c1["camera"]
    n = W * H
    ref = voxrnd.render(G, [cam], W, H, quality=quality, want_colors=False)
    buf = torch.full((n + 2,), 7, dtype=torch.int32, device="cuda")
    hit = torch.empty((3, n), dtype=torch.float32, device="cuda")
    out = voxrnd.render(G, [cam], W, H, quality=quality, want_colors=False,
                        out={"voxel": buf[1:n + 1], "hit": hit})
    torch.cuda.synchronize()
    assert buf[1:n + 1].data_ptr() % 16 == 4
    assert torch.equal(out["voxel"], ref["voxel"])
    assert int(buf[0]) == 7 and int(buf[n + 1]) == 7
    assert (ref["voxel"] == -1).float().mean() > 0.05


@pytest.mark.parametrize("quality", [0, 1, 2, 2 | 4, 2 | 8])
def test_render_three_pass_edge_cases(gpu, oracle, quality):
    """FAST, ADAPTIVE and ADAPTIVE with either kind of ray merging take
    the speculated passes (root searches, in-order replay, output; with
    plain merging the odd pixels only for the blocks whose merging stops);
    BEST the pre-pass and one pixel per unit.  Edge cases against the
    oracle: a camera looking away (no block queued), a camera inside the
    scene box (with plain merging, blocks that keep merging and blocks
    that stop), frames with an odd number of block rows, and 40 frames of
    16x4 pixels (one block row per frame, so a pre-pass tile spans many
    frames)."""
    from voxvision_amd import camera, voxrnd, voxtrees as VT
    c1 = _c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    bmin, bmax = c1["bbox"]
    ctr = (bmin + bmax) / 2
    diag = float(np.linalg.norm(bmax - bmin))
    fov = np.tan(np.radians(30))
    cases = []
    away = camera.make_camera(ctr + [0, -diag, 0], ctr + [0, -3 * diag, 0],
                              fov, 64, 48)
    inside = camera.make_camera(ctr, ctr + [diag, 0.3 * diag, 0.1 * diag],
                                fov, 64, 48)
    cases.append(([away], 64, 48))
    cases.append(([inside], 64, 48))
    cases.append(([inside, away, inside], 48, 44))
    rng = np.random.default_rng(3)
    tiny = []
    for k in range(40):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        tiny.append(camera.make_camera(ctr + 1.2 * diag * u, ctr, fov, 16, 4))
    cases.append((tiny, 16, 4))
    for cams, W, H in cases:
        out = voxrnd.render(G, cams, W, H, quality=quality, want_colors=False)
        torch.cuda.synchronize()
        vg = _np_u32(out["voxel"]).reshape(len(cams), -1)
        hg = out["hit"].cpu().numpy().reshape(3, len(cams), -1)
        for i, cam in enumerate(cams):
            _, vox_o, hit_o, _ = T.render(cam, W, H, quality)
            assert np.array_equal(vg[i], vox_o), (W, H, i)
            m = vox_o != MISS
            assert np.array_equal(hg[:, i, m].view(np.uint32),
                                  hit_o[:, m].view(np.uint32))
    # the away camera sees nothing; the inside one sees something
    _, v_away, _, _ = T.render(away, 64, 48, quality)
    _, v_in, _, st_in = T.render(inside, 64, 48, quality)
    assert (v_away == MISS).all() and (v_in != MISS).any()
    if quality == 2 | 4:
        # plain merging starts in every block and stops in some
        assert 0 < st_in["merge_blocks"] < st_in["blocks"]


@pytest.mark.parametrize("quality", [0, 2])
def test_cut_keeps_blocks_near_lone_voxels(gpu, oracle, quality):
    """The pre-pass finishes blocks whose rays miss every box of the tree's
    cut.  Here three lone voxels above a sparse plate sit in leaves one
    level above the cut's own: cameras aimed at them (and grazing past
    them) must still see them, as the oracle does."""
    from voxvision_amd import camera, voxrnd, voxtrees as VT
    rng = np.random.default_rng(5)
    ii = np.arange(40, dtype=np.float32)
    X, Y = np.meshgrid(ii, ii, indexing="ij")
    keep = rng.random(X.size) < 0.6
    plate = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size, np.float32)],
                     1)[keep]
    lone = np.array([[3, 4, 50], [30, 5, 60], [20, 33, 40]], np.float32)
    dots = np.concatenate([plate, lone])
    T = oracle.Tree(dots)
    G = VT.make_tree(dots)
    W, H = 320, 240
    fov = np.tan(np.radians(20))
    cams = []
    for p in lone:
        c = p + 0.5
        for off in ([60, 5, 3], [-40, 30, 10], [5, -70, -2]):
            cams.append(camera.make_camera(c + np.asarray(off, np.float32),
                                           c + rng.normal(0, 0.3, 3), fov,
                                           W, H))
    out = voxrnd.render(G, cams, W, H, quality=quality, want_colors=False)
    torch.cuda.synchronize()
    vg = _np_u32(out["voxel"]).reshape(len(cams), -1)
    hg = out["hit"].cpu().numpy().reshape(3, len(cams), -1)
    lone_ids = set()
    for i, cam in enumerate(cams):
        _, vox_o, hit_o, _ = T.render(cam, W, H, quality)
        assert np.array_equal(vg[i], vox_o), i
        m = vox_o != MISS
        assert np.array_equal(hg[:, i, m].view(np.uint32),
                              hit_o[:, m].view(np.uint32)), i
        lone_ids |= set(int(v) for v in np.unique(vox_o[m]))
    # the lone voxels were seen (their ids are among the hits)
    coords = G.voxel_coords(np.array(sorted(lone_ids), np.uint32))
    seen = {tuple(map(float, c)) for c in coords}
    assert all(tuple(map(float, p)) in seen for p in lone)
