"""Edited trees on the device: after inserts and deletes
(vox_insert_voxel / vox_delete_voxel), a committed version answers rays,
balls and renders exactly like the oracle's tree after the same edits;
earlier versions keep answering for the tree they were committed from;
the serving loop switches versions for later requests only."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


def _u32(t):
    return (t.cpu().numpy().astype(np.int64) & 0xFFFFFFFF).astype(np.uint32)


def _edits(rng, n, lo, hi):
    out = []
    for _ in range(n):
        p = np.floor(rng.uniform(lo, hi, 3))
        out.append(("ins" if rng.random() < 0.5 else "del", tuple(p)))
    return out


def test_committed_versions_match_oracle(gpu, oracle):
    from voxvision_amd import scenes, voxtrees as VT
    from voxvision_amd.edit import EditTree
    c1 = scenes.config_c1()
    dots = c1["dots"]
    bmin, bmax = c1["bbox"]
    E = EditTree(dots)
    T = oracle.Tree(dots)
    rng = np.random.default_rng(11)
    o, d = scenes.random_rays(60000, bmin, bmax, seed=12)
    cb, rb = scenes.random_balls(20000, bmin, bmax, seed=13)
    versions = []
    for round_ in range(3):
        for op, p in _edits(rng, 400, bmin - 2, bmax + 2):
            a = E.insert(p) if op == "ins" else E.delete(p)
            b = T.insert(p) if op == "ins" else T.delete(p)
            assert a == bool(b)
        # carve a slab through the middle: many deletions inside dense leaves
        z = int((bmin[2] + bmax[2]) / 2) + round_
        for x in range(int(bmin[0]), int(bmax[0])):
            for y in range(int(bmin[1]), int(bmax[1]), 2):
                assert E.delete((x, y, z)) == bool(T.delete((x, y, z)))
        assert E.hash() == T.hash()
        G = E.commit()
        assert G.hash() == T.hash()
        vox_o, hit_o = T.rays(o, d)
        vox_g, hit_g = VT.ray_tree_intersection(G, o, d)
        torch.cuda.synchronize()
        assert np.array_equal(_u32(vox_g), vox_o)
        m = vox_o != MISS
        assert np.array_equal(hit_g.cpu().numpy()[:, m].view(np.uint32),
                              hit_o[:, m].view(np.uint32))
        assert np.array_equal(VT.tree_ball_collidep(G, cb, rb).cpu().numpy(),
                              T.balls(cb, rb))
        versions.append((G, vox_o))
    # earlier versions are untouched by later edits and commits
    for G, vox_o in versions:
        vox_g, _ = VT.ray_tree_intersection(G, o, d)
        torch.cuda.synchronize()
        assert np.array_equal(_u32(vox_g), vox_o)


def test_render_after_edits_matches_oracle(gpu, oracle):
    from voxvision_amd import scenes, voxrnd
    from voxvision_amd.edit import EditTree
    c1 = scenes.config_c1()
    E = EditTree(c1["dots"])
    T = oracle.Tree(c1["dots"])
    bmin, bmax = c1["bbox"]
    rng = np.random.default_rng(5)
    for op, p in _edits(rng, 3000, bmin, bmax):
        assert (E.insert(p) if op == "ins" else E.delete(p)) == \
            bool(T.insert(p) if op == "ins" else T.delete(p))
    G = E.commit()
    W, H = c1["W"], c1["H"]
    for q in (0, 2 | 8):
        _, vox_o, hit_o, st = T.render(c1["camera"], W, H, q)
        out = voxrnd.render(G, [c1["camera"]], W, H, quality=q,
                            want_colors=False)
        cnt = voxrnd.render_counts(G, [c1["camera"]], W, H, quality=q)
        torch.cuda.synchronize()
        assert np.array_equal(_u32(out["voxel"]), vox_o)
        m = vox_o != MISS
        assert np.array_equal(out["hit"].cpu().numpy()[:, m].view(np.uint32),
                              hit_o[:, m].view(np.uint32))
        assert cnt["node_visits"] == st["node_visits"]


def test_server_switches_versions_for_later_requests(gpu, oracle):
    from voxvision_amd import scenes, server
    from voxvision_amd.edit import EditTree
    c1 = scenes.config_c1()
    E = EditTree(c1["dots"])
    T0 = oracle.Tree(c1["dots"])
    G0 = E.commit()
    bmin, bmax = c1["bbox"]
    # a wall of new voxels between the camera and the scene
    cam = c1["camera"]
    pos = np.array(cam["position"], np.float32)
    mid = (bmin + bmax) / 2
    wall_at = pos + 0.5 * (mid - pos)
    T1 = oracle.Tree(c1["dots"])
    for a in range(-20, 21):
        for b in range(-20, 21):
            p = wall_at + np.array([a, b, 0], np.float32)
            E.insert(p)
            T1.insert(p)
    G1 = E.commit()
    assert G1.hash() == T1.hash() and G1.hash() != G0.hash()
    W, H = 320, 240
    from voxvision_amd import camera
    cam_s = dict(cam)
    cam_s["mul"], cam_s["xsub"], cam_s["ysub"] = camera.set_window_size(
        cam["fov"], W, H)
    S = server.Server(streams=2)
    tid = S.add_tree(G0)
    t_old, out_old = S.submit_render(tid, cam_s, W, H)
    S.replace_tree(tid, G1)
    t_new, out_new = S.submit_render(tid, cam_s, W, H)
    S.wait(t_old)
    S.wait(t_new)
    _, vox0, _, _ = T0.render(cam_s, W, H, 0)
    _, vox1, _, _ = T1.render(cam_s, W, H, 0)
    assert not np.array_equal(vox0, vox1)
    assert np.array_equal(_u32(out_old["voxel"]), vox0)
    assert np.array_equal(_u32(out_new["voxel"]), vox1)
    S.close()
