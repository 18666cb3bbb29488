"""Batched multi-tree form: one persistent launch over many uneven trees
gives, ray for ray, the results of a plain launch per tree (and the
oracle's), also when captured in a HIP graph and replayed."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


def _small_c4():
    from voxvision_amd import scenes
    return scenes.config_c4(seed=5, n_scenes=9, edge_range=(16, 48),
                            ray_range=(256, 20000))


def _concat(c4):
    o = np.concatenate([s["o"] for s in c4])
    d = np.concatenate([s["d"] for s in c4])
    first = np.cumsum([0] + [len(s["o"]) for s in c4])[:-1]
    return o, d, first


def test_batch_equals_plain_and_oracle(gpu, oracle):
    from voxvision_amd import batch, voxtrees as VT
    c4 = _small_c4()
    c4.append(dict(dots=np.zeros((0, 3), np.float32),   # an empty tree
                   o=np.zeros((50, 3), np.float32),
                   d=np.ones((50, 3), np.float32)))
    o, d, first = _concat(c4)
    trees = [VT.make_tree(s["dots"]) for s in c4]
    items = [(t, f, len(s["o"])) for t, f, s in zip(trees, first, c4)]
    B = batch.Batch(items)
    n = len(o)
    so, sd = batch.soa(o), batch.soa(d)
    vox = torch.empty(n, dtype=torch.int32, device="cuda")
    leaf = torch.empty(n, dtype=torch.int32, device="cuda")
    hit = torch.full((3, n), float("nan"), device="cuda")
    B.ray(so, sd, vox, leaf, hit)
    torch.cuda.synchronize()
    vb = vox.cpu().numpy().view(np.uint32)
    hb = hit.cpu().numpy()
    for (t, f, cnt), s in zip(items, c4):
        sl = slice(f, f + cnt)
        vp, hp, lp = VT.ray_tree_intersection(t, s["o"], s["d"],
                                              want_leaf=True)
        vp = (vp.cpu().numpy() & 0xFFFFFFFF).astype(np.uint32)
        assert np.array_equal(vb[sl], vp)
        assert np.array_equal(leaf.cpu().numpy().view(np.uint32)[sl],
                              (lp.cpu().numpy() & 0xFFFFFFFF).astype(np.uint32))
        m = vp != MISS
        assert np.array_equal(hb[:, sl][:, m].view(np.uint32),
                              hp.cpu().numpy()[:, m].view(np.uint32))
        if len(s["dots"]) and cnt <= 8000:
            T = oracle.Tree(s["dots"])
            vo, ho = T.rays(s["o"], s["d"])
            assert np.array_equal(vb[sl], vo)


def test_batch_graph_replay(gpu):
    """Capture one batched launch in a graph; replay it three times with
    eager work in between; every replay matches the eager launch."""
    from voxvision_amd import batch, voxtrees as VT
    c4 = _small_c4()
    o, d, first = _concat(c4)
    trees = [VT.make_tree(s["dots"]) for s in c4]
    B = batch.Batch([(t, f, len(s["o"])) for t, f, s in zip(trees, first,
                                                            c4)])
    n = len(o)
    so, sd = batch.soa(o), batch.soa(d)
    ref = torch.empty(n, dtype=torch.int32, device="cuda")
    refh = torch.full((3, n), float("nan"), device="cuda")
    B.ray(so, sd, ref, None, refh)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    vox = torch.empty(n, dtype=torch.int32, device="cuda")
    hit = torch.full((3, n), float("nan"), device="cuda")
    with torch.cuda.stream(s):
        B.ray(so, sd, vox, None, hit, stream=s)  # set up the stream's workspace
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        B.ray(so, sd, vox, None, hit, stream=s)
    for _ in range(3):
        vox.fill_(7)
        hit.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(vox, ref)
        m = ref != -1
        assert torch.equal(hit[:, m], refh[:, m])


def test_batch_rejects_ray_ranges_that_would_wrap(gpu):
    """An item's first ray and ray count must each stay below 2^31, so
    that their sum cannot wrap around 2^64 into a small index."""
    import ctypes as C
    from voxvision_amd import _lib, voxtrees as VT
    G = VT.make_tree(np.zeros((1, 3), np.float32))
    L = _lib.load()
    for first, count in (((1 << 64) - 1, 2), ((1 << 63), (1 << 63) + 5),
                         (1 << 31, 0), (0, 1 << 31)):
        items = (_lib.BatchItem * 1)()
        items[0].tree = G.handle
        items[0].first = first
        items[0].count = count
        out = C.c_void_p()
        assert L.voxg_batch_create(1, items, C.byref(out)) == -34
        assert not out.value
