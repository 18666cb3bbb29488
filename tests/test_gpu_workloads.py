"""The batched and serving forms at their full workloads.

C4: 32 scenes of 32^3..512^3 (noise volumes, hollow shells, Menger
sponges; 93.9 M voxels), 64 K..8 M rays each (66.2 M in all, half camera
rays, half random), in one persistent launch: every ray equals the plain
per-scene launch (voxel id, leaf node, hit-point bits), every 61st ray of
each scene equals the oracle, and a graph replay of the launch equals it.

C5: 8 resident 256^3 CT trees; render requests from 320x240 to 3840x2160
in every quality and ball batches of 1 K to 1 M, submitted together from
three threads: each result equals the same request run alone through the
plain entry points (and, sampled, the oracle).
"""
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


def _u32(t):
    return t.view(torch.int32)


@pytest.fixture(scope="module")
def c4_full():
    from voxvision_amd import batch, scenes, voxtrees as VT
    c4 = scenes.config_c4()
    trees = [VT.make_tree(s["dots"]) for s in c4]
    o = np.concatenate([s["o"] for s in c4])
    d = np.concatenate([s["d"] for s in c4])
    first = np.cumsum([0] + [len(s["o"]) for s in c4])[:-1]
    items = [(t, int(f), len(s["o"])) for t, f, s in zip(trees, first, c4)]
    so, sd = batch.soa(o), batch.soa(d)
    return c4, trees, items, so, sd, len(o)


def test_c4_batched_equals_plain_and_oracle(gpu, oracle, c4_full):
    from voxvision_amd import batch, voxtrees as VT
    import ctypes as C
    from voxvision_amd import _lib
    c4, trees, items, so, sd, n = c4_full
    assert len(c4) == 32 and n == sum(len(s["o"]) for s in c4)
    assert n > 60_000_000
    B = batch.Batch(items)
    vox = torch.full((n,), -2, dtype=torch.int32, device="cuda")
    leaf = torch.full((n,), -2, dtype=torch.int32, device="cuda")
    hit = torch.full((3, n), float("nan"), device="cuda")
    B.ray(so, sd, vox, leaf, hit)
    # one plain launch per scene into separate outputs
    L = _lib.load()
    pv = torch.full((n,), -2, dtype=torch.int32, device="cuda")
    pl = torch.full((n,), -2, dtype=torch.int32, device="cuda")
    ph = torch.full((3, n), float("nan"), device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for t, f, cnt in items:
        off = f * 4
        _lib.check(L.voxg_ray_batch(
            t.handle, cnt, *[x.data_ptr() + off for x in so + sd],
            pv.data_ptr() + off, pl.data_ptr() + off,
            *[h.data_ptr() + off for h in ph], s), "voxg_ray_batch")
    assert L.voxg_synchronize(s) == 0
    assert torch.equal(vox, pv)
    assert torch.equal(leaf, pl)
    m = vox != -1
    assert not (vox == -2).any()
    assert torch.equal(_u32(hit)[:, m], _u32(ph)[:, m])
    del pv, pl, ph
    # a graph replay of the batched launch
    gs = torch.cuda.Stream()
    gv = torch.empty_like(vox)
    gh = torch.full_like(hit, float("nan"))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=gs):
        B.ray(so, sd, gv, None, gh, stream=gs)
    for _ in range(2):
        gv.fill_(-2)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(gv, vox)
        assert torch.equal(_u32(gh)[:, m], _u32(hit)[:, m])
    del g, gv, gh
    # every 61st ray of each scene against the oracle (trees built on
    # several threads; the oracle's build is single-threaded C)
    vh = vox.cpu().numpy().view(np.uint32)
    hh = hit.cpu().numpy()

    def check(k):
        sc = c4[k]
        _, f, cnt = items[k]
        sel = np.arange(0, cnt, 61)
        T = oracle.Tree(sc["dots"])
        vo, ho = T.rays(sc["o"][sel], sc["d"][sel], nthreads=2)
        T.close()
        g_v = vh[f + sel]
        ok = np.array_equal(g_v, vo)
        mm = vo != MISS
        ok = ok and np.array_equal(hh[:, f + sel][:, mm].view(np.uint32),
                                   ho[:, mm].view(np.uint32))
        return k, ok, len(sel)

    with ThreadPoolExecutor(max_workers=8) as ex:
        res = list(ex.map(check, range(len(c4))))
    bad = [k for k, ok, _ in res if not ok]
    assert not bad, bad
    assert sum(c for _, _, c in res) > 1_000_000


def _c5_trees():
    from voxvision_amd import scenes, voxtrees as VT
    with ThreadPoolExecutor(max_workers=8) as ex:
        dots = list(ex.map(
            lambda k: scenes.volume_to_dots(scenes.ct_volume(256,
                                                             seed=100 + k)),
            range(8)))
    return dots, [VT.make_tree(d) for d in dots]


def test_c5_serving_mix_equals_alone(gpu, oracle):
    from voxvision_amd import camera, scenes, server, voxrnd, voxtrees as VT
    dots, trees = _c5_trees()
    boxes = [scenes.bbox_of(d) for d in dots]
    rng = np.random.default_rng(41)
    sizes = [(320, 240), (640, 480), (1280, 720), (1920, 1080),
             (3840, 2160)]
    jobs = []
    for k in range(30):
        tid = k % 8
        bmin, bmax = boxes[tid]
        c = (bmin + bmax) / 2
        if k % 3 != 2:
            W, H = sizes[k % 5]
            a, el = rng.uniform(0, 2 * np.pi), rng.uniform(-0.4, 0.4)
            dist = float(rng.uniform(0.9, 1.5)) * \
                float(np.linalg.norm(bmax - bmin))
            pos = c + dist * np.array([np.cos(a) * np.cos(el),
                                       np.sin(a) * np.cos(el), np.sin(el)])
            cam = camera.make_camera(pos, c, np.tan(np.radians(30)), W, H)
            q = [0, 1, 2, 6, 10][k % 5]
            jobs.append(("r", tid, cam, W, H, q))
        else:
            nb = [1000, 30000, 300000, 1 << 20][(k // 3) % 4]
            bc, br = scenes.random_balls(nb, bmin, bmax, seed=200 + k)
            jobs.append(("b", tid, bc, br))
    assert any(j[0] == "r" and j[3] == 3840 for j in jobs)
    assert any(j[0] == "b" and len(j[2]) == 1 << 20 for j in jobs)
    # each request alone, through the plain entry points
    ref = []
    for j in jobs:
        if j[0] == "r":
            _, tid, cam, W, H, q = j
            out = voxrnd.render(trees[tid], [cam], W, H, quality=q,
                                want_colors=False)
            torch.cuda.synchronize()
            ref.append((out["voxel"].cpu(), out["hit"].cpu()))
        else:
            _, tid, bc, br = j
            ref.append(VT.tree_ball_collidep(trees[tid], bc, br).cpu())
    # the same requests, all in flight together, from three threads
    S = server.Server(streams=4)
    ids = [S.add_tree(t) for t in trees]
    results = {}
    lock = threading.Lock()

    def worker(part):
        for k in part:
            j = jobs[k]
            if j[0] == "r":
                tk, o = S.submit_render(ids[j[1]], j[2], j[3], j[4], j[5])
            else:
                tk, o = S.submit_balls(ids[j[1]], j[2], j[3])
            with lock:
                results[k] = (tk, o)

    th = [threading.Thread(target=worker, args=(list(range(p, 30, 3)),))
          for p in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    groups = set()
    for k, (tk, o) in sorted(results.items()):
        info = S.wait(tk, 120.0)
        assert info["status"] == 0, k
        groups.add(info["group"])
        if jobs[k][0] == "r":
            rv, rh = ref[k]
            assert torch.equal(o["voxel"], rv), k
            m = rv != -1
            assert torch.equal(o["hit"].view(3, -1)[:, m].view(torch.int32),
                               rh[:, m].view(torch.int32)), k
        else:
            assert torch.equal(o, ref[k]), k
    st = S.stats()
    S.close()
    assert st["completed"] == len(jobs)
    assert len(groups) < len(jobs)  # requests shared launch groups
    # sampled against the oracle: the first 320x240 render and 30K balls
    T = oracle.Tree(dots[0])
    _, tid, cam, W, H, q = jobs[0]
    _, vo, _, _ = T.render(cam, W, H, q, nthreads=8)
    assert np.array_equal(ref[0][0].numpy().view(np.uint32), vo)
    kb = next(k for k, j in enumerate(jobs)
              if j[0] == "b" and len(j[2]) <= 30000)
    Tb = oracle.Tree(dots[jobs[kb][1]])
    assert np.array_equal(ref[kb].numpy(), Tb.balls(jobs[kb][2], jobs[kb][3]))
