"""Benchmark of the voxvision ray-casting path on one MI355X.

Step = one pass of the hot path over one batch: 64 orbit poses of the
256^3 CT-like scene (1.34 M voxels) at 1920x1080, i.e. 132,710,400 primary
rays traced with the vox_render per-pixel loop in BEST mode, writing the
hit voxel id and hit point of every pixel.  One kernel launch per step,
inputs (tree, cameras) resident in HBM before the timed region.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--quick]

Prints one JSON line.  `roofline` follows the algorithmic-bytes model
(40 B per ray + 32 B per node visit + 12 B per voxel test, visits and
tests counted by the counting variant of the same kernel); `cpu_baseline`
times the CPU oracle (a C restatement of the reference, one task per 4x4
block on a thread pool) on a bounded sample of the same frames.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Mrays/s coherent+incoherent on 256³ scene @1080p; "
          "batched vs plain; p99 latency")
HBM_PEAK_GBS = 8000.0
BYTES_PER_RAY, BYTES_PER_VISIT, BYTES_PER_TEST = 40, 32, 12
BYTES_PER_BALL = 17  # centre 12 B + radius 4 B in, 1 B out
QUALITY_NAMES = {0: "best", 1: "fast", 2: "adaptive",
                 6: "adaptive_ray_merge", 10: "adaptive_ray_merge_accurate"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(dots, cams, W, H, budget_s):
    """Oracle render (BEST) on all usable host cores over consecutive poses
    until `budget_s` seconds of work are done (at least one frame)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    T = O.Tree(dots)
    cores = O.cpu_count()
    frames, rays = 0, 0
    t0 = time.perf_counter()
    while True:
        T.render(cams[frames % len(cams)], W, H, O.QUALITY_BEST,
                 nthreads=cores)
        frames += 1
        rays += W * H
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    T.render(cams[0], W, H, O.QUALITY_BEST, nthreads=1)
    dt1 = time.perf_counter() - t1
    # the other timings of the CPU restatement: ADAPTIVE frames, looped ray
    # queries (incoherent, as in the 16 M-ray workload) and ball queries,
    # on all cores and on one, each on a bounded sample
    import numpy as np
    from voxvision_amd import scenes
    other = {}
    t1 = time.perf_counter()
    T.render(cams[1 % len(cams)], W, H, O.QUALITY_ADAPTIVE, nthreads=cores)
    other["render_adaptive_mrays_s"] = round(
        W * H / (time.perf_counter() - t1) / 1e6, 2)
    bmin = dots.min(0).astype(np.float32)
    bmax = (dots.max(0) + 1).astype(np.float32)
    o, d = scenes.random_rays(1 << 20, bmin, bmax, seed=11)
    for name, n, th in (("rays_incoherent", 1 << 20, cores),
                        ("rays_incoherent_1core", 1 << 17, 1)):
        t1 = time.perf_counter()
        T.rays(o[:n], d[:n], nthreads=th)
        other[name + "_mrays_s"] = round(n / (time.perf_counter() - t1) / 1e6,
                                         2)
    c, r = scenes.random_balls(1 << 20, bmin, bmax, seed=13)
    for name, n, th in (("balls", 1 << 20, cores), ("balls_1core", 1 << 17, 1)):
        t1 = time.perf_counter()
        T.balls(c[:n], r[:n], nthreads=th)
        other[name + "_mqueries_s"] = round(
            n / (time.perf_counter() - t1) / 1e6, 2)
    T.close()
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": cores,
            "kind": "port",
            "build": "gcc -O3 -march=x86-64-v3 -ffp-contract=off "
                     "(oracle/Makefile: portable AVX2+FMA, so the library "
                     "built here runs on the GPU host; the reference's "
                     "RELEASE build is clang -O3 -march=native)",
            "sample": "%d of the 64 poses at %dx%d (%.1f s, %d threads, "
                      "oracle vox_render loop, BEST); 1 core: %.2f Mrays/s"
                      % (frames, W, H, dt, cores, W * H / dt1 / 1e6),
            "other": other}


def batched_leg(ev0, ev1, stream, log, fixed=False, split=False):
    """32 uneven scenes (edges 32..512, noise / shells / Menger sponges,
    64K..8M rays each, half camera rays, half random): one plain
    voxg_ray_batch launch per scene versus one persistent batched launch
    (and its HIP-graph replay), each form at its own autotuned
    configuration (fixed: at the library's default configurations, for
    profiling).  split: the camera halves of the 32 scenes and their random
    halves also timed alone, as batched launches."""
    import ctypes as C
    import numpy as np
    import torch
    from voxvision_amd import _lib, autotune, batch, scenes, voxtrees as VT
    L = _lib.load()
    t0 = time.perf_counter()
    c4 = scenes.config_c4()
    trees = [VT.make_tree(s["dots"]) for s in c4]
    log("[bench] batched leg: 32 scenes ready in %.1f s"
        % (time.perf_counter() - t0))
    o = np.concatenate([s["o"] for s in c4])
    d = np.concatenate([s["d"] for s in c4])
    first = np.cumsum([0] + [len(s["o"]) for s in c4])[:-1]
    counts = [len(s["o"]) for s in c4]
    n = len(o)
    so, sd = batch.soa(o), batch.soa(d)
    vox = torch.empty(n, dtype=torch.int32, device="cuda")
    hit = [torch.empty(n, device="cuda") for _ in range(3)]
    sp = C.c_void_p(stream.cuda_stream)

    def plain_runner(cfg):
        def run():
            for t, f, cnt in zip(trees, first, counts):
                off = int(f) * 4
                _lib.check(L.voxg_ray_batch(
                    t.handle, cnt, *[x.data_ptr() + off for x in so + sd],
                    vox.data_ptr() + off, None,
                    *[h.data_ptr() + off for h in hit], sp))
        return run

    def batch_runner(cfg):
        B = batch.Batch(list(zip(trees, first, counts)))
        return lambda: B.ray(so, sd, vox, None, hit)

    space = ["chunk", "waves_per_cu", "stack_lds", "refill", "leaf_batch",
             "finish_batch"]
    if fixed:
        pcfg, bcfg = VT.get_launch_config("rays"), VT.get_launch_config("batch")
        pms = autotune.time_config(plain_runner, pcfg, 3, "rays")
        bms = autotune.time_config(batch_runner, bcfg, 3, "batch")
    else:
        pcfg, pms, _ = autotune.autotune(plain_runner, params=space, reps=2,
                                         kind="rays")
        bcfg, bms, _ = autotune.autotune(batch_runner, params=space, reps=2,
                                         kind="batch")
    B = batch.Batch(list(zip(trees, first, counts)))
    g = torch.cuda.CUDAGraph()
    gs = torch.cuda.Stream()
    with torch.cuda.stream(gs):
        B.ray(so, sd, vox, None, hit, stream=gs)
    gs.synchronize()
    with torch.cuda.graph(g, stream=gs):
        B.ray(so, sd, vox, None, hit, stream=gs)
    g.replay()
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(5):
        g.replay()
    ev1.record(stream)
    torch.cuda.synchronize()
    gms = ev0.elapsed_time(ev1) / 5
    log("[bench] batched leg: plain %.2f ms, batched %.2f ms, graph %.2f ms"
        % (pms, bms, gms))
    halves = {}
    if split:
        for name, part in (
                ("camera", lambda s: (0, s["camera_pixels"])),
                ("random", lambda s: (s["camera_pixels"],
                                      len(s["o"]) - s["camera_pixels"]))):
            items = [(t, int(f) + part(s)[0], part(s)[1])
                     for t, f, s in zip(trees, first, c4)]
            Bh = batch.Batch(items)
            hms = autotune.time_config(
                lambda cfg, Bh=Bh: (lambda: Bh.ray(so, sd, vox, None, hit)),
                bcfg, 3, "batch")
            nr = sum(c for _, _, c in items)
            halves[name] = {"rays": nr, "ms": round(hms, 3),
                            "mrays_s": round(nr / hms / 1e3, 1)}
        log("[bench] batched halves: %s" % json.dumps(halves))
    return {"scenes": len(c4), "rays": n,
            "voxels": int(sum(len(s["dots"]) for s in c4)),
            "plain_mrays_s": round(n / pms / 1e3, 1),
            "batched_mrays_s": round(n / bms / 1e3, 1),
            "graph_mrays_s": round(n / gms / 1e3, 1),
            "speedup_batched_vs_plain": round(pms / min(bms, gms), 3),
            "plain_config": pcfg, "batched_config": bcfg,
            **({"halves": halves} if halves else {})}


def aux_leg(G, dots, W, H, nf, ev0, ev1, stream, log):
    """The calls on either side of the path: vox_redraw's square-to-row
    copy of the headline frames (with the clear), vox_cd_collide for 4 M
    bodies, and vox_read_raw_data's selection over a 512^3 one-byte
    volume (8% selected).  HBM bytes: redraw 12 B per pixel (read 4, write
    4 + clear 4); selection 1 B per sample read + 12 B per selected voxel
    written (the compulsory bytes; the count pass also writes and the
    write pass reads one selection bit per sample).  The selection is
    timed as the Python call makes it: count, the count to the host, exact
    allocation, corners.  Tree edits: 20,000 alternating inserts and deletes on
    the host tree of the headline scene (vox_insert_voxel /
    vox_delete_voxel), then one commit of the result as a new device
    version (flatten + upload)."""
    import numpy as np
    import torch
    from voxvision_amd import voxrnd, voxtrees as VT

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(reps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / reps

    res = {}
    npx = nf * W * H
    sq = torch.randint(0, 1 << 30, (npx,), dtype=torch.int32, device="cuda")
    px = torch.empty((nf, H, W), dtype=torch.int32, device="cuda")
    ms = timed(lambda: voxrnd.redraw(sq, W, H, pixels=px, clear=True))
    res["redraw_64x1080p_ms"] = round(ms, 4)
    res["redraw_gb_s"] = round(12.0 * npx / ms / 1e6, 1)
    del sq, px
    bmin, bmax = np.array(G.info["bbox_min"]), np.array(G.info["bbox_max"])
    nb = 1 << 22
    g = torch.Generator(device="cuda").manual_seed(17)
    lo = torch.tensor(bmin, device="cuda", dtype=torch.float32)
    ext = torch.tensor(bmax - bmin, device="cuda", dtype=torch.float32)
    pos0 = [lo[k] + ext[k] * torch.rand(nb, device="cuda", generator=g)
            for k in range(3)]
    last0 = [p + 1.0 for p in pos0]
    rad = 0.5 + 7.5 * torch.rand(nb, device="cuda", generator=g)
    P = [p.clone() for p in pos0]
    Lst = [p.clone() for p in last0]
    hit = torch.empty(nb, dtype=torch.uint8, device="cuda")

    def reset():  # every body moved again, so every body is tested
        for a, b in zip(P, pos0):
            a.copy_(b)
        for a, b in zip(Lst, last0):
            a.copy_(b)

    def collide():
        reset()
        voxrnd.cd_collide(G, P, Lst, rad, collided=hit)
    ms = timed(collide) - timed(reset)
    res["cd_collide_4M_ms"] = round(ms, 4)
    res["cd_collide_mbodies_s"] = round(nb / ms / 1e3, 1)
    res["cd_collide_moved_back"] = round(float(hit.float().mean()), 4)
    dim = (512, 512, 512)
    vol = torch.randint(0, 256, dim, dtype=torch.uint8, device="cuda",
                        generator=g)
    flat = vol.reshape(-1)
    xyz = [None]

    def select():
        xyz[0] = VT.raw_select(flat, dim, 1, lo=235, hi=256)
    ms = timed(select, reps=3)
    m = xyz[0].shape[0]
    res["raw_select_512cube_ms"] = round(ms, 4)
    res["raw_select_gsamples_s"] = round(flat.numel() / ms / 1e6, 1)
    res["raw_select_selected"] = int(m)
    res["raw_select_gb_s"] = round((1.0 * flat.numel() + 12.0 * m) / ms / 1e6,
                                   1)
    # the same selection in one voxg_raw_select call into an output of known
    # capacity: the device passes alone, no host round trip
    import ctypes as C
    from voxvision_amd import _lib
    L = _lib.load()
    d3 = (C.c_uint * 3)(*dim)
    one = (C.c_float * 3)(1.0, 1.0, 1.0)
    cnt = torch.empty(1, dtype=torch.int64, device="cuda")
    out = torch.empty((m, 3), dtype=torch.float32, device="cuda")
    sp = C.c_void_p(stream.cuda_stream)

    def select_one():
        _lib.check(L.voxg_raw_select(flat.data_ptr(), d3, 1, 235, 256, one,
                                     out.data_ptr(), m, cnt.data_ptr(), sp))
    ms1 = timed(select_one, reps=5)
    assert int(cnt.item()) == m and torch.equal(out, xyz[0])
    res["raw_select_one_call_ms"] = round(ms1, 4)
    res["raw_select_one_call_gb_s"] = round(
        (1.0 * flat.numel() + 12.0 * m) / ms1 / 1e6, 1)
    # tree edits on the headline scene: host-side vox_insert_voxel /
    # vox_delete_voxel, then one commit (flatten + upload of a version)
    from voxvision_amd.edit import EditTree
    t0 = time.perf_counter()
    E = EditTree(dots)
    res["edit_tree_build_s"] = round(time.perf_counter() - t0, 3)
    rng = np.random.default_rng(23)
    pts = np.floor(rng.uniform(bmin, bmax, (20000, 3))).astype(np.float32)
    t0 = time.perf_counter()
    changed = 0
    for i, p in enumerate(pts):
        changed += E.insert(p) if i % 2 == 0 else E.delete(p)
    dt = time.perf_counter() - t0
    res["edits_per_s"] = round(len(pts) / dt, 1)
    res["edits_changed"] = int(changed)
    t0 = time.perf_counter()
    G2 = E.commit()
    torch.cuda.synchronize()
    res["commit_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    res["commit_device_bytes"] = int(G2.info["device_bytes"])
    del G2, E
    log("[bench] around the path: %s" % json.dumps(res))
    return res


def serving_leg(log, seconds, rate, seed=31):
    """Serving: 8 resident 256^3 trees (CT-like, different seeds); a
    seeded Poisson stream of requests, 60% renders (320x240 .. 3840x2160,
    random resident tree, seeded pose, BEST/FAST/ADAPTIVE) and 40%
    collision batches (1K..1M balls, radii U[0.5, 8]); latency is server
    time from submission to results in host memory."""
    import queue
    import threading
    import numpy as np
    import torch
    from voxvision_amd import camera, scenes, server, voxtrees as VT
    t0 = time.perf_counter()
    trees, boxes = [], []
    for k in range(8):
        dots = scenes.volume_to_dots(scenes.ct_volume(256, seed=100 + k))
        trees.append(VT.make_tree(dots))
        boxes.append(scenes.bbox_of(dots))
    S = server.Server(streams=4)
    ids = [S.add_tree(t) for t in trees]
    rng = np.random.default_rng(seed)
    sizes = [(320, 240), (640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
    wts = [0.3, 0.25, 0.2, 0.15, 0.1]
    pools = {sz: queue.Queue() for sz in sizes}
    for sz, cnt in zip(sizes, (48, 48, 48, 32, 24)):
        for _ in range(cnt):
            pools[sz].put({"voxel": torch.empty(sz[0] * sz[1],
                                                dtype=torch.int32,
                                                pin_memory=True)})
    bc, br = scenes.random_balls(1 << 20, *boxes[0], seed=seed)
    plan, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        k = int(rng.integers(0, 8))
        if rng.random() < 0.6:
            sz = sizes[int(rng.choice(len(sizes), p=wts))]
            bmin, bmax = boxes[k]
            c = (bmin + bmax) / 2
            a, el = rng.uniform(0, 2 * np.pi), rng.uniform(-0.4, 0.4)
            dist = 1.5 * float(np.linalg.norm(bmax - bmin))
            pos = c + dist * np.array([np.cos(a) * np.cos(el),
                                       np.sin(a) * np.cos(el), np.sin(el)])
            cam = camera.make_camera(pos, c, np.tan(np.radians(30)), *sz)
            q = int(rng.choice([0, 1, 2, 6, 10]))
            plan.append((t, "r", ids[k], cam, sz, q))
        else:
            n = int(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
            off = int(rng.integers(0, (1 << 20) - n))
            plan.append((t, "b", ids[k], n, off))
    log("[bench] serving leg: %d requests planned over %.1f s, setup %.1f s"
        % (len(plan), seconds, time.perf_counter() - t0))
    def run(paced):
        """Submit the plan (at its arrival times when paced, else as fast
        as the result buffers come back) and collect every result."""
        tickets = queue.Queue()
        failure = []

        def collect(out):
            while True:
                item = tickets.get()
                if item is None:
                    return
                tk, kind, px, buf, sz = item
                try:
                    info = S.wait(tk, 120.0)
                except Exception as e:  # a request that did not complete
                    failure.append("%s request: %s" % (kind, e))
                    return
                if buf is not None:
                    pools[sz].put(buf)
                out.append((kind, px, info))

        def buffer(sz):
            while True:
                try:
                    return pools[sz].get(timeout=1.0)
                except queue.Empty:
                    if failure:
                        return None

        done = []
        th = threading.Thread(target=collect, args=(done,))
        th.start()
        start = time.perf_counter()
        for item in plan:
            if failure:
                break
            delay = item[0] - (time.perf_counter() - start)
            if paced and delay > 0:
                time.sleep(delay)
            if item[1] == "r":
                _, _, tid, cam, sz, q = item
                buf = buffer(sz)
                if buf is None:
                    break
                tk, _ = S.submit_render(tid, cam, sz[0], sz[1], q, out=buf)
                tickets.put((tk, "r", sz[0] * sz[1], buf, sz))
            else:
                _, _, tid, n, off = item
                tk, _ = S.submit_balls(tid, bc[off:off + n], br[off:off + n])
                tickets.put((tk, "b", n, None, None))
        tickets.put(None)
        th.join()
        if failure:
            raise RuntimeError(failure[0])
        lat = np.array([i["t_done"] - i["t_submit"] for _, _, i in done]) * 1e3
        lr = np.array([i["t_done"] - i["t_submit"] for k, _, i in done
                       if k == "r"]) * 1e3
        lb = np.array([i["t_done"] - i["t_submit"] for k, _, i in done
                       if k == "b"]) * 1e3
        first = min(i["t_submit"] for _, _, i in done)
        last = max(i["t_done"] for _, _, i in done)
        span = last - first
        px = sum(p for k, p, _ in done if k == "r")
        nb = sum(p for k, p, _ in done if k == "b")
        sizes_g = [i["group_size"] for _, _, i in done]
        # where the time goes: waiting for admission (the scheduler thread
        # forming and launching groups) and service (launch to results)
        wait = np.array([i["t_admit"] - i["t_submit"] for _, _, i in done])
        serve = np.array([i["t_done"] - i["t_admit"] for _, _, i in done])
        l4k = np.array([i["t_done"] - i["t_submit"] for k, p, i in done
                        if k == "r" and p >= 3840 * 2160]) * 1e3
        # the same over the first and the last half second of submissions
        t_sub = np.array([i["t_submit"] for _, _, i in done])
        t_first = t_sub.min()
        early = lat[t_sub - t_first < 0.5]
        late = lat[t_sub - t_first >= t_sub.max() - t_first - 0.5]
        return {"requests": len(done), "span_s": round(span, 3),
                "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "render_p50_ms": round(float(np.percentile(lr, 50)), 3),
                "render_p99_ms": round(float(np.percentile(lr, 99)), 3),
                "balls_p50_ms": round(float(np.percentile(lb, 50)), 3),
                "balls_p99_ms": round(float(np.percentile(lb, 99)), 3),
                "sustained_mrays_s": round(px / span / 1e6, 1),
                "sustained_mballs_s": round(nb / span / 1e6, 1),
                "requests_per_s": round(len(done) / span, 1),
                "mean_group_size": round(float(np.mean(sizes_g)), 2),
                "admit_wait_p50_ms": round(float(np.percentile(wait, 50)) * 1e3, 3),
                "admit_wait_p99_ms": round(float(np.percentile(wait, 99)) * 1e3, 3),
                "service_p50_ms": round(float(np.percentile(serve, 50)) * 1e3, 3),
                "service_p99_ms": round(float(np.percentile(serve, 99)) * 1e3, 3),
                "render_3840x2160_p50_ms": round(float(np.percentile(l4k, 50)), 3)
                if len(l4k) else None,
                "first_half_second_p99_ms": round(float(np.percentile(early, 99)), 3),
                "last_half_second_p99_ms": round(float(np.percentile(late, 99)), 3)}

    # capacity: the requests submitted without pacing (in flight at most as
    # many as there are result buffers, 200).  It runs first, so that every
    # stream has grown its buffers for the largest frames before the paced
    # run measures latency (growing one waits for the stream: a cold
    # server's first half second has a p99 of about 22 ms)
    try:
        sat = run(paced=False)
        st = S.stats()
        sat["peak_groups_inflight"] = st["peak_groups_inflight"]
        sat["peak_inflight"] = st["peak_inflight"]
        res = run(paced=True)
    except RuntimeError as e:
        # recorded, not hidden: the server is left as it is (closing it
        # would wait for the request that did not complete)
        log("[bench] serving leg failed: %s" % e)
        return {"error": str(e)}
    res["offered_rate_per_s"] = rate
    res["groups"] = S.stats()["groups"] - st["groups"]
    res["saturation"] = sat
    S.close()
    log("[bench] serving leg: %s" % json.dumps(res))
    return res


def c3_leg(G, bmin, bmax, ev0, ev1, stream, log, tune=False):
    """The incoherent workload on the headline scene: 16 M rays with
    origins uniform in the box grown by 10% of its extent per side and
    N(0, I) directions (voxg_ray_batch), and 4 M ball queries with radii
    U[0.5, 8] (voxg_ball_batch).  Each leg gets the roofline of its kernel:
    algorithmic bytes = 40 B per ray (17 B per ball) + 32 B per node visit
    + 12 B per voxel test, visits and tests counted by the counting launch
    on the same queries, over the launch's event time."""
    import ctypes as C
    import numpy as np
    import torch
    from voxvision_amd import _lib, scenes
    L = _lib.load()
    sp = C.c_void_p(stream.cuda_stream)
    res = {}
    n = 1 << 24
    o, d = scenes.random_rays(n, bmin, bmax, seed=11)
    so = [torch.from_numpy(np.ascontiguousarray(o[:, k])).cuda()
          for k in range(3)]
    sd = [torch.from_numpy(np.ascontiguousarray(d[:, k])).cuda()
          for k in range(3)]
    vo = torch.empty(n, dtype=torch.int32, device="cuda")
    ho = [torch.empty(n, device="cuda") for _ in range(3)]

    def inc():
        _lib.check(L.voxg_ray_batch(
            G.handle, n, *[t.data_ptr() for t in so + sd], vo.data_ptr(),
            None, *[t.data_ptr() for t in ho], sp), "voxg_ray_batch")
    if tune:
        from voxvision_amd import autotune
        icfg, _, _ = autotune.autotune(
            lambda cfg: inc, params=["leaf_batch", "finish_batch", "chunk",
                                     "refill"], reps=2, kind="rays")
        res["incoherent_launch"] = icfg
    inc()
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(5):
        inc()
    ev1.record(stream)
    torch.cuda.synchronize()
    ims = ev0.elapsed_time(ev1) / 5
    cnt = _lib.Counts()
    _lib.check(L.voxg_ray_batch_count(
        G.handle, n, *[t.data_ptr() for t in so + sd], C.byref(cnt), sp),
        "voxg_ray_batch_count")
    alg = (BYTES_PER_RAY * n + BYTES_PER_VISIT * cnt.node_visits +
           BYTES_PER_TEST * cnt.voxel_tests)
    tr = load_traffic("traffic_ct256_c3.json")
    res["incoherent_16M"] = {
        "ms": round(ims, 4), "mrays_s": round(n / ims / 1e3, 1),
        "roofline": {
            "bound": "hbm", "kernel": "rti_kernel<4, ChunkWork>",
            "achieved": round(alg / (ims / 1e3) / 1e9, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(alg / (ims / 1e3) / 1e9 / HBM_PEAK_GBS, 5),
            "traffic": tr.get("rays_hbm_bytes_per_launch") if tr else None,
            "algorithmic_bytes_per_launch": alg,
            "node_visits_per_ray": round(cnt.node_visits / n, 4),
            "voxel_tests_per_ray": round(cnt.voxel_tests / n, 4),
            "stack_spills": int(cnt.stack_spills)}}
    del so, sd, vo, ho
    nb = 1 << 22
    cb, rb = scenes.random_balls(nb, bmin, bmax, seed=13)
    cbt = [torch.from_numpy(np.ascontiguousarray(cb[:, k])).cuda()
           for k in range(3)]
    rbt = torch.from_numpy(rb).cuda()
    bo = torch.empty(nb, dtype=torch.uint8, device="cuda")

    def balls():
        _lib.check(L.voxg_ball_batch(
            G.handle, nb, *[t.data_ptr() for t in cbt], rbt.data_ptr(),
            bo.data_ptr(), sp), "voxg_ball_batch")
    balls()
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(5):
        balls()
    ev1.record(stream)
    torch.cuda.synchronize()
    bms = ev0.elapsed_time(ev1) / 5
    _lib.check(L.voxg_ball_batch_count(
        G.handle, nb, *[t.data_ptr() for t in cbt], rbt.data_ptr(),
        C.byref(cnt), sp), "voxg_ball_batch_count")
    alg = (BYTES_PER_BALL * nb + BYTES_PER_VISIT * cnt.node_visits +
           BYTES_PER_TEST * cnt.voxel_tests)
    res["balls_4M"] = {
        "ms": round(bms, 4), "mqueries_s": round(nb / bms / 1e3, 1),
        "hit_rate": round(float(bo.float().mean()), 4),
        "roofline": {
            "bound": "hbm", "kernel": "ball_kernel",
            "achieved": round(alg / (bms / 1e3) / 1e9, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(alg / (bms / 1e3) / 1e9 / HBM_PEAK_GBS, 5),
            "traffic": tr.get("balls_hbm_bytes_per_launch") if tr else None,
            "algorithmic_bytes_per_launch": alg,
            "node_visits_per_ball": round(cnt.node_visits / nb, 4),
            "voxel_tests_per_ball": round(cnt.voxel_tests / nb, 4)}}
    log("[bench] incoherent leg: %s" % json.dumps(res))
    return res


def fp64_accuracy():
    """The hit accuracy of the fp32 search against FP64, as recorded by
    tests/golden/make_fp64_accuracy.py (the GPU's results equal the CPU
    restatement's bit for bit, which the GPU tests check)."""
    p = os.path.join(ROOT, "tests", "golden", "fp64_accuracy.json")
    if not os.path.exists(p):
        return None
    r = json.load(open(p))
    return {k: {f: v[f] for f in ("rays", "tie_fraction",
                                  "voxel_agree_outside_ties",
                                  "hit_error_voxels_median",
                                  "hit_error_voxels_max",
                                  "hit_error_rel_max")}
            for k, v in r.items() if isinstance(v, dict)}


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p))
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--quick", action="store_true",
                    help="skip the incoherent and ball extras")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 32-scene batched-vs-plain leg")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving (request stream) leg")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-rate", type=float, default=1000.0,
                    help="Poisson arrival rate of the serving leg (req/s)")
    ap.add_argument("--quality", type=int, default=0,
                    help="render quality of the timed step (0 BEST, the "
                         "headline; 1 FAST, 2 ADAPTIVE, 6 / 10 with ray "
                         "merging), for profiling the other modes")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the headline render (and the incoherent "
                         "rays) before timing; by default the library's "
                         "default launch configurations, which the "
                         "committed profiles were taken at, are timed")
    ap.add_argument("--no-tune", action="store_true",
                    help="the default (kept for older command lines)")
    ap.add_argument("--legs", default="modes,c3,aux",
                    help="extra legs besides the headline step, comma "
                         "separated: modes (other render qualities), c3 "
                         "(16 M incoherent rays, 4 M balls), aux (calls "
                         "around the path); --quick runs none")
    ap.add_argument("--batched-fixed", action="store_true",
                    help="batched leg at the default launch configurations "
                         "instead of autotuned ones (for profiling)")
    ap.add_argument("--batched-split", action="store_true",
                    help="batched leg: also time the camera and the random "
                         "halves of the 32 scenes alone")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="run a plain device copy of this size first (the "
                         "reference kernel for HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the ray-casting path shards rays "
              "across GPUs with no communication, so only --gpus 1 is run "
              "here", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from voxvision_amd import _lib, scenes, voxrnd, voxtrees as VT

    L = _lib.load()
    _lib.check(L.voxg_init(0), "voxg_init")
    default_cfgs = VT.get_launch_configs()
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()

    log("[bench] generating the 256^3 scene and 64 poses")
    c2 = scenes.config_c2()
    dots, W, H = c2["dots"], c2["W"], c2["H"]
    G = VT.make_tree(dots)
    cams = voxrnd.cameras_tensor(c2["cameras"])
    nf = cams.shape[0]
    n = nf * W * H
    out = {"voxel": torch.empty(n, dtype=torch.int32, device="cuda"),
           "hit": torch.empty((3, n), dtype=torch.float32, device="cuda")}
    log("[bench] tree: %s" % json.dumps(G.info))

    def step():
        voxrnd.render(G, cams, W, H, quality=args.quality, out=out)

    tuned = None
    if args.tune:
        from voxvision_amd import autotune
        cfg, best_ms, hist = autotune.autotune(
            lambda cfg: step, params=["stack_lds", "waves_per_cu", "refill",
                                      "block", "leaf_batch", "finish_batch",
                                      "split_best"],
            reps=2, kind="render_best")
        tuned = {"config": cfg, "ms": round(best_ms, 3),
                 "candidates": len(hist)}
        log("[bench] autotuned render config %s (%.3f ms)" % (cfg, best_ms))

    if args.copy_calib_mib:
        nb = args.copy_calib_mib << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy calibration: %d bytes read, %d written" % (nb, nb))
        del src, dst

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    ms = ms_total / args.steps
    value = n / (ms / 1e3) / 1e6
    log("[bench] %d steps: %.3f ms/step (events), wall %.3f s"
        % (args.steps, ms, wall))

    hits = int((out["voxel"] != -1).sum().item())
    cnt = voxrnd.render_counts(G, cams, W, H, quality=args.quality)
    alg_bytes = (BYTES_PER_RAY * n + BYTES_PER_VISIT * cnt["node_visits"] +
                 BYTES_PER_TEST * cnt["voxel_tests"])
    achieved = alg_bytes / (ms / 1e3) / 1e9
    # the committed HBM traffic was measured on the BEST render
    tr = load_traffic("traffic_ct256_render.json") if args.quality == 0 \
        else None
    # one render call = the block pre-pass (block_cull_kernel) + the
    # traversal (rti_kernel); both are timed together by the events above
    roofline = {"bound": "hbm", "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": tr.get("hbm_bytes_per_launch") if tr else None,
                "algorithmic_bytes_per_launch": alg_bytes,
                # 40 B of I/O per ray alone (24 B in, 16 B out)
                "compulsory_gb_s": round(BYTES_PER_RAY * n / (ms / 1e3) / 1e9,
                                         2),
                "timed": "render call: block_cull_kernel + rti_kernel",
                "node_visits_per_ray": round(cnt["node_visits"] / n, 4),
                "voxel_tests_per_ray": round(cnt["voxel_tests"] / n, 4),
                # stack entries that left the LDS ring for the spill
                # buffer (20 B out and back each), counting launch
                "stack_spills": int(cnt["stack_spills"])}

    extra = {"hit_rate": round(hits / n, 4)}
    legs = set() if args.quick else set(x for x in args.legs.split(",") if x)
    if "modes" in legs:
        # the other quality modes on the same frames and outputs (default
        # configurations of their kinds); ADAPTIVE is the reference's
        # default (renderer.c:78)
        modes = {}
        for name, q in (("adaptive", 2), ("fast", 1),
                        ("adaptive_ray_merge", 2 | 4),
                        ("adaptive_ray_merge_accurate", 2 | 8)):
            voxrnd.render(G, cams, W, H, quality=q, out=out)
            torch.cuda.synchronize()
            ev0.record(stream)
            for _ in range(3):
                voxrnd.render(G, cams, W, H, quality=q, out=out)
            ev1.record(stream)
            torch.cuda.synchronize()
            qms = ev0.elapsed_time(ev1) / 3
            modes[name] = {"ms": round(qms, 3),
                           "mrays_s": round(n / qms / 1e3, 1)}
        extra["render_modes"] = modes
        log("[bench] other quality modes: %s" % json.dumps(modes))
    if "c3" in legs:
        bmin, bmax = c2["bbox"]
        c3 = c3_leg(G, bmin, bmax, ev0, ev1, stream, log, tune=args.tune)
        extra.update(c3)
        extra["incoherent_16M_mrays_s"] = c3["incoherent_16M"]["mrays_s"]
        extra["balls_4M_mqueries_s"] = c3["balls_4M"]["mqueries_s"]
    if "aux" in legs:
        extra["around_the_path"] = aux_leg(G, dots, W, H, nf, ev0, ev1,
                                           stream, log)
    acc = fp64_accuracy()
    if acc:
        extra["fp64_accuracy"] = acc
    if not args.no_batched:
        extra["batched"] = batched_leg(ev0, ev1, stream, log,
                                       fixed=args.batched_fixed,
                                       split=args.batched_split)
    if not args.no_serving:
        VT.set_launch_configs(default_cfgs)
        extra["serving"] = serving_leg(log, args.serve_seconds,
                                       args.serve_rate)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (%.0f s budget)" % args.cpu_seconds)
        cpu = cpu_baseline(dots, c2["cameras"], W, H, args.cpu_seconds)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "ct256_orbit64_1080p_" + QUALITY_NAMES.get(
                       args.quality, "q%d" % args.quality),
                   "scene": "256^3 4-octave value-noise CT shell+interior, "
                            "8%% occupancy, %d voxels" % len(dots),
                   "frames": nf, "W": W, "H": H, "rays_per_step": n,
                   "outputs": "voxel id + hit point per pixel"},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }
    line["config"]["launch"] = (tuned["config"] if tuned
                                else default_cfgs["render_best"])
    print(json.dumps(line), flush=True)
    if "error" in extra.get("serving", {}):
        # a serving request never completed: end without the runtime's
        # teardown, which would wait for it
        sys.stderr.flush()
        os._exit(1)


if __name__ == "__main__":
    main()
