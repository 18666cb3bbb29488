"""Failure handling on the GPU, every failure forced from the host so that
no kernel ever faults, traps or hangs:

* a render that fails after its block pre-pass (a failure injected at a
  host checkpoint, where a launch could fail) must leave the stream so
  that the next render is exact: its queue counts start from zero;
* a traversal whose stack bound is violated (the bound capped from the
  host below the tree's) must report VOXG_EIO at the next synchronising
  call, once, and launches after the cap is lifted must be exact again.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF
VOXG_EIO = -5


def _np_u32(t):
    return (t.cpu().numpy().astype(np.int64) & 0xFFFFFFFF).astype(np.uint32)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_render_after_a_failed_render_is_exact(gpu, oracle):
    from voxvision_amd import _lib, scenes, voxrnd, voxtrees as VT
    L = gpu
    c1 = scenes.config_c1()
    T = oracle.Tree(c1["dots"])
    G = VT.make_tree(c1["dots"])
    W, H, cam = c1["W"], c1["H"], c1["camera"]
    cams = voxrnd.cameras_tensor([cam])
    n = W * H
    vox = torch.empty(n, dtype=torch.int32, device="cuda")
    hit = torch.empty((3, n), dtype=torch.float32, device="cuda")
    modes = (oracle.QUALITY_BEST, oracle.QUALITY_FAST,
             oracle.QUALITY_ADAPTIVE,
             oracle.QUALITY_ADAPTIVE | oracle.QUALITY_RAY_MERGE)
    failed = 0
    for q in modes:
        _, vo, ho, _ = T.render(cam, W, H, q)
        m = vo != MISS
        for cp in (1, 2, 3, 4):
            p = _lib.RenderParams()
            p.quality = q
            _lib.check(L.voxg_test_fail_at(cp))
            rc = L.voxg_render(G.handle, 1, cams.data_ptr(), W, H,
                               C.byref(p), None, vox.data_ptr(),
                               hit.data_ptr(), _stream())
            _lib.check(L.voxg_test_fail_at(0))
            assert rc in (0, VOXG_EIO), rc
            failed += rc == VOXG_EIO
            hit.fill_(float("nan"))
            out = voxrnd.render(G, cams, W, H, quality=q, want_colors=False,
                                out={"voxel": vox, "hit": hit})
            torch.cuda.synchronize()
            assert np.array_equal(_np_u32(out["voxel"]), vo), (q, cp)
            hg = out["hit"].cpu().numpy()
            assert np.array_equal(hg[:, m].view(np.uint32),
                                  ho[:, m].view(np.uint32)), (q, cp)
            assert L.voxg_synchronize(_stream()) == 0
    # BEST: one checkpoint; FAST / ADAPTIVE: three; ray merging: four
    assert failed == 1 + 3 + 3 + 4


def test_stack_bound_violation_is_reported(gpu, oracle):
    from voxvision_amd import _lib, scenes, voxtrees as VT
    L = gpu
    c2 = scenes.config_c2(frames=1)
    T = oracle.Tree(c2["dots"])
    G = VT.make_tree(c2["dots"])
    o, d = scenes.random_rays(1 << 16, *c2["bbox"], seed=5)
    vo, ho = T.rays(o, d)
    so = [torch.from_numpy(np.ascontiguousarray(o[:, k])).cuda()
          for k in range(3)]
    sd = [torch.from_numpy(np.ascontiguousarray(d[:, k])).cuda()
          for k in range(3)]
    n = len(o)
    vox = torch.empty(n, dtype=torch.int32, device="cuda")
    hit = [torch.full((n,), float("nan"), device="cuda") for _ in range(3)]

    def run():
        return L.voxg_ray_batch(G.handle, n, *[t.data_ptr() for t in so + sd],
                                vox.data_ptr(), None,
                                *[t.data_ptr() for t in hit], _stream())

    with VT.saved_launch_configs():
        VT.set_launch_config("rays", stack_lds=2)
        _lib.check(L.voxg_test_stack_bound(2))
        try:
            _lib.check(run())
            rc = L.voxg_synchronize(_stream())
        finally:
            _lib.check(L.voxg_test_stack_bound(0))
        assert rc == VOXG_EIO, rc
        assert b"stack bound" in L.voxg_last_error()
        # reported once: the flag is cleared by the report
        assert L.voxg_synchronize(_stream()) == 0
        # rays that needed the dropped entries became misses; none gained
        # a hit it should not have
        vg = _np_u32(vox)
        lost = vg != vo
        assert lost.any()
        assert (vg[lost] == MISS).all()
        # the same launch with the true bound is exact, and so is the
        # counting launch
        _lib.check(run())
        assert L.voxg_synchronize(_stream()) == 0
        assert np.array_equal(_np_u32(vox), vo)
        m = vo != MISS
        hg = np.stack([h.cpu().numpy() for h in hit])
        assert np.array_equal(hg[:, m].view(np.uint32),
                              ho[:, m].view(np.uint32))
        cnt = VT.ray_counts(G, o, d)
        assert cnt["rays"] == n
