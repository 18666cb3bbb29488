"""voxg_ray_batch_noalloc (the entry point for callers that cannot allocate)
against voxg_ray_batch through the C ABI: two calls in a row, the second
with fewer rays, each bit-identical to voxg_ray_batch (voxel ids, leaf
nodes, hit-point bits, misses untouched); then a tree of the same scene
with a deep stack bound, and the empty call."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _soa(a):
    return [torch.from_numpy(np.ascontiguousarray(a[:, k])).cuda()
            for k in range(3)]


def _run(L, fn, G, so, sd, n):
    vox = torch.full((n,), -2, dtype=torch.int32, device="cuda")
    leaf = torch.full((n,), -2, dtype=torch.int32, device="cuda")
    hit = [torch.full((n,), float("nan"), device="cuda") for _ in range(3)]
    from voxvision_amd import _lib
    _lib.check(fn(G.handle, n, *[t.data_ptr() for t in so + sd],
                  vox.data_ptr(), leaf.data_ptr(),
                  *[h.data_ptr() for h in hit], _stream()), fn.__name__)
    _lib.check(L.voxg_synchronize(_stream()), "voxg_synchronize")
    return (vox.cpu().numpy(), leaf.cpu().numpy(),
            np.stack([h.cpu().numpy() for h in hit]).view(np.uint32))


def test_noalloc_matches_ray_batch(gpu):
    from voxvision_amd import scenes, voxtrees as VT
    L = gpu
    c2 = scenes.config_c2(frames=1)
    G = VT.make_tree(c2["dots"])
    bmin, bmax = c2["bbox"]
    o, d = scenes.random_rays(1 << 21, bmin, bmax, seed=41)
    # camera rays of one pose after the random ones
    cam_o, cam_d = scenes.camera_rays(c2["cameras"][0], 480, 272)
    o = np.concatenate([o, cam_o]).astype(np.float32)
    d = np.concatenate([d, cam_d]).astype(np.float32)
    so, sd = _soa(o), _soa(d)
    for n in (len(o), 100003):
        ref = _run(L, L.voxg_ray_batch, G, so, sd, n)
        got = _run(L, L.voxg_ray_batch_noalloc, G, so, sd, n)
        for a, b, what in zip(ref, got, ("voxel", "leaf", "hit")):
            assert np.array_equal(a, b), (n, what)
        assert (ref[0] != -2).all()
        assert (ref[0] != -1).any() and (ref[0] == -1).any()


def test_noalloc_small_ring_and_empty(gpu):
    """A 2-entry LDS ring makes the launch spill most of its pending
    entries into the scratch array, so the grid is cut to what the array
    holds; the results stay those of voxg_ray_batch."""
    from voxvision_amd import scenes, voxtrees as VT
    L = gpu
    c1 = scenes.config_c1()
    G = VT.make_tree(c1["dots"])
    o, d = scenes.random_rays(300000, *c1["bbox"], seed=43)
    so, sd = _soa(o), _soa(d)
    with VT.saved_launch_configs():
        VT.set_launch_config("rays", stack_lds=2, waves_per_cu=32)
        for n in (300000, 4097):
            ref = _run(L, L.voxg_ray_batch, G, so, sd, n)
            got = _run(L, L.voxg_ray_batch_noalloc, G, so, sd, n)
            for a, b in zip(ref, got):
                assert np.array_equal(a, b), n
    assert L.voxg_ray_batch_noalloc(G.handle, 0, None, None, None, None,
                                    None, None, None, None, None, None,
                                    None, _stream()) == 0
