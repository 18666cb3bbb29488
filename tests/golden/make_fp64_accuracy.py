"""Hit accuracy of the fp32 search against FP64, on the benchmark workloads.

Writes tests/golden/fp64_accuracy.json.  The fp32 answers are the oracle's
(the GPU's equal them bit for bit, which the GPU tests check); the FP64
answers come from the oracle's FP64 grid walk (orc_dda64_rays), which
tests every voxel box near the ray with the exact slab test.  Per
workload: rays checked, the tie fraction (rays whose first voxel fp32
rounding could change: entries within eps, grazed voxels, hit points on a
voxel edge), hit/miss and voxel agreement outside the tie set, and the hit
point error |fp32 - fp64| (max over xyz) in voxels: median and maximum.

  C1  64^3 blobs, the full 640x480 frame (BEST)
  C2  the 256^3 CT scene, 2 of the 64 orbit poses at 1920x1080, 60,000
      pixels each whose ray enters the scene box (seeded sample)
  C3  the same scene, 200,000 of the 16 M incoherent rays (the first ones
      of the benchmark's seed)

    python tests/golden/make_fp64_accuracy.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "fp64_accuracy.json")
EPS = 1e-6  # relative to the scene's coordinate scale


def _check(O, T, dots, o, d, vox, hit):
    """Compare fp32 answers (vox, hit (3, n)) with the FP64 grid walk."""
    idx, h64, tie = O.dda64_rays(dots, o, d, eps=EPS)
    ok = ~tie
    hit32 = vox != O.MISS
    hit64 = idx >= 0
    tab = T.voxel_table()
    both = ok & hit32 & hit64
    same_hit = np.array_equal(hit32[ok], hit64[ok])
    same_voxel = np.array_equal(tab[vox[both].astype(np.int64)],
                                dots[idx[both]])
    err = np.abs(hit[:, both].T.astype(np.float64) - h64[both]).max(1)
    scale = max(1.0, float(np.abs(dots).max()) + 1.0,
                float(np.abs(o).max()))
    return {
        "rays": int(len(o)),
        "hits": int(hit32.sum()),
        "tie_fraction": round(float(tie.mean()), 6),
        "hit_miss_agree_outside_ties": bool(same_hit),
        "voxel_agree_outside_ties": bool(same_voxel),
        "hit_error_voxels_median": float("%.3g" % np.median(err))
        if len(err) else 0.0,
        "hit_error_voxels_max": float("%.3g" % err.max()) if len(err) else 0.0,
        "hit_error_rel_max": float("%.3g" % (err.max() / scale))
        if len(err) else 0.0,
        "coordinate_scale": scale,
    }


def measure():
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from voxvision_amd import scenes
    res = {"eps_relative": EPS}
    # C1: the full frame
    c1 = scenes.config_c1()
    T = O.Tree(c1["dots"])
    cam, W, H = c1["camera"], c1["W"], c1["H"]
    _, vox, hit, _ = T.render(cam, W, H, O.QUALITY_BEST)
    d = O.camera_rays(cam, W, H)
    o = np.tile(np.asarray(cam["position"], np.float32), (len(d), 1))
    res["c1_frame_640x480"] = _check(O, T, c1["dots"], o, d, vox, hit)
    T.close()
    # C2: two 1080p poses, pixels whose rays enter the scene box
    c2 = scenes.config_c2(frames=64)
    dots = c2["dots"]
    T = O.Tree(dots)
    bmin, bmax = scenes.bbox_of(dots)
    rng = np.random.default_rng(29)
    parts = []
    for k in (5, 44):
        cam = c2["cameras"][k]
        d = O.camera_rays(cam, c2["W"], c2["H"])
        pos = np.asarray(cam["position"], np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            t0 = (bmin - pos) / d
            t1 = (bmax - pos) / d
        enter = (np.minimum(t0, t1).max(1) <= np.maximum(t0, t1).min(1))
        pick = rng.choice(np.nonzero(enter)[0], 60000, replace=False)
        parts.append((np.tile(pos.astype(np.float32), (60000, 1)), d[pick]))
    o = np.concatenate([p[0] for p in parts])
    d = np.concatenate([p[1] for p in parts])
    vox, hit = T.rays(o, d)
    res["c2_two_poses_1080p_120k"] = _check(O, T, dots, o, d, vox, hit)
    # C3: incoherent rays of the benchmark's seed
    o, d = scenes.random_rays(1 << 24, bmin, bmax, seed=11)
    o, d = o[:200000], d[:200000]
    vox, hit = T.rays(o, d)
    res["c3_incoherent_200k"] = _check(O, T, dots, o, d, vox, hit)
    T.close()
    return res


if __name__ == "__main__":
    r = measure()
    json.dump(r, open(OUT, "w"), indent=1, sort_keys=True)
    print(json.dumps(r, indent=1, sort_keys=True))
