"""Autotuner for the traversal launch configuration.

The tuned parameters are those of voxg_launch_config: entries of the
per-lane LDS stack ring (`stack_lds`), threads per workgroup (`block`),
persistent waves per CU (`waves_per_cu`), idle lanes that trigger a refill
(`refill`), rays a lane takes per fetch (`chunk`), the render pre-pass
(`render_cull`) and the lanes a wave gathers before it searches leaves
(`leaf_batch`) or finishes traversals (`finish_batch`), and the pixels of
a 4x4 block a lane takes at a time in BEST mode (`split_best`), for one
kind of launch ("rays", "render", "render_best", "batch").  The search is
coordinate descent from the kind's current configuration: each parameter in turn
is swept over its candidate values with the others fixed, keeping the
fastest, until a full pass changes nothing.  Timing uses HIP events on the
current stream; each candidate is warmed up once and timed over `reps`
runs.  Configurations whose stack ring (stack_lds * block * 20 B of LDS)
exceeds the 160 KiB of LDS of a gfx950 CU are skipped.
"""
import itertools
import json

import torch

from . import voxtrees as VT

SPACE = {
    "stack_lds": [2, 4, 8, 16, 32],
    "block": [64, 128, 256],
    "waves_per_cu": [8, 12, 16, 20, 24, 32],
    "refill": [1, 4, 8, 16, 32, 48, 64],
    "chunk": [1, 2, 4, 8, 16, 32],
    "render_cull": [0, 1],
    "leaf_batch": [1, 8, 16, 24, 32, 48, 64],
    "finish_batch": [1, 8, 16, 24, 32, 48, 64],
    "split_best": [0, 1, 2, 3, 4],
}


ENTRY_BYTES = 20  # one stack entry: node, entry point, pending list
LDS_BYTES = 160 * 1024


def _valid(cfg):
    return cfg["stack_lds"] * cfg["block"] * ENTRY_BYTES <= LDS_BYTES


def time_config(make_runner, cfg, reps=3, kind=None):
    """Milliseconds per run of the workload under `cfg` (set for `kind`, or
    for every kind when None)."""
    VT.set_launch_config(kind, **cfg)
    run = make_runner(cfg)
    run()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(make_runner, params=None, start=None, reps=3, max_passes=2,
             log=None, kind=None):
    """make_runner(cfg) -> zero-argument callable running the workload once
    (built after the configuration is set, for objects such as batches
    whose layout depends on it); `kind` is the kind of launch the workload
    makes.  Leaves the best configuration set for that kind and returns
    (best_cfg, best_ms, history)."""
    params = list(params or SPACE.keys())
    cfg = dict(start or VT.get_launch_config(kind or "rays"))
    best = time_config(make_runner, cfg, reps, kind)
    history = [(dict(cfg), best)]
    if log:
        log("start %s %.3f ms" % (json.dumps(cfg), best))
    seen = {json.dumps(cfg, sort_keys=True): best}
    for _ in range(max_passes):
        changed = False
        for p in params:
            for v in SPACE[p]:
                cand = dict(cfg)
                cand[p] = v
                key = json.dumps(cand, sort_keys=True)
                if key in seen or not _valid(cand):
                    continue
                ms = time_config(make_runner, cand, reps, kind)
                seen[key] = ms
                history.append((dict(cand), ms))
                if ms < best * 0.99:
                    best, cfg, changed = ms, cand, True
                    if log:
                        log("  %s=%s -> %.3f ms" % (p, v, ms))
        if not changed:
            break
    VT.set_launch_config(kind, **cfg)
    return cfg, best, history


def grid(params):
    """All combinations of the given parameter lists (small sweeps)."""
    keys = list(params)
    for vals in itertools.product(*[params[k] for k in keys]):
        yield dict(zip(keys, vals))
