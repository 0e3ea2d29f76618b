"""Launch-schedule autotuner for the persistent one-level step.

The predictor and corrector are persistent kernels: each XCD walks a contiguous range of the block
processing order, and the order groups the blocks in tiles of tile_x x tile_y blocks so that a block's
neighbours are mostly in the same XCD's L2.  Two choices set how well that works on a given mesh:
the tile shape, and how many workgroups per CU each kernel keeps resident (more hide latency, fewer
keep the L2 working set small).  Results do not depend on either (every block is computed the same
way and the time step is a minimum), so the tuner only times.

The search is coordinate-wise: tiles at the default workgroup counts, then the predictor's
workgroups per CU, then the corrector's, each candidate warmed up (first run captures its graph) and
timed as the best of a few repeats of `steps` steps.  Tuning advances the state it runs on, so it runs
on a scratch Domain or Batch (autotune_domain and autotune_batch build one from the caller's inputs
and free it).  Chosen schedules can be kept in a JSON cache keyed by the workload.
"""
import json
import os
import time

TILES = ((16, 12), (8, 8), (16, 16), (32, 8), (32, 16), (64, 4), (4, 4), (1 << 12, 1))
MAX_WG = 8


def _time(run, steps, repeats):
    run(max(2, steps // 4))  # warm-up: captures the candidate's graph
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        run(steps)
        best = min(best, time.perf_counter() - t0)
    return best


def tune(obj, run, steps=16, repeats=3, tiles=TILES, max_wg=MAX_WG, log=None):
    """Coordinate search over the schedules of obj (a Domain or Batch): obj.set_schedule(...) then
    run(n) advances obj by n steps and waits for them.  Returns the best schedule and the timings."""
    base = obj.schedule()
    trials = []

    def trial(tx, ty, wp, wc):
        obj.set_schedule(tx, ty, wp, wc)
        t = _time(run, steps, repeats)
        trials.append(dict(tile_x=tx, tile_y=ty, wg_pred=wp, wg_corr=wc, s_per_step=t / steps))
        if log:
            log("schedule tile %dx%d wg %d,%d: %.4f ms/step" % (tx, ty, wp, wc, 1e3 * t / steps))
        return t

    best = (base["tile_x"], base["tile_y"], 0, 0)
    best_t = trial(*best)
    for tx, ty in tiles:
        if (tx, ty) == best[:2]:
            continue
        t = trial(tx, ty, 0, 0)
        if t < best_t:
            best, best_t = (tx, ty, 0, 0), t
    for which in (2, 3):  # predictor, then corrector workgroups per CU
        for w in range(1, max_wg + 1):
            cand = list(best)
            cand[which] = w
            t = trial(*cand)
            if t < best_t:
                best, best_t = tuple(cand), t
    obj.set_schedule(*best)
    return dict(zip(("tile_x", "tile_y", "wg_pred", "wg_corr"), best)), trials


def cache_key(kind, precision, nblk, extra=""):
    """Workload key: kind ("domain"/"batch"), precision and the active block count rounded to a power
    of two (the schedule depends on the mesh size, not on its values)."""
    b = 1
    while b < nblk:
        b <<= 1
    return "%s|%s|%d%s" % (kind, precision, b, ("|" + extra) if extra else "")


def load_cache(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def store_cache(path, key, sched):
    data = load_cache(path)
    data[key] = sched
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune_domain(param, blocks, zb, h, u, v, cf, steps=16, repeats=3, cache=None, log=None, **kw):
    """Schedule for a one-level Domain with this mesh and state, tuned on a scratch Domain."""
    from . import Domain, Loop

    key = cache_key("domain", param["precision"], int(param["nblk"]))
    hit = load_cache(cache).get(key)
    if hit:
        return hit, []
    d = Domain(param, blocks)
    try:
        d.set_state(zb, h, u, v, cf)
        loop = d.init_dt(Loop(nextoutputtime=1e30))

        def run(n):
            d.run(loop, n)

        sched, trials = tune(d, run, steps=steps, repeats=repeats, log=log, **kw)
    finally:
        d.close()
    if cache:
        store_cache(cache, key, sched)
    return sched, trials


def autotune_batch(common, max_blocks, max_members, members, steps=16, repeats=3, cache=None, log=None, **kw):
    """Schedule for a Batch of these members, each a dict of Batch.add's arguments, tuned on a scratch
    Batch whose members never end during tuning."""
    from . import Batch

    nblk = sum(int(m["param"]["nblk"]) for m in members)
    key = cache_key("batch", common["precision"], nblk, "%d" % len(members))
    hit = load_cache(cache).get(key)
    if hit:
        return hit, []
    b = Batch(common, max_blocks, max_members)
    try:
        for m in members:
            args = dict(m)
            args["endtime"] = 1e30
            b.add(**args)
        sched, trials = tune(b, b.run, steps=steps, repeats=repeats, log=log, **kw)
    finally:
        b.close()
    if cache:
        store_cache(cache, key, sched)
    return sched, trials
