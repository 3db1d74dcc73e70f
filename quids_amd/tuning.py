"""Autotuner of the step's launch and layout parameters (qd_tuning).

The knobs: the dedup workgroup's LDS table size and the mean candidates per
dedup bucket (how full that table runs: the load factor of the interference
hash table), the bits of the first radix pass, and the generate kernel's
grid.  Combinations whose bucket target passes half the table size are
skipped.  ``autotune`` times a
representative step for each combination on the caller's context and keeps
the fastest; results are remembered per (device, workload key) in memory and,
when ``cache_path`` is given, in a JSON file.
"""
import itertools
import json
import os
import time

from . import UINT64_MAX, it_t, simulate

DEFAULT_GRID = {
    "dedup_cap": (512, 1024, 2048),
    "dedup_bucket_target": (256, 384, 512, 1024),
    "first_pass_bits": (8, 9),
    "generate_grid": (4096, 16384),
}

_memo = {}


def _time_step(ctx, state, rules, max_num_object, reps):
    """Best wall time of simulate(rule) for every rule of `rules` applied to
    a copy of `state` (the state is not modified)."""
    best = float("inf")
    for _ in range(reps):
        cur = state.copy()
        ctx.synchronize()
        t = time.perf_counter()
        for r in rules:
            nxt = it_t(ctx=ctx)
            simulate(cur, r, nxt, None, max_num_object)
            cur.close()
            cur = nxt
        ctx.synchronize()
        best = min(best, time.perf_counter() - t)
        cur.close()
    return best


def autotune(ctx, state, rules, max_num_object=UINT64_MAX, grid=None, reps=2, key=None, cache_path=None,
             verbose=False):
    """Picks the qd_tuning values (of `grid`, default DEFAULT_GRID) that run
    `rules` fastest on `state`, sets them on ctx and returns
    {"tuning": {...}, "seconds": best, "default_seconds": t0, "tried": n}.
    The rules run on a copy of state; results are bit-identical for every
    tuning (only launch shapes and bucket sizes change)."""
    grid = grid or DEFAULT_GRID
    dev_key = (ctx.device, key) if key is not None else None
    if dev_key is not None and dev_key in _memo:
        ctx.set_tuning(**_memo[dev_key]["tuning"])
        return _memo[dev_key]
    if cache_path and key is not None and os.path.exists(cache_path):
        with open(cache_path) as f:
            disk = json.load(f)
        hit = disk.get(f"{ctx.device}:{key}")
        if hit:
            ctx.set_tuning(**hit["tuning"])
            _memo[dev_key] = hit
            return hit
    start = ctx.get_tuning()
    names = list(grid)
    _time_step(ctx, state, rules, max_num_object, 1)  # warm-up: scratch allocation
    t0 = _time_step(ctx, state, rules, max_num_object, reps)
    best_t, best = t0, {k: start[k] for k in names}
    tried = 0
    for values in itertools.product(*[grid[k] for k in names]):
        cand = dict(zip(names, values))
        if cand.get("dedup_bucket_target", start["dedup_bucket_target"]) * 2 > cand.get("dedup_cap", start["dedup_cap"]):
            continue
        ctx.set_tuning(**cand)
        t = _time_step(ctx, state, rules, max_num_object, reps)
        tried += 1
        if verbose:
            print(f"autotune {cand}: {t * 1e3:.2f} ms", flush=True)
        if t < best_t:
            best_t, best = t, cand
    ctx.set_tuning(**best)
    res = {"tuning": best, "seconds": best_t, "default_seconds": t0, "tried": tried}
    if dev_key is not None:
        _memo[dev_key] = res
        if cache_path:
            disk = {}
            if os.path.exists(cache_path):
                with open(cache_path) as f:
                    disk = json.load(f)
            disk[f"{ctx.device}:{key}"] = res
            with open(cache_path, "w") as f:
                json.dump(disk, f, indent=1)
    return res
