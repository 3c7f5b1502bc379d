"""Autotuner for the path pool and kernel configuration.

The knobs change speed only, never results (every film is bit-identical
across them): the path-pool size (paths in flight, i.e. the queue capacity
of every bounce stage), the waves per SIMD the shading kernel is compiled
for, where the traversal stack lives (private scratch or LDS) and whether
the bounce iterations are replayed from a captured hipGraph.  The
tuner times short trial renders of a representative job list on the
current device, keeps the fastest configuration and caches it per device
and workload signature (JSON under $XDG_CACHE_HOME/3d-ray_amd or
~/.cache/3d-ray_amd).
"""
import hashlib
import json
import os
import time

from . import native

POOLS = (1 << 22, 1 << 23, 1 << 24, 1 << 25)
SHADE_WAVES = (2, 3)
STACKS = (1, 2)   # 1 private scratch, 2 LDS
GRAPHS = (1, 2)   # 1 direct launches, 2 hipGraph replay of the bounce iterations


def _cache_path():
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "3d-ray_amd", "autotune.json")


def signature(items):
    """Workload signature: resolutions, spp, depth and scene sizes."""
    h = hashlib.sha1()
    for r, w, hh, *_ in items:
        info = r.scene.info()
        h.update(f"{w}x{hh}:{r.settings.samples_per_pixel}:{r.settings.max_depth}:{info.num_primitives};".encode())
    return h.hexdigest()[:16]


def _trial(ctx, items, pool, waves, stack, graph, batch):
    ctx.tune(waves, stack, graph)
    t0 = time.perf_counter()
    if batch:
        native.render_batch(items, pool_paths=pool, films=False)
    else:
        for r, w, h, *_ in items:
            r.settings.paths_in_flight = pool
            r.render(w, h)
    return time.perf_counter() - t0


def tune(items, batch=True, trial_spp=4, pools=POOLS, shade_waves=SHADE_WAVES, stacks=STACKS, graphs=GRAPHS,
         use_cache=True, log=None, repeats=2):
    """Picks (pool, shade waves, stack, graph) for `items` = [(Renderer, w,
    h, ...)] from timed trial renders at `trial_spp` samples per pixel (the
    fastest of `repeats` runs per configuration), applies it to the
    renderers' context and returns it as a dict."""
    ctx = items[0][0].scene.ctx
    key = f"{signature(items)}:{'batch' if batch else 'plain'}"
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if batch:
        stacks = (1,)   # batched kernels always keep the traversal stack in private scratch
    if use_cache and key in cache:
        best = cache[key]
    else:
        saved = [r.settings.samples_per_pixel for r, *_ in items]
        for r, *_ in items:
            r.settings.samples_per_pixel = min(trial_spp, r.settings.samples_per_pixel)
        trials = []
        try:
            _trial(ctx, items, pools[len(pools) // 2], shade_waves[0], stacks[0], graphs[0], batch)   # warm-up
            for pool in pools:
                for waves in shade_waves:
                    for stack in stacks:
                        for graph in graphs:
                            dt = min(_trial(ctx, items, pool, waves, stack, graph, batch) for _ in range(repeats))
                            trials.append((dt, pool, waves, stack, graph))
                            if log:
                                log(f"autotune pool={pool} shade_waves={waves} stack={stack} graph={graph}: "
                                    f"{dt * 1e3:.1f} ms")
        finally:
            for (r, *_), spp in zip(items, saved):
                r.settings.samples_per_pixel = spp
        dt, pool, waves, stack, graph = min(trials)
        best = {"pool": pool, "shade_waves": waves, "stack": stack, "graph": graph, "trial_ms": dt * 1e3}
        if use_cache:
            cache[key] = best
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                with open(path, "w") as f:
                    json.dump(cache, f, indent=1)
            except OSError:
                pass
    ctx.tune(best["shade_waves"], best["stack"], best.get("graph", 1))
    if not batch:
        for r, *_ in items:
            r.settings.paths_in_flight = best["pool"]
    return best
