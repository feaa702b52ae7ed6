"""Autotuning of the tile step path with a persistent cache.

`autotune(target, particles)` runs the library's autotuner (swe_autotune:
every compiled tile shape of the step kernel, then on the fastest every tile
order in {0, 1, 4, 16} with and without graphs and every heavy-first mode,
timed on the context's own workload from a saved state that is restored
afterwards) and
records the winner in a JSON cache keyed by GPU architecture, terrain sizes,
particles per terrain and the parameters that shape the launches; a later
call with the same key applies the cached setting without timing.  Settings
change speed only, never results.
"""
import ctypes
import json
import os

from . import capi

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "simplewinderosion_amd", "tuning.json")


def device_arch(target):
    buf = ctypes.create_string_buffer(128)
    target._check(target.lib.swe_device_arch(target.ctx, buf, 128), "swe_device_arch")
    return buf.value.decode()


def _key(target, particles):
    p = target.params()
    sizes = getattr(target, "sizes", None) or [(target.nx, target.ny)]
    return json.dumps({
        "arch": device_arch(target), "sizes": [list(s) for s in sizes], "particles": [int(n) for n in particles],
        "window": int(p.window), "layers": int(p.layers), "track": int(p.track), "periodic": int(p.periodic),
        "deposit": int(p.deposit_on_death), "max_age": int(p.max_age),
    }, sort_keys=True)


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(target, particles, trial_cycles=2, cache=DEFAULT_CACHE, retune=False):
    """Tune a World or Batch for `particles` (an int, or one per terrain);
    returns {"shape", "tile_order", "graph", "heavy_first", "ms_per_cycle", "candidates", "cached"}."""
    if isinstance(particles, int):
        particles = [particles]
    key = _key(target, particles)
    table = _load(cache) if cache else {}
    if not retune and key in table:
        hit = table[key]
        target.set_tune(capi.SWE_TUNE_SHAPE, hit.get("shape", 0))
        target.set_tune(capi.SWE_TUNE_TILE_ORDER, hit["tile_order"])
        target.set_tune(capi.SWE_TUNE_GRAPH, hit["graph"])
        target.set_tune(capi.SWE_TUNE_HEAVY_FIRST, hit.get("heavy_first", 1))
        return dict(hit, cached=True)
    target._push()
    arr = (ctypes.c_int64 * len(particles))(*particles)
    res = capi.SweTuneResult()
    target._check(target.lib.swe_autotune(target.ctx, arr, int(trial_cycles), ctypes.byref(res)), "swe_autotune")
    out = {"shape": res.shape, "tile_order": res.tile_order, "graph": res.graph, "heavy_first": res.heavy_first,
           "ms_per_cycle": res.ms_per_cycle, "candidates": res.candidates}
    if cache:
        table = _load(cache)
        table[key] = out
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return dict(out, cached=False)
