"""Autotuner: launch shapes and schedules chosen by timing them on the device.

* tune_plan(options): views per block (4 or 8) of the line->bin kernel for one
  geometry, keyed by (H, W, V, D, scan_type, dtype).
* tune_batch(batch): resident blocks per CU and item order of the persistent
  heterogeneous kernel, keyed by the batch's shape multiset.

Results persist in a JSON file ($CTLIB_AMD_TUNE_CACHE, default
~/.cache/ctlib_amd/tune.json) and are applied automatically: plans through
apply_cached(options) (called by tune_plan and usable before a run), batches
when a HeteroBatch is built.
"""
import json
import os
import threading

import torch

from . import _C

_LOCK = threading.Lock()


def cache_path():
    return os.environ.get("CTLIB_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "ctlib_amd", "tune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(section, key, value):
    with _LOCK:
        data = _load()
        data.setdefault(section, {})[key] = value
        path = cache_path()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".tmp{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def plan_key(options, dtype=torch.float32):
    g = _C.parse_options(torch.as_tensor(options).detach().to("cpu", torch.float64))
    return f"{g['height']}x{g['width']}|V{g['views']}|D{g['dets']}|st{g['scan_type']}|{str(dtype).split('.')[-1]}"


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        a.record()
        fn()
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def apply_cached(options, device="cuda:0", dtype=torch.float32):
    """Apply a cached plan tuning, if any; returns the chosen views-per-block or None."""
    hit = _load().get("plan", {}).get(plan_key(options, dtype))
    if hit is None:
        return None
    dev = torch.device(device)
    _C.set_plan_tuning(torch.as_tensor(options).to("cpu"), dev.index or 0, int(hit["fwd_G"]))
    return int(hit["fwd_G"])


def tune_plan(options, batch=8, device="cuda:0", dtype=torch.float32, reps=5, candidates=(4, 8)):
    """Time projection with each views-per-block candidate; keep and cache the fastest."""
    dev = torch.device(device)
    o = torch.as_tensor(options).to("cpu")
    g = _C.parse_options(o.to(torch.float64))
    x = torch.rand(batch, 1, g["height"], g["width"], device=dev, dtype=dtype)
    times = {}
    for G in candidates:
        _C.set_plan_tuning(o, dev.index or 0, G)
        times[G] = _time(lambda: _C.projection(x, o), reps)
    best = min(times, key=times.get)
    _C.set_plan_tuning(o, dev.index or 0, best)
    _store("plan", plan_key(options, dtype), {"fwd_G": best, "ms": times})
    return best, times


def cached_batch_schedule(signature):
    hit = _load().get("batch", {}).get(signature)
    return None if hit is None else (hit["blocks_per_cu"], hit["order"])


def tune_batch(hb, reps=3, candidates=((1, 0), (2, 0), (2, 1), (3, 0), (4, 0))):
    """Time the persistent kernel under each (blocks_per_cu, order); keep and cache the fastest."""
    use_graph = hb.use_graph
    hb.use_graph = False
    times = {}
    for bpc, order in candidates:
        hb.set_schedule(bpc, order)
        times[(bpc, order)] = _time(hb.launch, reps)
    best = min(times, key=times.get)
    hb.set_schedule(*best)
    hb.use_graph = use_graph
    _store("batch", hb.signature(),
           {"blocks_per_cu": best[0], "order": best[1], "ms": {f"{k[0]},{k[1]}": v for k, v in times.items()}})
    return best, times
