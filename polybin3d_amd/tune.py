"""Autotuner for launch shapes (no counterpart in the reference).

A tunable call offers a few candidate shapes that all give the same result (up to FP32 rounding
order): the split-BF16 triplet GEMM's LDS stage length and A rows per launch, the number of Q maps
per fused pointing chain of the bispectrum MC Fisher, the number of power-spectrum seeds per batched
pass.  With tuning on, the first call of a (device, knob, shape) times every candidate with HIP events
on the current stream (one warm-up run each, then the median of a few runs, normalised per unit of
work) and keeps the fastest; the choice and the timings go to a JSON cache file, which later calls
and later processes read.  With tuning off, a cached choice is used if there is one, else the
built-in default.

  PB3D_AUTOTUNE=1        tune on first use (or tune.enable())
  PB3D_TUNE_CACHE=path   cache file (default ~/.cache/polybin3d_amd/tune.json)
"""
import json
import os
import tempfile
import threading

import torch

_lock = threading.Lock()
_state = {'enabled': None, 'cache': None, 'path': None}


def cache_path():
    return os.environ.get("PB3D_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "polybin3d_amd", "tune.json"))


def enable(flag=True):
    """Switch tuning on (True), off (False) or back to the PB3D_AUTOTUNE setting (None)."""
    _state['enabled'] = flag


def enabled():
    if _state['enabled'] is not None:
        return bool(_state['enabled'])
    return os.environ.get("PB3D_AUTOTUNE", "0") not in ("", "0")


def _load():
    path = cache_path()
    if _state['cache'] is None or _state['path'] != path:
        data = {}
        try:
            with open(path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
        _state['cache'], _state['path'] = data, path
    return _state['cache']


def _save():
    path = _state['path']
    d = os.path.dirname(path)
    try:
        os.makedirs(d, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=d, prefix=".tune")
        with os.fdopen(fd, "w") as f:
            json.dump(_state['cache'], f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs re-tuning


def device_key(device):
    p = torch.cuda.get_device_properties(device)
    return "%s|%s|%d" % (p.name, getattr(p, 'gcnArchName', ''), p.multi_processor_count)


def key_of(knob, shape, device):
    return "%s|%s|%s" % (device_key(device), knob, json.dumps(list(shape)))


def lookup(knob, shape, device):
    with _lock:
        e = _load().get(key_of(knob, shape, device))
    return None if e is None else _decode(e['choice'])


def _decode(v):
    return tuple(v) if isinstance(v, list) else v


def choose(knob, shape, candidates, run, device, default, work=None, reps=3):
    """The candidate to use for this (knob, shape): cached, or timed now when tuning is on, else
    `default`.  run(c) launches the work of candidate c on the current stream and may raise
    ValueError for a candidate that does not apply; work(c) is the amount of work c does (the time
    is divided by it; default 1)."""
    cands = [_decode(c) for c in candidates]
    c = lookup(knob, shape, device)
    if c is not None and c in cands:
        return c
    if not enabled():
        return default
    timings = {}
    for cand in cands:
        try:
            run(cand)
            ts = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                run(cand)
                e1.record()
                e1.synchronize()
                ts.append(e0.elapsed_time(e1))
        except ValueError:
            continue
        ts.sort()
        timings[cand] = ts[len(ts) // 2] / (work(cand) if work else 1.)
    if not timings:
        return default
    best = min(timings, key=timings.get)
    with _lock:
        _load()[key_of(knob, shape, device)] = {
            'choice': list(best) if isinstance(best, tuple) else best,
            'default': list(default) if isinstance(default, tuple) else default,
            'ms_per_unit': {json.dumps(list(k) if isinstance(k, tuple) else k): round(v, 5)
                            for k, v in timings.items()}}
        _save()
    return best


def entries():
    """The cache as a dict (for reports)."""
    with _lock:
        return dict(_load())
