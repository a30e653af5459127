"""Tiny autotuner: time candidate launch configurations with HIP events and
cache the winner per (device arch, kernel, mode, interpolator, size class).

The cache is a JSON file (``BPMI_TUNE_CACHE`` or
``~/.cache/bp_project_amd/tune.json``).  ``tune`` runs each candidate
``warmup`` + ``repeats`` times and keeps the lowest median event time.
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

import numpy as np

from .device import Event, device_info, elapsed_ms, synchronize


def cache_path():
    env = os.environ.get("BPMI_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "bp_project_amd" / "tune.json"


def _load():
    p = cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store(cache):
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass


def size_class(n):
    """Power-of-four bucket of a work size."""
    return int(round(math.log(max(int(n), 1), 4)))


def make_key(kernel, mode, interp, work):
    arch = device_info(0)["arch"].split(":")[0]
    return f"{arch}/{kernel}/{mode}/{interp}/s{size_class(work)}"


def time_candidate(run, candidate, warmup=1, repeats=3):
    for _ in range(warmup):
        run(candidate)
    synchronize()
    times = []
    for _ in range(repeats):
        a, b = Event(), Event()
        a.record()
        run(candidate)
        b.record()
        b.synchronize()
        times.append(elapsed_ms(a, b))
    return float(np.median(times))


def tune(key, candidates, run, warmup=1, repeats=3, force=False):
    """Return (best_candidate, {candidate: ms}) and cache the winner."""
    cache = _load()
    if not force and key in cache and cache[key]["best"] in [_enc(c) for c in candidates]:
        best = cache[key]["best"]
        return _dec(best, candidates), cache[key].get("times", {})
    times = {}
    for c in candidates:
        times[_enc(c)] = time_candidate(run, c, warmup, repeats)
    best = min(times, key=times.get)
    cache[key] = {"best": best, "times": times}
    _store(cache)
    return _dec(best, candidates), times


def _enc(c):
    return json.dumps(c, sort_keys=True)


def _dec(s, candidates):
    for c in candidates:
        if _enc(c) == s:
            return c
    return json.loads(s)
