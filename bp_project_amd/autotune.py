"""Autotuner for launch shapes: time candidates with HIP events and cache the
winner per (device arch, program, mode, interpolators, size class).

Two searches share one JSON cache (``BPMI_TUNE_CACHE`` or
``~/.cache/bp_project_amd/tune.json``):

* ``tune`` -- one knob, whole-program time per candidate.
* ``tune_config`` -- several knob groups searched one group at a time
  (coordinate search): for each group every candidate is run with the other
  groups at their current best, and each launch stage of the program is
  timed separately (events between the launches).  A group either picks the
  candidate with the least total time over its stages, or -- for knobs that
  can differ per stage, such as one launch shape per merge height -- the
  best candidate of each stage on its own.

The reference chose its launch parameters (leaf size, oversampling, region)
by hand sweeps (fast_bp/resident_experiment.py:25-31); here the sweep runs
on the device the program will use, and its record is kept beside the
choice.
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

import numpy as np


def cache_path():
    env = os.environ.get("BPMI_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "bp_project_amd" / "tune.json"


def _load():
    p = cache_path()
    try:
        cache = json.loads(p.read_text())
    except (OSError, ValueError):
        return {}
    return cache if isinstance(cache, dict) else {}


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


def make_key(kernel, mode, interp, work, arch=None, extra=""):
    """Cache key; ``arch`` defaults to device 0's architecture."""
    if arch is None:
        from .device import device_info
        arch = device_info(0)["arch"]
    arch = str(arch).split(":")[0]
    key = f"{arch}/{kernel}/{mode}/{interp}/s{size_class(work)}"
    return f"{key}/{extra}" if extra else key


def time_candidate(run, candidate, warmup=1, repeats=3):
    """Median event time (ms) of ``run(candidate)``."""
    from .device import Event, elapsed_ms, synchronize
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


def stage_times(execute, warmup=1, repeats=3):
    """Median time (ms) of each launch stage of ``execute(mark)``.

    ``execute`` enqueues the program and calls ``mark(name)`` after each
    launch; the stage ending at a mark is named after it, a name that recurs
    getting its occurrence number (``merges1``, ``merges2``, ..).
    """
    from .device import Event, elapsed_ms, synchronize
    for _ in range(warmup):
        execute(None)
    synchronize()
    runs, names = [], None
    for _ in range(repeats):
        first = Event()
        first.record()
        events, labels = [first], []

        def mark(name):
            e = Event()
            e.record()
            events.append(e)
            labels.append(name)
        execute(mark)
        events[-1].synchronize()
        runs.append([elapsed_ms(a, b) for a, b in zip(events, events[1:])])
        names = labels
    med = np.median(np.asarray(runs, dtype=np.float64), axis=0)
    return {k: float(v) for k, v in zip(stage_keys(names), med)}


def stage_keys(names):
    """Unique stage names: a name that occurs more than once gets its
    occurrence number (1-based), a name that occurs once stays as it is."""
    counts = {n: names.count(n) for n in names}
    seen, keys = {}, []
    for n in names:
        if counts[n] > 1:
            seen[n] = seen.get(n, 0) + 1
            keys.append(f"{n}{seen[n]}")
        else:
            keys.append(n)
    return keys


def tune(key, candidates, run, warmup=1, repeats=3, force=False, timer=None):
    """One knob: return (best_candidate, {candidate: ms}) and cache the winner.

    ``timer(run, candidate, warmup, repeats)`` measures one candidate
    (default: ``time_candidate``)."""
    timer = timer or time_candidate
    cache = _load()
    entry = cache.get(key)
    if (not force and isinstance(entry, dict) and
            entry.get("best") in [_enc(c) for c in candidates]):
        return _dec(entry["best"], candidates), entry.get("times", {})
    times = {}
    for c in candidates:
        times[_enc(c)] = float(timer(run, c, warmup, repeats))
    best = min(times, key=times.get)
    cache[key] = {"best": best, "times": times}
    _store(cache)
    return _dec(best, candidates), times


class Group:
    """One knob group of a coordinate search.

    name: the config field it sets; candidates: its values; stages: the
    launch stages it affects; per_stage: choose one value per stage (the
    config field then holds a list, one value per stage) instead of one
    value for all of them.
    """

    def __init__(self, name, candidates, stages, per_stage=False):
        if not candidates:
            raise ValueError(f"group {name!r} has no candidates")
        if not stages:
            raise ValueError(f"group {name!r} affects no stage")
        self.name = name
        self.candidates = [_norm(c) for c in candidates]
        self.stages = list(stages)
        self.per_stage = bool(per_stage)

    def valid(self, value):
        """Whether a (cached) value fits this group's candidates and stages."""
        encs = {_enc(c) for c in self.candidates}
        if self.per_stage:
            return (isinstance(value, list) and len(value) == len(self.stages) and
                    all(_enc(v) in encs for v in value))
        return _enc(value) in encs


def coordinate_search(groups, measure, base):
    """Search ``groups`` in order, each with the others at their current
    values.  ``measure(config) -> {stage: ms}``; ``base`` is the starting
    config (every group's field).  Returns (config, report) where
    report[group] = {"stages": [...], "times": {candidate: [ms per stage]},
    "choice": value}."""
    config = {k: _norm(v) for k, v in base.items()}
    for g in groups:
        if g.name not in config or not g.valid(config[g.name]):
            raise ValueError(f"base config has no valid value for {g.name!r}")
    report = {}
    for g in groups:
        times = {}
        for c in g.candidates:
            trial = dict(config)
            trial[g.name] = [c] * len(g.stages) if g.per_stage else c
            t = measure(trial)
            missing = [s for s in g.stages if s not in t]
            if missing:
                raise RuntimeError(f"measurement lacks stages {missing}")
            times[_enc(c)] = [float(t[s]) for s in g.stages]
        if g.per_stage:
            choice = [min(g.candidates, key=lambda c, i=i: times[_enc(c)][i])
                      for i in range(len(g.stages))]
        else:
            choice = min(g.candidates, key=lambda c: sum(times[_enc(c)]))
        config[g.name] = choice
        report[g.name] = {"stages": g.stages, "times": times, "choice": choice}
    return config, report


def tune_config(key, groups, measure, base, force=False):
    """Cached ``coordinate_search``: returns (config, report, cached).  A
    cached config is reused when every group's value is still one of its
    candidates (and, per stage, one per stage); otherwise, or with
    ``force``, the search runs and its result replaces the entry."""
    cache = _load()
    entry = cache.get(key)
    if not force and isinstance(entry, dict) and isinstance(entry.get("config"), dict):
        cfg = entry["config"]
        if all(g.name in cfg and g.valid(cfg[g.name]) for g in groups):
            return cfg, entry.get("report", {}), True
    config, report = coordinate_search(groups, measure, base)
    cache[key] = {"config": config, "report": report}
    _store(cache)
    return config, report, False


def _norm(c):
    """Candidates as JSON sees them (tuples become lists)."""
    return json.loads(json.dumps(c))


def _enc(c):
    return json.dumps(c, sort_keys=True)


def _dec(s, candidates):
    for c in candidates:
        if _enc(c) == s:
            return c
    return json.loads(s)
