"""Autotuner of the launch shapes (upf_set_tuning): block size and grid cap
of the grid-stride passes, candidates per lane and LDS staging depth of the
neighbour-list build, how far the host loop enqueues PCISPH iterations
ahead, and the list format (packed, delta-coded, or chunk unions staged in
LDS) and load width of the correction loop.  None of them changes results (every per-particle sum keeps its
order), so candidates are compared on the same steps from the same initial
state.  The best choice is kept per (workload, particle count) in a JSON
cache: $UPF_TUNING_CACHE, else ~/.cache/upf/tuning.json.
"""
import ctypes as C
import json
import os
import time

from . import UPF_PCISPH_STATIC, UPFError
from ._abi import Tuning

DEFAULT = dict(loop_block=256, grid_cap=8192, nbr_rounds=4, nbr_stage=0, speculate=4, loop_list=1, loop_width=8,
               nbr_tile=0, spec_first=4, xcd_remap=1, nbr_skip=0, nbr_fuse=1)
CANDIDATES = dict(loop_block=(64, 128, 256), grid_cap=(2048, 4096, 8192, 16384), nbr_rounds=(2, 4, 8),
                  nbr_stage=(0, 32, 64), speculate=(1, 2, 4, 8), loop_list=(0, 1, 2),
                  loop_width=(4, 8), nbr_tile=(0, 256, 512), spec_first=(0, 2, 4, 8, 64),
                  xcd_remap=(0, 1), nbr_skip=(0, 1), nbr_fuse=(0, 1))


def cache_path():
    return os.environ.get("UPF_TUNING_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "upf", "tuning.json"))


def cache_key(workload, n):
    return "%s:%d" % (workload, int(n))


def load(workload, n, path=None):
    """The cached tuning (dict) for (workload, n), or None."""
    path = path or cache_path()
    try:
        with open(path) as f:
            entry = json.load(f).get(cache_key(workload, n))
    except (OSError, ValueError):
        return None
    # entries written before a field existed take its default
    return None if entry is None else dict(DEFAULT, **entry["tuning"])


def store(workload, n, tuning, ms_per_step, trials, path=None):
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    data[cache_key(workload, n)] = {"tuning": dict(tuning), "ms_per_step": ms_per_step, "trials": trials}
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def coordinate_search(measure, start=None, candidates=None, sweeps=1, log=None):
    """Minimise measure(tuning dict) -> seconds one parameter at a time,
    starting from `start`; returns (best dict, best time, number of trials).
    `log`, a list, receives (tuning, seconds) of every trial."""
    cand = candidates or CANDIDATES
    best = dict(start or DEFAULT)
    seen = {}

    def run(t):
        k = tuple(sorted(t.items()))
        if k not in seen:
            try:
                seen[k] = measure(dict(t))
            except UPFError:  # a combination the library rejects (e.g. union lists need 256-thread blocks)
                seen[k] = float("inf")
            if log is not None:
                log.append((dict(t), seen[k]))
        return seen[k]

    best_t = run(best)
    for _ in range(sweeps):
        for name, values in cand.items():
            for v in values:
                t = dict(best)
                t[name] = v
                tt = run(t)
                if tt < best_t:
                    best, best_t = t, tt
    return best, best_t, len(seen)


def tune_simulation(sim, pos, types, workload, steps=4, warmup=1, settle=20, method=UPF_PCISPH_STATIC,
                    use_cache=True, path=None, log=None):
    """Tunes `sim` (a Simulation) on the scene (pos, types) and applies the
    result.  The scene first runs `settle` steps (a resting start needs few
    correction iterations, a moving fluid many more); every trial uploads that
    settled state afresh and times `steps` steps after `warmup`, so the caller
    re-uploads its own state afterwards.  Returns (tuning dict, ms/step)."""
    n = len(types)
    if use_cache:
        t = load(workload, n, path)
        if t is not None:
            sim.set_tuning(Tuning.from_dict(t))
            return t, None
    hip = C.CDLL("libamdhip64.so")
    sim.set_tuning(Tuning.from_dict(DEFAULT))
    sim.upload(pos, types)
    vel = None
    if settle > 0:
        sim.step(settle, method)
        st = sim.state()
        pos, vel = st["pos"], st["vel"]

    def measure(t):
        sim.set_tuning(Tuning.from_dict(t))
        sim.upload(pos, types, vel)
        sim.step(warmup, method)
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        sim.step(steps, method)
        hip.hipDeviceSynchronize()
        return time.perf_counter() - t0

    best, sec, trials = coordinate_search(measure, log=log)
    sim.set_tuning(Tuning.from_dict(best))
    ms = sec / steps * 1e3
    store(workload, n, best, ms, trials, path)
    return best, ms


def tune_batch(batch, pos, types, workload, steps=3, use_graph=True, use_cache=True, path=None, log=None):
    """Tunes a Batch on its scenes (pos, types concatenated); every trial
    re-uploads them, runs the eager first step and a graph capture, then
    times `steps` graph-replayed steps.  The caller re-uploads afterwards."""
    n = len(types)
    cand = dict(CANDIDATES)
    if use_graph:
        cand.pop("speculate")  # the graph never speculates
        cand.pop("spec_first")
    if use_cache:
        t = load(workload, n, path)
        if t is not None:
            batch.set_tuning(Tuning.from_dict(t))
            return t, None
    hip = C.CDLL("libamdhip64.so")

    def measure(t):
        batch.set_tuning(Tuning.from_dict(t))
        batch.upload(pos, types)
        batch.step(2, use_graph=use_graph)
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        batch.step(steps, use_graph=use_graph)
        hip.hipDeviceSynchronize()
        return time.perf_counter() - t0

    best, sec, trials = coordinate_search(measure, candidates=cand, log=log)
    batch.set_tuning(Tuning.from_dict(best))
    ms = sec / steps * 1e3
    store(workload, n, best, ms, trials, path)
    return best, ms
