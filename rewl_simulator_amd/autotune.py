"""Launch-shape autotuner for the sweep kernel.

For a workload (list of systems + params) it times short sweeps of a scratch
context for each candidate launch shape and returns the fastest:

  * walkers per wavefront group: the LDS-fitting maximum (0), 32 or 16 --
    fewer lanes per group means less LDS per workgroup, so more workgroups per
    CU for large batches, at the cost of idle lanes;
  * producer wavefronts per workgroup (integer/Philox path only): none (-1),
    1 or 2 -- threshold generation on other SIMDs versus in the walker wave.

Results are cached per workload signature (in memory, and in a JSON file when
a path is given).  The choice changes only speed: every launch shape runs
the same chains (tests/test_gpu_parity.py::test_fast_path_producers_bit_exact).
"""
import hashlib
import json
import os

from . import abi
from .engine import Engine

_CACHE = {}


def signature(systems, params):
    items = []
    for s in systems:
        items.append((s["model"], s["L"], s["J"], s["K"], s["field"] is not None if "field" in s else False,
                      s["num_windows"], s["replicas_per_window"], round(s["e_max"] - s["e_min"], 6), s["bin_size"]))
    items.sort()
    h = hashlib.sha1(json.dumps([items, params.get("rng_mode", 0)], default=str).encode()).hexdigest()[:16]
    return h


def candidates(systems, params):
    intpath = all(s.get("field") is None and s.get("h", 0.0) == 0.0 for s in systems)
    fast = intpath and params.get("rng_mode", abi.RNG_PHILOX) == abi.RNG_PHILOX
    out = []
    for wpg in (0, 32, 16):
        for prod in ((-1, 1, 2) if fast else (0,)):
            out.append(dict(walkers_per_group=wpg, producers=prod))
    return out


def measure(systems, params, shape, sweeps):
    p = dict(params, **shape)
    with Engine(systems, p) as eng:
        eng.adjust_to_range()
        eng.sweeps(max(1, sweeps // 4))
        st0 = eng.stats()
        eng.sweeps(sweeps)
        st1 = eng.stats()
    ms = st1["kernel_ms_sweep"] - st0["kernel_ms_sweep"]
    att = st1["attempts"] - st0["attempts"]
    return att / (ms * 1e-3) if ms > 0 else 0.0


def autotune(systems, params, sweeps=20, cache_path=None, verbose=False):
    """Returns (best_shape, {shape: attempts/s}) and remembers the choice."""
    key = signature(systems, params)
    if key in _CACHE:
        return _CACHE[key], {}
    if cache_path and os.path.exists(cache_path):
        disk = json.load(open(cache_path))
        if key in disk:
            _CACHE[key] = disk[key]
            return disk[key], {}
    results = {}
    best, best_rate = None, -1.0
    for shape in candidates(systems, params):
        try:
            rate = measure(systems, params, shape, sweeps)
        except abi.REWLError:
            continue            # e.g. the shape does not fit in LDS
        results[json.dumps(shape, sort_keys=True)] = rate
        if verbose:
            print(f"autotune {shape}: {rate:.3e} attempts/s", flush=True)
        if rate > best_rate:
            best, best_rate = shape, rate
    _CACHE[key] = best
    if cache_path:
        disk = json.load(open(cache_path)) if os.path.exists(cache_path) else {}
        disk[key] = best
        json.dump(disk, open(cache_path, "w"), indent=1)
    return best, results
