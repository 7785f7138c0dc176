"""Launch-shape autotuner for the sweep kernels.

For a workload (list of systems + params) it times short sweeps of a scratch
context for each candidate launch shape and returns the fastest:

  * integer energies, Philox draws: the row-unrolled kernel's variants
    (sweep_kernel 0 = slot speculation, 2 = none, 4 = read-ahead) with 7 or 3
    producer wavefronts, and the row kernel (sweep_kernel 1);
  * field models, Philox draws: the field kernel (producers 0) or the generic
    kernel (producers -1), with the LDS-fitting group width or 16 walkers;
  * replay mode: the generic kernel at the LDS-fitting group width or 32.

Results are cached per workload signature (in memory, and in a JSON file when
a path is given).  The choice changes only speed: every launch shape runs
the same chains (tests/test_gpu_parity.py checks each kernel against the
oracle bit for bit).
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


TUNED_L = (16, 24, 32, 48, 64)


def candidates(systems, params):
    intpath = all(s.get("field") is None and s.get("h", 0.0) == 0.0 for s in systems)
    philox = params.get("rng_mode", abi.RNG_PHILOX) == abi.RNG_PHILOX
    if not philox:
        return [dict(walkers_per_group=w, producers=0, sweep_kernel=0) for w in (0, 32)]
    if not intpath:
        return [dict(walkers_per_group=w, producers=p, sweep_kernel=0) for p in (0, -1) for w in (0, 16)]
    out = []
    if all(s["L"] in TUNED_L for s in systems):
        for sk in (0, 2, 4):
            for prod in (0, 3):
                out.append(dict(walkers_per_group=0, producers=prod, sweep_kernel=sk))
    out.append(dict(walkers_per_group=0, producers=0, sweep_kernel=1))
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
    disk = json.load(open(cache_path)) if cache_path and os.path.exists(cache_path) else {}
    if key in _CACHE or key in disk:
        best = _CACHE[key] if key in _CACHE else disk[key]
        _CACHE[key] = best
        if cache_path and key not in disk:
            disk[key] = best
            json.dump(disk, open(cache_path, "w"), indent=1)
        return best, {}
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
        disk[key] = best
        json.dump(disk, open(cache_path, "w"), indent=1)
    return best, results
