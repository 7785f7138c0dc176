"""Python front end of the HIP REWL engine over the C ABI (include/rewl_c.h).

Mirrors the reference's REWL_simulation surface (rewl_simulation.hpp:21-66):
construct from model and REWL parameters, adjust walkers into their windows,
``simulate`` (run to the final ln f), then export the stitched and normalised
ln g.  Every call goes to the GPU; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import abi

_SYSTEM_KEYS = ("model", "L", "J", "K", "h", "e_min", "e_max", "bin_size", "num_windows",
                "replicas_per_window", "overlap", "seed")


def _desc_array(systems, keep):
    arr = (abi.SystemDesc * len(systems))()
    for d, s in zip(arr, systems):
        for k in _SYSTEM_KEYS:
            setattr(d, k, s[k])
        field = s.get("field")
        if field is not None:
            buf = np.ascontiguousarray(field, dtype=np.float64)
            keep.append(buf)
            d.field = buf.ctypes.data_as(C.POINTER(C.c_double))
        ex = s.get("exclude_bins")
        if ex is not None:
            buf = np.ascontiguousarray(ex, dtype=np.uint8)
            keep.append(buf)
            d.exclude_bins = buf.ctypes.data_as(C.POINTER(C.c_uint8))
        ws = s.get("walker_seeds")
        if ws is not None:
            buf = np.ascontiguousarray(ws, dtype=np.uint64)
            keep.append(buf)
            d.walker_seeds = buf.ctypes.data_as(C.POINTER(C.c_uint64))
    return arr


def _params(p):
    out = abi.Params()
    for k, v in p.items():
        setattr(out, k, v)
    return out


# SA_Parameters (simulated_annealing_parameters.cxx:9-20); max_blocks and
# chains have no counterpart there.
ANNEAL_DEFAULTS = dict(initial_temperature=25.0, final_temperature=0.25, num_temperatures=250, block_size=1000,
                       stdev_tolerance=0.01, frozen_criterion=0.0, max_blocks=100, chains=1)


def anneal_params(**kw):
    p = dict(ANNEAL_DEFAULTS)
    p.update(kw)
    return p


def _anneal_call(fn, systems, params, keep_spins):
    """Shared marshalling of rewl_anneal and its oracle restatement."""
    keep = []
    descs = _desc_array([dict(s) for s in systems], keep)
    ap = abi.AnnealParams()
    for k, v in params.items():
        setattr(ap, k, v)
    n = len(systems) * int(params["chains"])
    emin = np.empty(n, np.float64)
    efin = np.empty(n, np.float64)
    sw = np.empty(n, np.uint64)
    ndof = sum(int(s["L"]) ** 2 * (1 if s["model"] == abi.ISING2D else 2) for s in systems) * int(params["chains"])
    spins = np.empty(ndof, np.int8) if keep_spins else None
    rc = fn(descs, len(systems), C.byref(ap), emin.ctypes.data_as(C.POINTER(C.c_double)),
            efin.ctypes.data_as(C.POINTER(C.c_double)), sw.ctypes.data_as(C.POINTER(C.c_uint64)),
            spins.ctypes.data_as(C.POINTER(C.c_int8)) if keep_spins else None)
    return rc, dict(min_energy=emin, final_energy=efin, sweeps=sw, spins=spins)


def anneal(systems, params, spins=True):
    """Simulated annealing of params["chains"] chains per system on the GPU
    (rewl_anneal; Simulated_Annealer::simulate_annealing).  Returns arrays
    indexed system * chains + chain: min_energy, final_energy, sweeps, and the
    final spins in the reference layout, chains of each system in order."""
    lib = abi.load_engine()
    rc, out = _anneal_call(lib.rewl_anneal, systems, params, spins)
    abi.check(rc)
    return out


class Engine:
    """One rewl_ctx: every walker of the given systems on the current GPU."""

    def __init__(self, systems, params, autotune=False):
        """autotune: before creating the context, time the candidate launch
        shapes of this workload on scratch contexts (autotune.autotune: kernel
        variant, producer wavefronts, walkers per wavefront) and create it
        with the fastest; True, or a JSON path that caches the choice across
        processes.  Every shape runs the same chains."""
        self.lib = abi.load_engine()
        self._keep = []
        self.systems = [dict(s) for s in systems]
        self.params = dict(params)
        if autotune:
            from . import autotune as tuner
            best, _ = tuner.autotune(self.systems, self.params,
                                     cache_path=autotune if isinstance(autotune, str) else None)
            if best:
                self.params.update(best)
                self.tuned_shape = dict(best)
        descs = _desc_array(self.systems, self._keep)
        pp = _params(self.params)
        ctx = C.c_void_p()
        abi.check(self.lib.rewl_create(descs, len(systems), C.byref(pp), C.byref(ctx)))
        self.ctx = ctx
        n = C.c_int32()
        abi.check(self.lib.rewl_num_walkers(self.ctx, C.byref(n)))
        self.num_walkers = n.value
        abi.check(self.lib.rewl_num_windows(self.ctx, C.byref(n)))
        self.num_windows = n.value
        self.windows = [self.window_info(k) for k in range(self.num_windows)]
        self.walker_window = np.empty(self.num_walkers, dtype=np.int64)
        for k, w in enumerate(self.windows):
            R = self.systems[w["system"]]["replicas_per_window"]
            self.walker_window[w["first_walker"]:w["first_walker"] + R] = k

    def close(self):
        if self.ctx:
            self.lib.rewl_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- steps -------------------------------------------------------------
    def adjust_to_range(self):
        abi.check(self.lib.rewl_adjust_to_range(self.ctx))

    def sweeps(self, n):
        abi.check(self.lib.rewl_sweeps(self.ctx, int(n)))

    def exchange(self):
        abi.check(self.lib.rewl_exchange(self.ctx))

    def check(self):
        abi.check(self.lib.rewl_check(self.ctx))

    def run(self, max_checks):
        """Returns True when every window reached the final ln f."""
        rc = self.lib.rewl_run(self.ctx, int(max_checks))
        if rc == abi.REWL_ERR_ITER_CAP:
            return False
        abi.check(rc)
        return True

    def finalize(self):
        abi.check(self.lib.rewl_finalize(self.ctx))

    def refine(self, sweeps, sweeps_per_exchange=None, chunk=None, progress=None):
        """Multicanonical refinement of the finished run (rewl_refine): fixed
        weights 1/g for `sweeps` sweeps, then ln g += ln H per window.  With
        `chunk`, the sweeps run in parts of that many (rewl_refine_begin /
        _sweeps / _end, the same result) and progress(done) is called after
        each."""
        spe = self.params["sweeps_per_exchange"] if sweeps_per_exchange is None else sweeps_per_exchange
        if not chunk:
            abi.check(self.lib.rewl_refine(self.ctx, int(sweeps), int(spe)))
            return
        abi.check(self.lib.rewl_refine_begin(self.ctx))
        done = 0
        while done < sweeps:
            k = min(int(chunk), int(sweeps) - done)
            abi.check(self.lib.rewl_refine_sweeps(self.ctx, k, int(spe)))
            done += k
            if progress:
                progress(done)
        abi.check(self.lib.rewl_refine_end(self.ctx))

    def synchronize(self):
        abi.check(self.lib.rewl_synchronize(self.ctx))

    # ---- queries -------------------------------------------------------------
    def window_info(self, k):
        wi = abi.WindowInfo()
        abi.check(self.lib.rewl_get_window_info(self.ctx, int(k), C.byref(wi)))
        return {f: getattr(wi, f) for f, _ in abi.WindowInfo._fields_}

    def stats(self):
        st = abi.Stats()
        abi.check(self.lib.rewl_get_stats(self.ctx, C.byref(st)))
        return {f: getattr(st, f) for f, _ in abi.Stats._fields_}

    def _nbins(self, g):
        return self.windows[self.walker_window[g]]["num_bins"]

    def logdos(self, g):
        n = self._nbins(g)
        out = np.empty(n, dtype=np.float64)
        abi.check(self.lib.rewl_get_logdos(self.ctx, int(g), out.ctypes.data_as(C.POINTER(C.c_double)), n))
        return out

    def counts(self, g):
        n = self._nbins(g)
        out = np.empty(n, dtype=np.uint32)
        abi.check(self.lib.rewl_get_counts(self.ctx, int(g), out.ctypes.data_as(C.POINTER(C.c_uint32)), n))
        return out

    def visits(self, g):
        n = self._nbins(g)
        out = np.empty(n, dtype=np.uint64)
        abi.check(self.lib.rewl_get_visits(self.ctx, int(g), out.ctypes.data_as(C.POINTER(C.c_uint64)), n))
        return out

    def ndof(self, g):
        s = self.systems[self.windows[self.walker_window[g]]["system"]]
        return s["L"] ** 2 * (1 if s["model"] == abi.ISING2D else 2)

    def spins(self, g):
        n = self.ndof(g)
        out = np.empty(n, dtype=np.int8)
        abi.check(self.lib.rewl_get_spins(self.ctx, int(g), out.ctypes.data_as(C.POINTER(C.c_int8)), n))
        return out

    def energy(self, g):
        e = C.c_double()
        abi.check(self.lib.rewl_get_energy(self.ctx, int(g), C.byref(e)))
        return e.value

    def flavor(self, g):
        f = C.c_int32()
        abi.check(self.lib.rewl_get_flavor(self.ctx, int(g), C.byref(f)))
        return f.value

    def num_obs(self, model):
        corr = self.params.get("observables", 0) == 2
        if model == abi.ISING2D:
            return abi.NUM_OBS_ISING_CORR if corr else abi.NUM_OBS_ISING
        return abi.NUM_OBS_AT_CORR if corr else abi.NUM_OBS_AT

    def correlator_slots(self, model):
        """Columns that hold Fourier-correlator samples (observables = 2)."""
        if self.params.get("observables", 0) != 2:
            return []
        return [3] if model == abi.ISING2D else [3, 7, 11, 15]

    def observables(self, g):
        """Observable sums [bin, ob] of walker g (observables on); the mean of a
        bin is the sum over visits(g) + counts(g) of that bin, or for a
        correlator slot over ceil((visits + counts) / 1000) samples."""
        n = self._nbins(g)
        s = self.systems[self.windows[self.walker_window[g]]["system"]]
        nobs = self.num_obs(s["model"])
        out = np.empty((n, nobs), dtype=np.float64)
        abi.check(self.lib.rewl_get_observables(self.ctx, int(g), out.ctypes.data_as(C.POINTER(C.c_double)), n, nobs))
        return out

    def system_observables(self, system=0):
        """Per-bin observable means of one system on its global bin grid:
        every walker's sums pooled over the windows and replicas that hold the
        bin, divided by their pooled update counts (NaN where no walker has
        been).  The reference averages the windows' means by overlap count
        (concatenate_rewl_tables.hpp); pooling weights them by samples."""
        s = self.systems[system]
        nobs = self.num_obs(s["model"])
        corr = self.correlator_slots(s["model"])
        nb = int(round((s["e_max"] - s["e_min"]) / s["bin_size"]))
        tot = np.zeros((nb, nobs))
        cnt = np.zeros(nb)
        samples = np.zeros(nb)
        for w in self.windows:
            if w["system"] != system:
                continue
            g0 = int(round((w["e_min"] - s["e_min"]) / s["bin_size"]))
            for r in range(s["replicas_per_window"]):
                g = w["first_walker"] + r
                n = self.counts(g).astype(np.float64) + self.visits(g)
                tot[g0:g0 + w["num_bins"]] += self.observables(g)
                cnt[g0:g0 + w["num_bins"]] += n
                samples[g0:g0 + w["num_bins"]] += np.ceil(n / 1000.0)
        with np.errstate(invalid="ignore", divide="ignore"):
            means = tot / cnt[:, None]
            for c in corr:
                means[:, c] = tot[:, c] / samples
            return s["e_min"] + s["bin_size"] * np.arange(nb), means

    def densities(self, g):
        """Ashkin-Teller sigma-tau density counts [bin, 1024] of walker g
        (densities on): cell (tau axis) * 32 + (sigma axis)."""
        n = self._nbins(g)
        out = np.empty((n, abi.DENSITY_CELLS), dtype=np.uint32)
        abi.check(self.lib.rewl_get_densities(self.ctx, int(g), out.ctypes.data_as(C.POINTER(C.c_uint32)), n))
        return out

    def rng_cursor(self, g):
        c = C.c_uint64()
        abi.check(self.lib.rewl_get_rng_cursor(self.ctx, int(g), C.byref(c)))
        return c.value

    # ---- state import ----------------------------------------------------------
    def set_logdos(self, g, lng):
        buf = np.ascontiguousarray(lng, dtype=np.float64)
        abi.check(self.lib.rewl_set_logdos(self.ctx, int(g), buf.ctypes.data_as(C.POINTER(C.c_double)), buf.size))

    def set_window_lnf(self, k, lnf):
        abi.check(self.lib.rewl_set_window_lnf(self.ctx, int(k), float(lnf)))

    def set_spins(self, g, spins):
        buf = np.ascontiguousarray(spins, dtype=np.int8)
        abi.check(self.lib.rewl_set_spins(self.ctx, int(g), buf.ctypes.data_as(C.POINTER(C.c_int8)), buf.size))

    def set_system_spins(self, system, spins):
        """Every walker of one system starts from `spins` (import_DoFs after
        the master's broadcast, main.cpp:194-210); adjust before sweeping."""
        buf = np.ascontiguousarray(spins, dtype=np.int8)
        abi.check(self.lib.rewl_set_system_spins(self.ctx, int(system), buf.ctypes.data_as(C.POINTER(C.c_int8)),
                                                 buf.size))

    # ---- results ------------------------------------------------------------------
    def stitched(self, system=0, states_per_site=None, join="reference"):
        """Stitched ln g of one system; normalised to states_per_site^N when
        given.  join="reference": the reference's derivative-matching join
        (rewl_stitch); "overlap": the mean difference over each overlap
        (rewl_stitch_overlap, unvisited bins NaN)."""
        cap = sum(w["num_bins"] for w in self.windows if w["system"] == system)
        E = np.empty(cap, dtype=np.float64)
        lg = np.empty(cap, dtype=np.float64)
        n = C.c_int32()
        sps = -1.0 if states_per_site is None else float(states_per_site)
        fn = {"reference": self.lib.rewl_stitch, "overlap": self.lib.rewl_stitch_overlap}[join]
        abi.check(fn(self.ctx, int(system), sps, E.ctypes.data_as(C.POINTER(C.c_double)),
                     lg.ctypes.data_as(C.POINTER(C.c_double)), cap, C.byref(n)))
        return E[:n.value].copy(), lg[:n.value].copy()


def thermodynamics(energies, logdos, system_size, t_min=0.01, t_max=50.01, num_t=50000):
    """FP64 canonical averages on the GPU: columns F/N, U/N, U^2/N^2, S/N, C/N."""
    lib = abi.load_engine()
    E = np.ascontiguousarray(energies, dtype=np.float64)
    lg = np.ascontiguousarray(logdos, dtype=np.float64)
    out = np.empty((num_t, 5), dtype=np.float64)
    abi.check(lib.rewl_thermodynamics(E.ctypes.data_as(C.POINTER(C.c_double)),
                                      lg.ctypes.data_as(C.POINTER(C.c_double)), E.size, int(system_size),
                                      float(t_min), float(t_max), int(num_t),
                                      out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def device_count():
    lib = abi.load_engine()
    n = C.c_int()
    lib.rewl_device_count(C.byref(n))
    return n.value


def thermo_observables(energies, logdos, observables, t_min=0.01, t_max=50.01, num_t=50000):
    """Canonical averages [T, ob] of per-bin observable means on the GPU."""
    lib = abi.load_engine()
    E = np.ascontiguousarray(energies, dtype=np.float64)
    lg = np.ascontiguousarray(logdos, dtype=np.float64)
    ob = np.ascontiguousarray(observables, dtype=np.float64).reshape(len(E), -1)
    out = np.empty((num_t, ob.shape[1]), dtype=np.float64)
    dp = C.POINTER(C.c_double)
    abi.check(lib.rewl_thermo_observables(E.ctypes.data_as(dp), lg.ctypes.data_as(dp), ob.ctypes.data_as(dp),
                                          len(E), ob.shape[1], t_min, t_max, num_t, out.ctypes.data_as(dp)))
    return out
