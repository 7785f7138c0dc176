"""Learned warm start for batched random-strain Ashkin-Teller REWL.

Pipeline (every step on the GPU):

  1. ``generate``: converged ln g curves of random-strain AT realisations
     (the batched configuration: 8 windows x 4 walkers, 512 bins over
     [E(all-up), 0), visited-support flatness 0.8, ln f halved to
     1e-6), many realisations per context, stitched and normalised to
     sum g = 4^N.
  2. ``train``: the MLP of include/rewl_mlp.h, inputs (E/N, 1/L, K, sigma),
     target ln g/N, trained with the library's HIP forward/backward/Adam
     kernels on every visited bin of every curve.
  3. ``evaluate``: held-out realisations (new disorder seeds) run from flat
     ln g at ln f = 1 (plain WL, the fallback) and from the network's ln g at
     ln f = 1e-3 (warm start), both to ln f < 1e-6; wall time of each run and
     the ln g error of each against a reference run taken further (ln f
     1e-8).  ln g errors are max |difference| over the visited bins after a
     common anchor, per realisation.

    python -m rewl_simulator_amd.warmstart --out DIR [--quick]
"""
import argparse
import json
import os
import time

import numpy as np

from . import abi, configs
from .engine import Engine
from .mlp import MLP, features, warm_start

K_AT = 0.5
STATES_PER_SITE = 4.0


def rfat_systems(specs, bins=512, windows=8, replicas=4):
    """One random-strain AT system per (L, sigma, seed), over [E(all-up), 0):
    the reference's default half range, which holds every energy that carries
    weight at positive temperature.  The upper half ends in the checkerboard
    states, single configurations a gap away from the rest of the spectrum;
    once one replica has visited one, visited-support flatness waits for
    every replica to revisit it in every stage."""
    systems = []
    for L, sigma, seed in specs:
        h = configs.gaussian_strain(L, sigma, seed)
        e0, _ = configs.rfat_range(L, h, K=K_AT)
        bs = (0.0 - e0) / bins
        systems.append(configs.at_system(L, e0, 0.0, bs, windows, replicas, K=K_AT, field=h, seed=seed))
    return systems


def rfat_params(lnf_final=1e-6, sweeps_per_check=250, average_mode=1):
    return configs.params(flatness_mode=abi.FLAT_MIN_OVER_MEAN_VISITED, flatness=0.8, schedule=abi.SCHED_HALVE,
                          lnf_final=lnf_final, sweeps_per_check=sweeps_per_check,
                          sweeps_per_exchange=sweeps_per_check, adjust_mode=1, average_mode=average_mode)


def stitch_visited(windows, e_min, bin_size):
    """Join window curves on the system's global bin grid, over visited bins
    only.  `windows` holds (g0, ln g, visited) per window in energy order (g0
    = the window's first global bin).  Each window is shifted onto the curve
    so far by the mean difference over the bins both have visited, and takes
    over from the middle of that common stretch.  The reference's
    derivative-matching join (rewl_stitch) assumes every bin of the overlap
    holds states; clean systems with bins finer than their level spacing have
    empty bins there, whose ln g is no value at all.  Returns (E, ln g,
    visited); unvisited bins hold NaN."""
    nb = max(g0 + len(l) for g0, l, _ in windows)
    lng = np.full(nb, np.nan)
    vis = np.zeros(nb, bool)
    for k, (g0, lk, vk) in enumerate(windows):
        idx = g0 + np.arange(len(lk))
        if k == 0:
            lng[idx[vk]] = lk[vk]
            vis[idx[vk]] = True
            continue
        common = vk & vis[idx]
        if not common.any():
            raise ValueError(f"window {k} shares no visited bin with the windows below it")
        shift = float(np.mean(lng[idx[common]] - lk[common]))
        take = vk & (idx >= int(np.median(idx[common])))
        lng[idx[take]] = lk[take] + shift
        vis[idx[take]] = True
    E = e_min + bin_size * np.arange(nb)
    return E, lng, vis


def normalise(lng, vis, N):
    """Shift ln g so that the visited bins hold STATES_PER_SITE^N states."""
    v = lng[vis]
    m = v.max()
    return lng - (m + np.log(np.exp(v - m).sum()) - N * np.log(STATES_PER_SITE))


def system_curve(eng, system):
    """Stitched, normalised (E, ln g, visited) of one system after finalize:
    the replicas of a window are equal then, so replica 0 carries the window."""
    s = eng.systems[system]
    parts = []
    for w in eng.windows:
        if w["system"] != system:
            continue
        v = np.zeros(w["num_bins"], dtype=np.uint64)
        for r in range(s["replicas_per_window"]):
            v += eng.visits(w["first_walker"] + r) + eng.counts(w["first_walker"] + r).astype(np.uint64)
        g0 = int(round((w["e_min"] - s["e_min"]) / s["bin_size"]))
        parts.append((g0, eng.logdos(w["first_walker"]), v > 0))
    E, lng, vis = stitch_visited(parts, s["e_min"], s["bin_size"])
    return E, normalise(lng, vis, s["L"] ** 2), vis


def curves(eng, specs):
    """(L, sigma, E, ln g, visited) per system, ln g normalised to 4^N over
    the visited bins."""
    out = []
    for i, (L, sigma, seed) in enumerate(specs):
        E, lng, vis = system_curve(eng, i)
        out.append(dict(L=L, sigma=sigma, seed=seed, E=E, lng=lng, visited=vis))
    return out


def run_until_done(eng, max_checks, log=None, every=50, name="", max_seconds=None):
    """Run check cycles until every window is at its final ln f; returns
    (converged, seconds).  Runs in chunks of `every` checks so that progress
    (the slowest window's ln f) can be logged; with max_seconds, no new chunk
    starts after that much wall time."""
    eng.synchronize()
    t0 = time.perf_counter()
    done, checks = False, 0
    while checks < max_checks and not done and (max_seconds is None or time.perf_counter() - t0 < max_seconds):
        k = min(every, max_checks - checks)
        done = eng.run(k)
        checks += k
        if log is not None:
            lnf = max(eng.window_info(w)["lnf"] for w in range(eng.num_windows))
            log(f"  {name}: {checks} checks, slowest ln f {lnf:.2e}, {time.perf_counter() - t0:.1f} s")
    eng.synchronize()
    return done, time.perf_counter() - t0


def generate(specs, lnf_final=1e-6, batch=256, max_checks=20000, log=print, max_seconds=None, bins=512):
    """Converged curves for every spec, `batch` realisations per context
    (with max_seconds, each context stops after that much wall time and its
    windows keep the ln g they have)."""
    out = []
    for b in range(0, len(specs), batch):
        part = specs[b:b + batch]
        t0 = time.perf_counter()
        with Engine(rfat_systems(part, bins=bins), rfat_params(lnf_final)) as eng:
            eng.adjust_to_range()
            done, secs = run_until_done(eng, max_checks, log=log, name="generate", max_seconds=max_seconds)
            eng.finalize()
            out.extend(curves(eng, part))
            st = eng.stats()
        log(f"generate: {b + len(part)}/{len(specs)} curves, converged={done}, {secs:.1f} s run, "
            f"{st['sweeps']} sweeps, {time.perf_counter() - t0:.1f} s total")
    return out


def dataset(curve_list):
    xs, ts = [], []
    for c in curve_list:
        L = c["L"]
        N = L * L
        m = c["visited"]
        xs.append(features(c["E"][m], N, L, c.get("K", K_AT), c["sigma"]))
        ts.append(c["lng"][m] / N)
    return np.concatenate(xs), np.concatenate(ts)


def train(curve_list, steps=4000, batch=16384, seed=3, log=print):
    x, t = dataset(curve_list)
    mlp = MLP(seed=seed)
    mlp.fit_standardisation(x, t)
    hist = mlp.train(x, t, steps=steps, batch=batch, lr=2e-3, lr_final=1e-5, seed=seed + 1, warmup=100)
    log(f"train: {len(x)} points, {steps} steps, {mlp.last_ms():.0f} ms on the GPU, "
        f"loss {hist[:10].mean():.3e} -> {hist[-50:].mean():.3e}")
    return mlp, hist


def predict_lng(mlp, E, L, sigma):
    N = L * L
    return N * mlp.predict(features(E, N, L, K_AT, sigma))


def anchored_max_error(a, b, mask):
    d = (a - b)[mask]
    d = d - d[0]
    return float(np.max(np.abs(d)))


def evaluate(mlp, specs, lnf_final=1e-6, lnf_warm=1e-3, ref_lnf=1e-8, max_checks=40000, log=print, bins=512):
    """Plain (flat ln g, ln f = 1) and warm-started (network ln g, ln f =
    lnf_warm) runs of the same held-out systems to ln f < lnf_final.  The
    plain result is the yardstick of the warm one (max |difference| over commonly visited
    bins, anchored); with ref_lnf > 0 both are also compared to a plain run
    taken to ref_lnf."""
    systems = rfat_systems(specs, bins=bins)
    res = {}
    ref = None
    if ref_lnf > 0:
        # the same disorder with independent walker streams: a reference that
        # shares the compared runs' streams would repeat their errors
        ref_systems = [dict(s, seed=s["seed"] + 0x9E3779B9) for s in systems]
        with Engine(ref_systems, rfat_params(ref_lnf)) as eng:
            eng.adjust_to_range()
            done, secs = run_until_done(eng, max_checks, log=log, name="reference")
            eng.finalize()
            ref = curves(eng, specs)
        log(f"evaluate: reference run to ln f {ref_lnf:g}: converged={done}, {secs:.1f} s")
        res["reference"] = dict(lnf=ref_lnf, seconds=secs, converged=bool(done))
    out = {}
    for name in ("plain", "warm"):
        with Engine(systems, rfat_params(lnf_final)) as eng:
            eng.adjust_to_range()
            t_init = 0.0
            if name == "warm":
                t0 = time.perf_counter()
                warm_start(eng, mlp, lambda i: specs[i][1], lnf_warm)
                t_init = time.perf_counter() - t0
            done, secs = run_until_done(eng, max_checks, log=log, name=name)
            st = eng.stats()
            eng.finalize()
            out[name] = curves(eng, specs)
        res[name] = dict(seconds=secs, init_seconds=t_init, converged=bool(done), sweeps=st["sweeps"],
                         flat_events=st["flat_events"], lnf_start=lnf_warm if name == "warm" else 1.0)
        if ref is not None:
            errs = [anchored_max_error(g["lng"], r["lng"], r["visited"] & g["visited"]) for g, r in zip(out[name], ref)]
            res[name].update(lng_error_max=float(np.max(errs)), lng_error_median=float(np.median(errs)))
        log(f"evaluate: {name}: {secs:.1f} s (+{t_init:.2f} s init), converged={done}, {st['sweeps']} sweeps")
    dev = [anchored_max_error(w["lng"], p["lng"], w["visited"] & p["visited"]) for w, p in zip(out["warm"], out["plain"])]
    res["warm_vs_plain_max"] = float(np.max(dev))
    res["warm_vs_plain_median"] = float(np.median(dev))
    # the network's own error before any sampling, against the best curve at hand
    yardstick = ref if ref is not None else out["plain"]
    pe = [anchored_max_error(predict_lng(mlp, r["E"], L, sigma), r["lng"], r["visited"])
          for r, (L, sigma, seed) in zip(yardstick, specs)]
    res["prediction_error_max"] = float(np.max(pe))
    res["prediction_error_median"] = float(np.median(pe))
    res["_reference_curves"] = ref
    res["speedup_wall"] = res["plain"]["seconds"] / (res["warm"]["seconds"] + res["warm"]["init_seconds"])
    res["speedup_sweeps"] = res["plain"]["sweeps"] / max(res["warm"]["sweeps"], 1)
    log(f"evaluate: speedup {res['speedup_wall']:.2f}x wall, {res['speedup_sweeps']:.2f}x sweeps; "
        f"warm vs plain max |d ln g| {res['warm_vs_plain_max']:.3f}; network error {res['prediction_error_max']:.3f}")
    return res


def evaluate_refine(mlp, specs, ref, plain, first=2000, sweeps_per_exchange=250, log=print, bins=512):
    """The network's ln g as fixed weights from the start: rounds of
    multicanonical refinement (rewl_refine; each samples with ln f = 0 and
    then corrects ln g by ln H), the round length doubling from `first`,
    until the ln g error against the reference (max and median over the
    realisations, anchored, over commonly visited bins) is no worse than
    plain Wang-Landau's at its final ln f, or the sweeps reach plain's.  The
    network's prediction replaces the whole Wang-Landau run: without it, the
    flat start's fixed weights would leave the walkers at the most probable
    energies."""
    systems = rfat_systems(specs, bins=bins)
    rounds = []
    with Engine(systems, rfat_params(1e-6)) as eng:
        eng.adjust_to_range()
        t0 = time.perf_counter()
        warm_start(eng, mlp, lambda i: specs[i][1], 1e-3)
        t_init = time.perf_counter() - t0
        total, r, secs = 0, int(first), 0.0
        while total < plain["sweeps"]:
            t1 = time.perf_counter()
            eng.refine(r, sweeps_per_exchange)
            secs += time.perf_counter() - t1
            total += r
            c = curves(eng, specs)
            errs = [anchored_max_error(g["lng"], rr["lng"], rr["visited"] & g["visited"]) for g, rr in zip(c, ref)]
            rounds.append(dict(sweeps=total, seconds=secs, lng_error_max=float(np.max(errs)),
                               lng_error_median=float(np.median(errs))))
            log(f"refine from the network: {total} sweeps, {secs:.1f} s, ln g error max {np.max(errs):.3f}, "
                f"median {np.median(errs):.3f}")
            if np.max(errs) <= plain["lng_error_max"] and np.median(errs) <= plain["lng_error_median"]:
                break
            r *= 2
    last = rounds[-1]
    reached = last["lng_error_max"] <= plain["lng_error_max"] and last["lng_error_median"] <= plain["lng_error_median"]
    return dict(rounds=rounds, init_seconds=t_init, reached_plain_error=bool(reached),
                speedup_sweeps=plain["sweeps"] / last["sweeps"],
                speedup_wall=plain["seconds"] / (last["seconds"] + t_init))


def budget_run(mlp, specs, checks, lnf_warm=1e-3, log=print, bins=512):
    """Plain and warm-started runs of the same systems for a fixed number of
    check cycles: the slowest and median window ln f each reaches."""
    systems = rfat_systems(specs, bins=bins)
    res = {}
    for name in ("plain", "warm"):
        with Engine(systems, rfat_params(1e-6)) as eng:
            eng.adjust_to_range()
            if name == "warm":
                warm_start(eng, mlp, lambda i: specs[i][1], lnf_warm)
            done, secs = run_until_done(eng, checks, log=log, name=name + " (budget)")
            lnf = np.array([eng.window_info(k)["lnf"] for k in range(eng.num_windows)])
        res[name] = dict(checks=checks, seconds=secs, converged=bool(done), slowest_lnf=float(lnf.max()),
                         median_lnf=float(np.median(lnf)), frac_windows_done=float((lnf < 1e-6).mean()))
        log(f"budget: {name}: slowest ln f {lnf.max():.2e}, median {np.median(lnf):.2e}, "
            f"{(lnf < 1e-6).mean():.2f} of windows done after {checks} checks")
    return res


def record_stages(reach, lnf, sweeps):
    """reach[k, j] = sweeps for every window k whose ln f is now at or below
    2^-j and that has no entry for j yet (j up to reach.shape[1] - 1)."""
    stage = np.minimum(np.floor(-np.log2(np.asarray(lnf, dtype=np.float64)) + 1e-9).astype(np.int64),
                       reach.shape[1] - 1)
    for k in range(reach.shape[0]):
        row = reach[k, :stage[k] + 1]
        row[row < 0] = sweeps


def stage_trace(specs, seconds, mlp=None, lnf_warm=1e-3, every=20, log=print, windows=8, bins=512):
    """Plain (flat ln g, ln f = 1) or, with `mlp`, warm-started (network ln g,
    ln f = lnf_warm) Wang-Landau on the systems for a wall-clock budget,
    recording per window the sweeps after which its ln f was first at or
    below 2^-j, j = 0..20 (2^-20 < 1e-6, the final ln f): the time each
    window spends in each stage.  reach[k, j] = -1 where window k has not got
    there; j at or below the starting stage gets the first check's sweeps."""
    systems = rfat_systems(specs, bins=bins, windows=windows)
    params = rfat_params(1e-6)
    spc = params["sweeps_per_check"]
    with Engine(systems, params) as eng:
        eng.adjust_to_range()
        if mlp is not None:
            warm_start(eng, mlp, lambda i: specs[i][1], lnf_warm)
        nw = eng.num_windows
        index = np.array([w["window"] for w in eng.windows])
        system = np.array([w["system"] for w in eng.windows])
        reach = np.full((nw, 21), -1, dtype=np.int64)
        eng.synchronize()
        t0 = time.perf_counter()
        checks, done = 0, False
        while not done and time.perf_counter() - t0 < seconds:
            done = eng.run(every)
            checks += every
            lnf = np.array([eng.window_info(k)["lnf"] for k in range(nw)])
            record_stages(reach, lnf, checks * spc)
            slow = ", ".join(f"{lnf[index == w].max():.1e}" for w in range(index.max() + 1))
            log(f"stage trace: {checks * spc} sweeps, {time.perf_counter() - t0:.0f} s; slowest ln f per window "
                f"index [{slow}]; windows done {(lnf < 1e-6).mean():.3f}")
        eng.synchronize()
        secs = time.perf_counter() - t0
    summary = []
    for w in range(index.max() + 1):
        sel = index == w
        row = {}
        for j in (1, 10, 20):
            got = reach[sel, j][reach[sel, j] >= 0]
            row[f"lnf_2^-{j}"] = dict(windows=int(got.size), of=int(sel.sum()),
                                      median_sweeps=float(np.median(got)) if got.size else None)
        summary.append(row)
    return dict(sweeps=checks * spc, seconds=secs, converged=bool(done), start="flat" if mlp is None else "network",
                per_window_index=summary, reach=reach.tolist(), window=index.tolist(), system=system.tolist())


def ising_systems(sizes=(16, 24, 32, 48), bin_size=8.0, replicas=16):
    """Clean 2D Ising at each L in the layout of the L = 64 workload: bins of
    width 8 over the full spectrum (top level included, unreachable levels
    excluded), windows of 75% overlap, L / 2 of them."""
    return [configs.ising_full_range(L, bin_size, max(4, L // 2), replicas) for L in sizes]


def ising_params(lnf_final=1e-6):
    return configs.params(flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE,
                          lnf_final=lnf_final, sweeps_per_check=1000, sweeps_per_exchange=1000, average_mode=1)


def ising_curves(sizes=(16, 24, 32, 48), lnf_final=1e-6, max_checks=5000, log=print):
    """Engine-produced ln g curves of the clean 2D Ising model at each L (one
    context, every L run to ln f < lnf_final), stitched by the reference's
    join and normalised to 2^N, as training curves (K = 0, sigma = 0)."""
    systems = ising_systems(sizes)
    t0 = time.perf_counter()
    with Engine(systems, ising_params(lnf_final)) as eng:
        eng.adjust_to_range()
        done, secs = run_until_done(eng, max_checks, log=log, every=20, name="ising curves")
        eng.finalize()
        out = []
        for i, L in enumerate(sizes):
            E, lng = eng.stitched(i, states_per_site=2.0)
            out.append(dict(L=L, K=0.0, sigma=0.0, seed=0, E=E, lng=lng, visited=np.isfinite(lng)))
    log(f"ising curves: L = {list(sizes)}, converged={done}, {secs:.1f} s run, {time.perf_counter() - t0:.1f} s total")
    return out, done, secs


def training_specs(n, sizes=(16, 24, 32, 48), sigmas=(0.0, 0.25, 0.5, 0.75, 1.0), seed0=50000):
    """n (L, sigma, seed) triples cycling sizes and disorder strengths; the
    disorder seeds are disjoint from the held-out ones."""
    return [(sizes[i % len(sizes)], sigmas[(i // len(sizes)) % len(sigmas)], seed0 + i) for i in range(n)]


def config5(n=64, bins=512, seed0=900000):
    """The learned-warm-start workload: n held-out random-strain AT
    realisations at L = 48 (sigma 0.5), 8 windows x 4 walkers, `bins` bins
    over [E(all-up), 0), visited-support flatness 0.8, halving to ln f 1e-6;
    warm-started windows begin at ln f = 1e-3.  At L = 48 bins of 22.5 (the
    stated 512) are wider than one excitation at the low-energy end and stall
    the lowest windows; 1024 bins do not.  Returns (specs, systems, params)."""
    specs = heldout_specs(n, L=48, sigma=0.5, seed0=seed0)
    return specs, rfat_systems(specs, bins=bins), rfat_params(1e-6)


def heldout_specs(n, L=48, sigma=0.5, seed0=900000):
    return [(L, sigma, seed0 + i) for i in range(n)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="warmstart_out")
    ap.add_argument("--curves", type=int, default=256)
    ap.add_argument("--sizes", default="12,16,20,24")
    ap.add_argument("--sigmas", default="0,0.25,0.5,0.75,1")
    ap.add_argument("--heldout", type=int, default=64)
    ap.add_argument("--heldout-L", type=int, default=24)
    ap.add_argument("--lnf-final", type=float, default=1e-6)
    ap.add_argument("--ref-lnf", type=float, default=0.0)
    ap.add_argument("--lnf-warm", type=float, default=1e-3, help="ln f the warm-started windows begin at")
    ap.add_argument("--mlp-in", default="", help="load this trained network (mlp.npz) instead of generating and training")
    ap.add_argument("--no-eval", action="store_true", help="generate and train only")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--budget-L", type=int, default=0, help="also a fixed-budget comparison at this L")
    ap.add_argument("--budget-checks", type=int, default=200)
    ap.add_argument("--heldout-seed0", type=int, default=900000, help="disorder seed of the first held-out realisation")
    ap.add_argument("--generate-seconds", type=float, default=0.0,
                    help="stop generating after this much wall time (windows keep the ln g they have)")
    ap.add_argument("--stage-trace", type=float, default=0.0, metavar="SECONDS",
                    help="instead of the comparison: per-window stage times of the held-out realisations for "
                         "SECONDS, from a flat ln g at ln f = 1 or from the network's at --lnf-warm (--stage-start)")
    ap.add_argument("--stage-start", choices=("flat", "network"), default="network")
    ap.add_argument("--stage-windows", type=int, default=8, help="windows per held-out realisation in --stage-trace")
    ap.add_argument("--bins", type=int, default=512, help="bins per realisation over [E(all-up), 0)")
    ap.add_argument("--stage-bins", type=int, default=0, help="bins per held-out realisation in --stage-trace "
                                                             "(default: --bins)")
    ap.add_argument("--refine-first", type=int, default=0, metavar="SWEEPS",
                    help="also run the network's ln g through rounds of multicanonical refinement (first round "
                         "SWEEPS, doubling) until it matches plain Wang-Landau's error (needs --ref-lnf)")
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    logf = open(os.path.join(a.out, "log.txt"), "a")

    def log(msg):
        print(msg, flush=True)
        logf.write(msg + "\n")
        logf.flush()

    sizes = tuple(int(v) for v in a.sizes.split(","))
    sigmas = tuple(float(v) for v in a.sigmas.split(","))
    train_specs = training_specs(a.curves, sizes=sizes, sigmas=sigmas)
    held = heldout_specs(a.heldout, L=a.heldout_L, seed0=a.heldout_seed0)
    if a.stage_trace and a.stage_start == "flat":
        res = stage_trace(held, a.stage_trace, log=log, windows=a.stage_windows, bins=a.stage_bins or a.bins)
        log(json.dumps(res["per_window_index"]))
        json.dump(res, open(os.path.join(a.out, "stage_trace_flat.json"), "w"))
        return
    if a.mlp_in:
        mlp = MLP()
        mlp.load(a.mlp_in)
        info = dict(mlp_in=a.mlp_in)
    else:
        t0 = time.perf_counter()
        cl = generate(train_specs, lnf_final=a.lnf_final, log=log, max_seconds=a.generate_seconds or None, bins=a.bins)
        gen_s = time.perf_counter() - t0
        mlp, hist = train(cl, steps=a.steps, log=log)
        info = dict(generate_seconds=gen_s, train_ms=mlp.last_ms(), curves=len(cl),
                    train_points=int(sum(c["visited"].sum() for c in cl)), train_sizes=list(sizes))
        mlp.save(os.path.join(a.out, "mlp.npz"))
        if a.no_eval:
            log(json.dumps(info))
            return
    if a.stage_trace:
        res = stage_trace(held, a.stage_trace, mlp=mlp, lnf_warm=a.lnf_warm, log=log, windows=a.stage_windows,
                          bins=a.stage_bins or a.bins)
        res.update(info)
        log(json.dumps(res["per_window_index"]))
        json.dump(res, open(os.path.join(a.out, "stage_trace_network.json"), "w"))
        return
    res = evaluate(mlp, held, lnf_final=a.lnf_final, lnf_warm=a.lnf_warm, ref_lnf=a.ref_lnf, log=log, bins=a.bins)
    ref = res.pop("_reference_curves")
    if a.refine_first and ref is not None:
        res["refine_from_network"] = evaluate_refine(mlp, held, ref, res["plain"], first=a.refine_first, log=log,
                                                      bins=a.bins)
    res.update(info)
    res["heldout"] = dict(count=len(held), L=held[0][0], sigma=held[0][1])
    if a.budget_L:
        res["budget"] = budget_run(mlp, heldout_specs(a.heldout, L=a.budget_L), a.budget_checks, log=log, bins=a.bins)
        res["budget"]["L"] = a.budget_L
    log(json.dumps(res))
    json.dump(res, open(os.path.join(a.out, "result.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
