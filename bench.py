"""Benchmark: spin-flip attempts/s of the REWL walker sweep on one MI355X.

Workload: 2D Ising, L = 64 periodic (N = 4096), 2049 bins of width 8 over
[-8192, 8192], 32 windows at 75% overlap x 64 walkers = 2048 walkers, one lane
per walker, flatness min >= 0.8 mean, ln f halved toward 1e-8.
A step is one check cycle: sweeps_per_check sweeps of every walker, the replica
exchange and the flatness/averaging check.  Inputs (lattices, histograms) are
resident in HBM before timing starts.

Prints one JSON line with the throughput, the sweep kernel's roofline figures
(HIP-event timing of the kernel on the engine's own stream) and a bounded
CPU-baseline sample of the oracle on the host cores.
"""
import argparse
import json
import os
import signal
import subprocess
import sys
import tempfile
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# LDS, where the histograms live: nominal 256 CUs x 128 B/clk x 2.4 GHz; the
# bench measures the device's own figure (rewl_selfcheck(2)) and uses that
LDS_NOMINAL_GBS = 256 * 128 * 2.4
BYTES_PER_ATTEMPT = 32.25   # 8 B read ln g[new] + 16 B RMW ln g[cur] + 8 B RMW count + 0.25 B lattice
# The reference program as built by oracle/ref_build.sh (rewl_simulator_amd/build.py):
# at this workload's cadence without the correlator (the headline baseline),
# and with its defaults.
REFERENCE_RUNS = {
    "workload": (os.path.join(ROOT, "oracle", "_ref", "rewl_ref_ising_L64_plain"),
                 "an exchange every 1000 sweeps, CORRELATION_LENGTHS off; it still updates its magnetisation "
                 "moments at every attempt"),
    "defaults": (os.path.join(ROOT, "oracle", "_ref", "rewl_ref_ising_L64"),
                 "its defaults: an exchange every sweep, the correlator on"),
}
MPIEXEC = "/opt/conda/bin/mpiexec"
# Measured HBM bytes of one sweep launch of this workload with the default
# kernel, keyed by sweeps per launch: rocprofv3 --pmc FETCH_SIZE (x 2.0) and
# WRITE_SIZE (x 1.0), the scales read off a plain 256 MiB copy
# (profiles/k_sweep_tuned_ising_L64_counters.md).
MEASURED_TRAFFIC = {1000: (2.0 * 5618.125 + 1.0 * 7537.90625) * 1024}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sweeps-per-step", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target length of the CPU-baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU-baseline ranks / threads; 0 = the host CPUs this process may use")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-reference", action="store_true",
                    help="time only the oracle as the CPU baseline, not the reference program")
    ap.add_argument("--cadence-sweeps", type=int, default=200,
                    help="also time this many sweeps with an exchange after every sweep (the reference's cadence); "
                         "0 to skip")
    ap.add_argument("--walkers-per-group", type=int, default=0)
    ap.add_argument("--producers", type=int, default=0, help="producer wavefronts (0 auto, -1 none)")
    ap.add_argument("--sweep-kernel", type=int, default=0,
                    help="0 auto, 1 row kernel, 2 row-unrolled, 3 + slot speculation, 4 + read-ahead")
    ap.add_argument("--saturate", type=int, default=4, metavar="R",
                    help="also time R independent copies of the workload in one context (R x 2048 walkers, "
                         "enough to fill the chip); 0 to skip")
    ap.add_argument("--workload", default="config2", choices=["config2", "config3", "config4"],
                    help="config2: L=64 Ising REWL (the headline); config3: L=32 Ashkin-Teller REWL, 512 "
                         "walkers; config4: 256 batched random-strain Ashkin-Teller realisations in one context")
    ap.add_argument("--observables", type=int, default=0, choices=[0, 1, 2],
                    help="per-bin observables during the sweeps (1: moments, 2: with the Fourier correlator); "
                         "the headline figure is measured with 0, as the workload's reference runs them")
    ap.add_argument("--converge", type=float, default=0.0, metavar="LNF",
                    help="instead of the throughput steps, time the run from the first sweep until every "
                         "window's ln f is below LNF")
    ap.add_argument("--max-checks", type=int, default=200000)
    ap.add_argument("--time-to-lnf", type=float, default=1e-6, metavar="LNF",
                    help="also time the workload from ln f = 1 to LNF and report it in the line (the metric's "
                         "second half); 0 to skip")
    ap.add_argument("--warm-start-lnf", type=float, default=0.0, metavar="LNF",
                    help="also run the learned warm start end to end at a small size (training curves, the "
                         "network, held-out random-strain realisations at L = 12, plain vs warm-started to LNF, "
                         "about a minute) and report the factor in the line; 0 (the default) to skip")
    ap.add_argument("--no-warm-start-config2", dest="warm_start_config2", action="store_false",
                    help="skip the learned warm start of this workload (by default, with --time-to-lnf, the line also "
                         "times this workload to that target from the network's ln g at ln f = 1e-3, the network "
                         "trained on engine-produced clean-Ising curves at L = 16-48: about three minutes)")
    ap.add_argument("--warm-start-source", choices=("network", "exact"), default="network",
                    help="the warm start of this workload seeds from the network (the learned warm start) or from the "
                         "exact DoS (the bound any start at ln f = 1e-3 can reach)")
    ap.add_argument("--refine-copies", type=int, default=1, metavar="R",
                    help="with --refine: refine in a context with R times the replicas per window, every walker "
                         "seeded with the converged ln g and configurations")
    ap.add_argument("--refine", type=int, default=0, metavar="SWEEPS",
                    help="with --converge: then SWEEPS sweeps of multicanonical refinement (rewl_refine) and "
                         "the ln g error after it")
    ap.add_argument("--no-retire", action="store_true",
                    help="with --workload config4 --converge: keep finished realisations in the launches")
    ap.add_argument("--half-range", action="store_true",
                    help="with --workload config3 --converge: the reference's own Ashkin-Teller range [-5N, 0) "
                         "instead of the full spectrum")
    ap.add_argument("--average-mode", type=int, default=0,
                    help="replica averaging: 0 count-weighted (the reference), 1 equal weights")
    ap.add_argument("--dump", default="", metavar="PATH",
                    help="with --converge: also save E, ln g and the exact ln g to PATH (.npz)")
    ap.add_argument("--one-over-t", action="store_true",
                    help="the reference's ONE_OVER_T_ALGORITHM ln f schedule instead of halving")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (counter calibration) and exit")
    return ap.parse_args()


def calibrate_copy(nbytes):
    """A plain copy kernel of known size: rocprof's FETCH_SIZE / WRITE_SIZE of
    this launch give the scale that converts those counters into bytes."""
    import torch
    n = nbytes // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert torch.equal(dst[:16], src[:16])
    print(json.dumps({"calibration_copy_bytes": n * 4, "launches": 3}), flush=True)


def host_cpus():
    """The host CPUs this process may use: the cgroup CPU quota when one is set
    (this machine's share of its host), else the affinity mask; with nproc and
    the CPU model for the record."""
    info = {"os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    info["cgroup_quota"] = quota
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        info["nproc"] = None
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("CPU(s):"):
                info["lscpu_cpus"] = int(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    info["used"] = min(info["affinity"], quota) if quota else info["affinity"]
    return info


def reference_cpu_baseline(target_s, cores, run="workload"):
    """The reference program itself (built from its sources by
    oracle/ref_build.sh: L = 64 Ising, one walker per rank, the 1/t schedule;
    REFERENCE_RUNS names the two builds) under MPICH's mpiexec with one rank
    per host core, stopped after about target_s seconds.  Its rank 0 prints
    each iteration's update rate (sweeps x N / seconds,
    rewl_simulation.hpp:600-614); every rank sweeps as often, so the job's rate
    is that figure times the rank count.  None when the binary or MPICH is
    missing."""
    exe, what = REFERENCE_RUNS[run]
    if not (os.path.exists(exe) and os.path.exists(MPIEXEC)):
        return None
    rates = []
    with tempfile.TemporaryDirectory() as d:
        work = os.path.join(d, "work")
        os.makedirs(work)
        proc = subprocess.Popen([MPIEXEC, "-launcher", "fork", "-n", str(cores), exe], cwd=work,
                                stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, start_new_session=True)
        lines = []
        reader = threading.Thread(target=lambda: lines.extend(proc.stdout), daemon=True)
        reader.start()
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < target_s and proc.poll() is None:
            time.sleep(0.5)
        for sig, wait in ((signal.SIGINT, 5), (signal.SIGKILL, 5)):
            if proc.poll() is None:
                try:
                    os.killpg(proc.pid, sig)
                except ProcessLookupError:
                    pass
                try:
                    proc.wait(timeout=wait)
                except subprocess.TimeoutExpired:
                    pass
        reader.join(timeout=5)
        elapsed = time.perf_counter() - t0
    for line in lines:
        if "Update rate" in line:
            rates.append(float(line.split("=")[1].split()[0]))
    if len(rates) < 2:
        return None
    per_rank = sum(rates[1:]) / len(rates[1:])   # the first iteration includes the start-up
    return dict(value=per_rank * cores, unit="spin-flip attempts/s", cores=cores, kind="reference",
                sample=f"the reference's L=64 Ising program ({what}; {cores} MPI ranks = {cores} windows x 1 "
                       f"walker) for {elapsed:.0f} s: {len(rates) - 1} iteration rate(s) of rank 0 after the "
                       f"first, {per_rank:.3e} attempts/s per rank")


def cpu_baseline(systems, params, target_s, threads):
    """The oracle (a single-threaded-per-walker restatement of the reference
    update, FP64 spins) timed on the host cores over a bounded sample of the
    same workload: every walker swept the same number of times."""
    from oracle.oracle import Oracle
    ora = Oracle(systems, params)
    ora.adjust_to_range()
    sec, att = ora.bench_sweeps(1, threads)
    sweeps = max(1, int(target_s / max(sec, 1e-3)))
    sec, att = ora.bench_sweeps(sweeps, threads)
    ora.close()
    return dict(value=att / sec, unit="spin-flip attempts/s", cores=threads, kind="port",
                sample=f"{sweeps} sweep(s) of all {len(range(ora.num_walkers))} walkers from the adjusted "
                       f"initial state, ln f = 1, {att} attempts in {sec:.2f} s")


def cpu_baselines(a, systems, params, gpu_rate):
    """cpu_baseline: the reference program on the host cores at this
    workload's cadence (kind "reference"), or the oracle when the reference
    cannot run there (kind "port"); beside it the reference with its
    defaults, the oracle on the workload, and the oracle calibrated against
    the reference on the same work (moment observables at every attempt, as
    the reference build keeps them)."""
    cpus = host_cpus()
    cores = a.cpu_threads or cpus["used"]
    port = cpu_baseline(systems, params, a.cpu_seconds, cores)
    port["gpu_over_cpu"] = gpu_rate / port["value"]
    ref = None if a.no_reference else reference_cpu_baseline(a.cpu_seconds, cores, "workload")
    if ref is None:
        port["host"] = cpus
        return port
    ref["gpu_over_cpu"] = gpu_rate / ref["value"]
    ref["host"] = cpus
    ref["port"] = port
    same = cpu_baseline(systems, dict(params, observables=1), a.cpu_seconds, cores)
    ref["calibration"] = {"port_with_moments": same["value"], "reference": ref["value"],
                          "port_over_reference": same["value"] / ref["value"],
                          "sample": same["sample"] + " (observables = 1)"}
    dflt = reference_cpu_baseline(a.cpu_seconds, cores, "defaults")
    if dflt is not None:
        dflt["gpu_over_cpu"] = gpu_rate / dflt["value"]
        ref["reference_defaults"] = dflt
    return ref


def lds_peak():
    """The device's measured LDS read bandwidth (GB/s), rewl_selfcheck(2)."""
    import ctypes as C
    from rewl_simulator_amd import abi
    lib = abi.load_engine()
    out = C.c_double()
    abi.check(lib.rewl_selfcheck(2, C.byref(out)))
    return out.value


def reference_cadence(systems, params, sweeps):
    """The workload at the reference's exchange cadence, an exchange after
    every sweep (rewl_parameters.hpp:28): one sweep launch and one exchange
    launch per sweep, the histograms loaded into LDS and written back each
    time.  Reported beside the headline, which exchanges once per step."""
    from rewl_simulator_amd.engine import Engine
    p = dict(params, sweeps_per_exchange=1)
    eng = Engine(systems, p)
    eng.adjust_to_range()
    for _ in range(10):
        eng.sweeps(1)
        eng.exchange()
    eng.synchronize()
    st0 = eng.stats()
    t0 = time.perf_counter()
    for _ in range(sweeps):
        eng.sweeps(1)
        eng.exchange()
    eng.synchronize()
    t1 = time.perf_counter()
    st1 = eng.stats()
    eng.close()
    att = st1["attempts"] - st0["attempts"]
    return {"sweeps_per_exchange": 1, "sweeps": sweeps, "attempts_per_s": att / (t1 - t0),
            "ms_per_sweep_and_exchange": (t1 - t0) * 1e3 / sweeps,
            "exchanges_accepted": st1["exchanges_accepted"] - st0["exchanges_accepted"]}


def saturating_batch(systems, params, copies, peak, steps=2):
    """The same sweep on `copies` independent copies of the workload (own
    seeds) in one context: 2048 walkers fill 63 of 256 CUs, so this shows the
    kernel's throughput on a full chip.  Reported beside the workload's own
    figure, never as it."""
    from rewl_simulator_amd.engine import Engine
    batch = [dict(s, seed=s["seed"] + 7919 * r) for r in range(copies) for s in systems]
    eng = Engine(batch, params)
    eng.adjust_to_range()
    n = params["sweeps_per_check"]
    eng.sweeps(n)
    eng.synchronize()
    st0 = eng.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.sweeps(n)
    eng.synchronize()
    t1 = time.perf_counter()
    st1 = eng.stats()
    eng.close()
    att = st1["attempts"] - st0["attempts"]
    kms = st1["kernel_ms_sweep"] - st0["kernel_ms_sweep"]
    rate = att / (t1 - t0)
    return {"copies": copies, "walkers": 2048 * copies, "sweeps_per_launch": n, "launches": steps,
            "attempts_per_s": rate, "kernel_attempts_per_s": att / (kms / 1e3),
            "lds_GBs": rate * BYTES_PER_ATTEMPT / 1e9, "lds_frac": rate * BYTES_PER_ATTEMPT / 1e9 / peak}


def exact_binned_lng(E_bins, bin_size):
    """Exact ln g of the L=64 periodic Ising model summed over each bin
    [E, E + bin_size), from the committed fixture (exact integer DoS);
    -inf for bins without states."""
    import numpy as np
    f = np.load(os.path.join(ROOT, "tests", "golden", "ising_L64_lng.npz"))
    e, lg = f["E"], f["lng"]
    out = np.full(len(E_bins), -np.inf)
    b = np.floor((e - E_bins[0]) / bin_size).astype(np.int64)
    for k in range(len(E_bins)):
        m = b == k
        if m.any():
            mx = lg[m].max()
            out[k] = mx + np.log(np.exp(lg[m] - mx).sum())
    return out


AT_WORKLOADS = {
    "config3": dict(metric="spin-flip attempts/s (Ashkin-Teller L=32)", walkers=512,
                    data="synthetic (all-up start, Philox draws)", kernel="k_sweep_tuned<AT, L=32>",
                    workload="2D Ashkin-Teller REWL, L=32, J=1, K=0.5, 1024 bins of width 8, 16 windows x 32 walkers"),
    "config4": dict(metric="spin-flip attempts/s (batched random-strain Ashkin-Teller)", walkers=8192,
                    data="synthetic (Gaussian strain per realisation, Philox draws)", kernel="k_sweep_field<AT, 3 producers>",
                    workload="256 random-strain AT realisations, L in {16,24,32,48}, 8 windows x 4 walkers, "
                             "512 bins, one launch per step"),
}


def batched(a):
    """The Ashkin-Teller workloads.  Config 3: one L=32 system, 16 windows x
    32 walkers, row-unrolled integer kernel.  Config 4: 256 random-strain
    realisations (L in {16, 24, 32, 48}, 8 windows x 4 walkers each, 8192
    walkers) swept in one launch per step by the field-model kernel.  A step is
    sweeps_per_step sweeps + exchange + check.  Attempts/s counts sigma and tau
    proposals (2N per sweep)."""
    from rewl_simulator_amd import configs
    from rewl_simulator_amd.engine import Engine
    w = AT_WORKLOADS[a.workload]
    systems, params = configs.config3() if a.workload == "config3" else configs.config4()
    n = a.sweeps_per_step
    params = dict(params, sweeps_per_check=n, sweeps_per_exchange=n, producers=a.producers,
                  sweep_kernel=a.sweep_kernel, walkers_per_group=a.walkers_per_group, observables=a.observables)
    eng = Engine(systems, params)
    eng.adjust_to_range()

    def step():
        eng.sweeps(n)
        eng.exchange()
        eng.check()

    for _ in range(a.warmup):
        step()
    eng.synchronize()
    st0 = eng.stats()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    eng.synchronize()
    t1 = time.perf_counter()
    st1 = eng.stats()
    eng.close()
    att = st1["attempts"] - st0["attempts"]
    rate = att / (t1 - t0)
    launches = st1["sweep_launches"] - st0["sweep_launches"]
    avg_launch_s = (st1["kernel_ms_sweep"] - st0["kernel_ms_sweep"]) / 1e3 / max(launches, 1)
    achieved = att / max(launches, 1) * 32.5 / avg_launch_s / 1e9
    peak = lds_peak()
    line = {
        "metric": w["metric"], "value": rate,
        "unit": "spin-flip attempts/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": (t1 - t0) * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 energies and ln g" if a.workload == "config4" else "int32 energies, f64 ln g", "data": w["data"],
        "config": {"workload": w["workload"], "walkers": w["walkers"], "sweeps_per_step": n,
                   "observables": a.observables},
        "roofline": {"bound": "lds", "achieved": achieved, "peak": peak, "unit": "GB/s",
                     "frac": achieved / peak, "traffic": None, "kernel": w["kernel"],
                     "peak_source": "rewl_selfcheck(2)", "avg_launch_ms": avg_launch_s * 1e3,
                     "bytes_per_attempt": 32.5,
                     "hbm": {"peak": HBM_PEAK_GBS, "algorithmic_frac": achieved / HBM_PEAK_GBS}},
    }
    if not a.no_cpu:
        cpus = host_cpus()
        threads = a.cpu_threads or cpus["used"]
        line["cpu_baseline"] = cpu_baseline(systems, params, a.cpu_seconds, threads)
        line["cpu_baseline"]["gpu_over_cpu"] = rate / line["cpu_baseline"]["value"]
        line["cpu_baseline"]["host"] = cpus
    print(json.dumps(line), flush=True)


def converge_batch(a):
    """Time to ln f < a.converge for every realisation of the batched
    random-strain workload (config 4) in one context: checks of
    sweeps_per_step sweeps, an exchange and a check; with retirement (the
    default) a realisation whose windows are all done leaves the launches.
    Reports when the last realisation of each lattice size finished, the
    sweep-launch time per check at the start and at the end (what retirement
    saves), and how many realisations finished within --max-checks."""
    from rewl_simulator_amd import configs
    from rewl_simulator_amd.engine import Engine
    systems, params = configs.config4(lnf_final=a.converge)
    n = a.sweeps_per_step
    params = dict(params, sweeps_per_check=n, sweeps_per_exchange=n, producers=a.producers,
                  sweep_kernel=a.sweep_kernel, walkers_per_group=a.walkers_per_group,
                  retire=0 if a.no_retire else 1)
    eng = Engine(systems, params)
    eng.adjust_to_range()
    wins = {}
    for k in range(eng.num_windows):
        wins.setdefault(eng.windows[k]["system"], []).append(k)
    sizes = [s["L"] for s in systems]
    finished = {}
    launch_ms = []
    eng.synchronize()
    st0 = eng.stats()
    t0 = time.perf_counter()
    checks = 0
    while checks < a.max_checks and len(finished) < len(systems):
        ms0 = eng.stats()["kernel_ms_sweep"]
        eng.sweeps(n)
        eng.exchange()
        eng.check()
        checks += 1
        launch_ms.append(eng.stats()["kernel_ms_sweep"] - ms0)
        if checks % 5 == 0:
            now = time.perf_counter() - t0
            for si, ks in wins.items():
                if si not in finished and all(eng.window_info(k)["done"] for k in ks):
                    finished[si] = now
            if checks % 20 == 0:
                print(f"# check {checks}: {len(finished)} of {len(systems)} realisations done, {now:.1f} s",
                      file=sys.stderr, flush=True)
    eng.synchronize()
    t1 = time.perf_counter()
    st1 = eng.stats()
    # per lattice size and window index: windows at the final ln f, and
    # windows past their first stage (ln f halved at least once)
    per_window = {}
    infos = [eng.window_info(k) for k in range(eng.num_windows)]
    for L in sorted(set(sizes)):
        rows = []
        for w in range(max(i["window"] for i in infos) + 1):
            sel = [i for i in infos if sizes[i["system"]] == L and i["window"] == w]
            rows.append({"windows": len(sel), "done": sum(int(i["done"]) for i in sel),
                         "past_first_stage": sum(int(i["lnf"] < 1.0) for i in sel)})
        per_window[str(L)] = rows
    eng.close()
    by_L = {}
    for L in sorted(set(sizes)):
        idx = [i for i, x in enumerate(sizes) if x == L]
        done = [finished[i] for i in idx if i in finished]
        by_L[str(L)] = {"realisations": len(idx), "finished": len(done),
                        "last_finished_s": max(done) if len(done) == len(idx) else None}
    head = launch_ms[:10]
    tail = launch_ms[-10:]
    print(json.dumps({
        "metric": "time to ln f < %g, every realisation of the batched random-strain workload" % a.converge,
        "value": t1 - t0 if len(finished) == len(systems) else None, "unit": "s", "higher_is_better": False,
        "converged": len(finished) == len(systems), "checks": checks, "sweeps_per_check": n,
        "retire": not a.no_retire, "sweeps": st1["sweeps"] - st0["sweeps"],
        "attempts": st1["attempts"] - st0["attempts"], "wall_s": t1 - t0,
        "by_L": by_L, "per_window_index": per_window,
        "sweep_launch_ms_per_check": {"first_10_mean": sum(head) / max(len(head), 1),
                                      "last_10_mean": sum(tail) / max(len(tail), 1)},
        "config": {"workload": AT_WORKLOADS["config4"]["workload"]},
    }), flush=True)


def converge_at(a):
    """Time to ln f < a.converge on the Ashkin-Teller workload (config 3):
    wall time from the first sweep until the slowest window's ln f is below
    the threshold.  No exact density of states is known for it; the stitched,
    normalised ln g is checked at the ground bin, which holds only the four
    ground states (the lowest excitation costs 8J + 4K = 12 > the bin width
    8), so ln g there must be ln 4 (with the full range, where the
    normalisation is exact), and at the step to bin 1, which holds only the
    8N single-flip excitations: ln g(1) - ln g(0) = ln 2N for either range."""
    import numpy as np
    from rewl_simulator_amd import configs
    from rewl_simulator_amd.engine import Engine
    systems, params = configs.config3(lnf_final=a.converge)
    if a.half_range:
        N = 32 * 32
        systems = [dict(systems[0], e_max=0.0)]
        assert systems[0]["e_min"] == -5.0 * N
    params = dict(params, average_mode=a.average_mode, observables=a.observables)
    eng = Engine(systems, params)
    eng.adjust_to_range()
    per_check = params["sweeps_per_check"] // params["sweeps_per_exchange"]
    eng.synchronize()
    st0 = eng.stats()
    t0 = time.perf_counter()
    checks = 0
    lnf = 1.0
    while checks < a.max_checks:
        for _ in range(per_check):
            eng.sweeps(params["sweeps_per_exchange"])
            eng.exchange()
        eng.check()
        checks += 1
        if checks % 10 == 0 or eng.stats()["all_done"]:
            lnf = max(eng.window_info(k)["lnf"] for k in range(eng.num_windows))
            if checks % 100 == 0:
                print(f"# check {checks}: max ln f {lnf:.3e}, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            if lnf < a.converge:
                break
    eng.synchronize()
    t1 = time.perf_counter()
    st1 = eng.stats()
    eng.finalize()
    E, lng = eng.stitched(0, states_per_site=4.0)
    eng.close()
    print(json.dumps({
        "metric": "time to ln f < %g (Ashkin-Teller L=32)" % a.converge, "value": t1 - t0, "unit": "s",
        "higher_is_better": False, "converged": bool(lnf < a.converge), "checks": checks,
        "sweeps": st1["sweeps"] - st0["sweeps"], "attempts": st1["attempts"] - st0["attempts"],
        "attempts_per_s": (st1["attempts"] - st0["attempts"]) / (t1 - t0),
        "ground_bin_lng": float(lng[0]), "ground_bin_error": float(abs(lng[0] - np.log(4.0))),
        # bin 1 holds only the single-flip excitations E0 + 8J + 4K: 4 ground states x 2N flips
        "lowest_gap_error": float(abs((lng[1] - lng[0]) - np.log(2.0 * 32 * 32))),
        "config": {"workload": AT_WORKLOADS["config3"]["workload"] + (", range [-5N, 0)" if a.half_range else ""),
                   "schedule": "halving",
                   "average_mode": params["average_mode"], "observables": a.observables},
    }), flush=True)


def widen(eng, systems, params, copies):
    """A context like eng's (one system) with `copies` times the replicas per
    window: every walker starts from eng's finalised ln g of its window and
    the configuration of one of eng's walkers there.  More walkers sample the
    same fixed weights in the same time while the chip has room for them."""
    from rewl_simulator_amd.engine import Engine
    s = dict(systems[0])
    r0 = s["replicas_per_window"]
    s["replicas_per_window"] = r0 * copies
    big = Engine([s], params)
    for k, w in enumerate(big.windows):
        w0 = eng.windows[k]
        lng = eng.logdos(w0["first_walker"])
        for r in range(s["replicas_per_window"]):
            g = w["first_walker"] + r
            big.set_logdos(g, lng)
            big.set_spins(g, eng.spins(w0["first_walker"] + r % r0))
    big.adjust_to_range()
    return big


def converge(eng, params, a, systems=None, print_line=True):
    """Time to ln f < a.converge: wall time from the first sweep until the
    slowest window's ln f drops below the threshold (the reference's
    f < final_increment test, rewl_simulation.hpp:594, 636-671); then the
    stitched, normalised ln g against the exact DoS."""
    per_check = params["sweeps_per_check"] // params["sweeps_per_exchange"]
    eng.synchronize()
    st0 = eng.stats()
    t0 = time.perf_counter()
    checks = 0
    lnf = 1.0
    while checks < a.max_checks:
        for _ in range(per_check):
            eng.sweeps(params["sweeps_per_exchange"])
            eng.exchange()
        eng.check()
        checks += 1
        if checks % 10 == 0 or eng.stats()["all_done"]:
            lnf = max(eng.window_info(k)["lnf"] for k in range(eng.num_windows))
            if checks % 200 == 0:
                print(f"# check {checks}: max ln f {lnf:.3e}, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            if lnf < a.converge:
                break
    eng.synchronize()
    t1 = time.perf_counter()
    st1 = eng.stats()
    import numpy as np
    eng.finalize()

    def errors(e):
        E, lng = e.stitched(0, states_per_site=2.0)
        ex = exact_binned_lng(E, params_bin_size(e))
        ok = np.isfinite(ex)
        d = (lng - ex)[ok]
        E2, lng2 = e.stitched(0, states_per_site=2.0, join="overlap")
        ok2 = np.isfinite(ex[:len(lng2)]) & np.isfinite(lng2)
        d2 = (lng2 - ex[:len(lng2)])[ok2]
        return E, lng, ex, ok, d, d2

    E, lng, ex, ok, d, d2 = errors(eng)
    refined = None
    if a.refine > 0:
        t2 = time.perf_counter()
        target = widen(eng, systems, params, a.refine_copies) if a.refine_copies > 1 else eng
        t2b = time.perf_counter()

        def progress(done):
            print(f"# refinement: {done} of {a.refine} sweeps, {time.perf_counter() - t2b:.1f} s", file=sys.stderr, flush=True)

        target.refine(a.refine, params["sweeps_per_exchange"], chunk=25000, progress=progress)
        target.synchronize()
        t3 = time.perf_counter()
        _, _, _, _, r1, r2 = errors(target)
        refined = {"sweeps": a.refine, "walkers": target.num_walkers, "seconds": t3 - t2,
                   "seconds_to_seed_the_walkers": t2b - t2,
                   "lng_max_abs_error_normalised": float(np.max(np.abs(r1))),
                   "lng_max_abs_error_overlap_join": float(np.max(np.abs(r2))),
                   "total_seconds": (t1 - t0) + (t3 - t2)}
        if target is not eng:
            target.close()
    if a.dump:
        nbmax = max(w["num_bins"] for w in eng.windows)
        wl = np.full((eng.num_windows, nbmax), np.nan)
        for k, w in enumerate(eng.windows):
            wl[k, :w["num_bins"]] = eng.logdos(w["first_walker"])
        np.savez(a.dump, E=E, lng=lng, exact=ex, window_emin=np.array([w["e_min"] for w in eng.windows]),
                 window_lng=wl)
    result = {
        "metric": "time to ln f < %g" % a.converge, "value": t1 - t0, "unit": "s", "higher_is_better": False,
        "converged": bool(lnf < a.converge), "checks": checks, "sweeps": st1["sweeps"] - st0["sweeps"],
        "attempts": st1["attempts"] - st0["attempts"],
        "attempts_per_s": (st1["attempts"] - st0["attempts"]) / (t1 - t0),
        "flat_events": st1["flat_events"], "max_lnf": lnf,
        "lng_max_abs_error_normalised": float(np.max(np.abs(d))),
        "lng_max_abs_error_anchored": float(np.max(np.abs(d - d[0]))),
        "lng_bins_compared": int(ok.sum()),
        "lng_max_abs_error_overlap_join": float(np.max(np.abs(d2))),
        "refined": refined,
        "config": {"workload": "2D Ising REWL, L=64 periodic, 32 windows x 64 walkers, flatness 0.8",
                   "schedule": "1/t" if a.one_over_t else "halving", "average_mode": params["average_mode"]},
    }
    if print_line:
        print(json.dumps(result), flush=True)
    return result


def time_to_lnf_leg(systems, params, a):
    """The metric's second half for the line: a fresh context of the same
    workload run from ln f = 1 until every window's ln f is below
    a.time_to_lnf (halving, equal-weight averaging, an exchange and a check
    every 1000 sweeps), with the stitched ln g against the exact DoS.  The
    learned warm start is not part of this workload (plain flat init)."""
    from rewl_simulator_amd.engine import Engine
    p = dict(params, sweeps_per_check=1000, sweeps_per_exchange=1000, average_mode=1)
    leg = argparse.Namespace(converge=a.time_to_lnf, max_checks=2000, refine=0, refine_copies=1, dump=None,
                             one_over_t=False)
    with Engine(systems, p) as eng:
        eng.adjust_to_range()
        r = converge(eng, p, leg, systems, print_line=False)
    return {k: r[k] for k in ("metric", "value", "unit", "converged", "checks", "sweeps", "attempts",
                              "lng_max_abs_error_normalised", "lng_max_abs_error_overlap_join")}


def warm_start_leg(a):
    """The metric's warm-start comparison, sized to run in about a minute:
    the learned warm start end to end.  Training curves: 24 random-strain AT
    realisations at L = 8 and 12 (sigma 0, 0.5, 1) run to ln f <
    a.warm_start_lnf; the network trained on them by the library's kernels;
    8 held-out L = 12 realisations (sigma 0.5) run from a flat ln g at
    ln f = 1 (plain) and from the network's ln g at 1e-3 (warm) to the same
    final ln f.  Sweep counts are deterministic (Philox streams, fixed
    seeds); wall times are the engine's."""
    from rewl_simulator_amd import warmstart as W

    def log(msg):
        print("# " + msg.strip(), file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    curves = W.generate(W.training_specs(24, sizes=(8, 12), sigmas=(0.0, 0.5, 1.0)), lnf_final=a.warm_start_lnf,
                        batch=24, max_checks=4000, log=log)
    mlp, _ = W.train(curves, steps=6000, batch=8192, log=log)
    res = W.evaluate(mlp, W.heldout_specs(8, L=12, sigma=0.5), lnf_final=a.warm_start_lnf, ref_lnf=0.0,
                     max_checks=4000, log=log)
    plain, warm = res["plain"], res["warm"]
    return {
        "metric": "sweeps and wall time to ln f < %g, learned warm start vs plain" % a.warm_start_lnf,
        "workload": "8 held-out random-strain AT realisations, L=12, sigma=0.5, 8 windows x 4 walkers, 512 bins "
                    "over [E(all-up), 0); network trained on 24 engine-produced curves at L=8, 12",
        "plain": {"sweeps": plain["sweeps"], "seconds": plain["seconds"], "converged": plain["converged"]},
        "warm": {"sweeps": warm["sweeps"], "seconds": warm["seconds"], "init_seconds": warm["init_seconds"],
                 "converged": warm["converged"], "lnf_start": warm["lnf_start"]},
        "factor_sweeps": res["speedup_sweeps"], "factor_wall": res["speedup_wall"],
        "warm_vs_plain_max_abs_lng": res["warm_vs_plain_max"],
        "network_error_max": res["prediction_error_max"],
        "leg_seconds": time.perf_counter() - t0,
    }


def config2_warm_leg(systems, params, a):
    """The metric's warm start on this workload.  Training curves: the clean
    2D Ising model at L = 16, 24, 32, 48 from the engine (bins of width 8
    over the full spectrum, L/2 windows x 16 walkers, to ln f <
    a.time_to_lnf); the network trained on them by the library's kernels
    (inputs E/N and 1/L; K and sigma are 0).  Then this workload starts from
    the network's ln g at ln f = 1e-3 and is timed to ln f < a.time_to_lnf
    exactly as the time_to_lnf leg times the plain start (halving,
    equal-weight averaging, an exchange and a check every 1000 sweeps), with
    the same ln g error against the exact DoS."""
    import numpy as np
    from rewl_simulator_amd import warmstart as W
    from rewl_simulator_amd.engine import Engine
    from rewl_simulator_amd.mlp import features, warm_start

    def log(msg):
        print("# " + msg.strip(), file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    sizes = (16, 24, 32, 48)
    exact = a.warm_start_source == "exact"
    if not exact:
        curves, gen_done, gen_s = W.ising_curves(sizes, lnf_final=a.time_to_lnf, log=log)
        mlp, _ = W.train(curves, steps=6000, batch=8192, log=log)
    L = systems[0]["L"]
    N = L * L
    E = systems[0]["e_min"] + systems[0]["bin_size"] * np.arange(
        int(round((systems[0]["e_max"] - systems[0]["e_min"]) / systems[0]["bin_size"])))
    ex = exact_binned_lng(E, systems[0]["bin_size"])
    ok = np.isfinite(ex)
    pred = np.where(ok, ex, 0.0) if exact else N * mlp.predict(features(E, N, L, 0.0, 0.0))
    dp = (pred - ex)[ok]
    p = dict(params, sweeps_per_check=1000, sweeps_per_exchange=1000, average_mode=1)
    leg = argparse.Namespace(converge=a.time_to_lnf, max_checks=2000, refine=0, refine_copies=1, dump=None,
                             one_over_t=False)
    with Engine(systems, p) as eng:
        eng.adjust_to_range()
        t1 = time.perf_counter()
        if exact:
            # every walker's ln g = the exact binned ln g over its window's bins
            e0, bs = systems[0]["e_min"], systems[0]["bin_size"]
            for k, w in enumerate(eng.windows):
                g0 = int(round((w["e_min"] - e0) / bs))
                lng = pred[g0:g0 + w["num_bins"]].copy()
                for r in range(systems[0]["replicas_per_window"]):
                    eng.set_logdos(w["first_walker"] + r, lng)
                eng.set_window_lnf(k, 1e-3)
        else:
            warm_start(eng, mlp, lambda i: 0.0, 1e-3)
        init_s = time.perf_counter() - t1
        r = converge(eng, p, leg, systems, print_line=False)
    out = {k: r[k] for k in ("metric", "value", "unit", "converged", "checks", "sweeps", "attempts",
                             "lng_max_abs_error_normalised", "lng_max_abs_error_overlap_join")}
    out.update(lnf_start=1e-3, init_seconds=init_s, source=a.warm_start_source,
               start_error_before_sampling={"max_abs_anchored": float(np.max(np.abs(dp - dp[0]))),
                                            "max_abs_mean_removed": float(np.max(np.abs(dp - dp.mean())))},
               leg_seconds=time.perf_counter() - t0)
    if not exact:
        out["training"] = {"sizes": list(sizes), "curves_converged": bool(gen_done), "generate_seconds": gen_s,
                           "points": int(sum(int(c["visited"].sum()) for c in curves))}
    return out


# The walker's instruction stream of the default launch: 24 instructions
# between successive attempts of the unrolled L = 64 row (the median over the
# row in the ISA of k_sweep_tuned<0, 64, true, 1, 0>), about 16 of them VALU.
# A wavefront issues at most one VALU instruction every 4 cycles, so the
# attempt cannot take fewer than about 4 cycles per instruction.
WALKER_INSTRUCTIONS_PER_ATTEMPT = 24


def latency_model(cycles, clock_ghz, walkers, walkers_per_wave, default_launch):
    """Cycles per attempt per walker (the walker's chain is the bound of this
    workload), and for the default launch the issue floor of its instruction
    stream and the fraction of it reached."""
    out = {"cycles_per_attempt_per_walker": cycles, "clock_ghz_nominal": clock_ghz, "cus": 256}
    if default_launch:
        floor = 4.0 * WALKER_INSTRUCTIONS_PER_ATTEMPT
        out.update({"walker_wavefronts": -(-walkers // 33), "walkers_per_wavefront": walkers_per_wave,
                    "issue_floor": {"instructions_per_attempt": WALKER_INSTRUCTIONS_PER_ATTEMPT,
                                    "cycles_per_instruction": 4, "floor_cycles": floor,
                                    "frac": floor / cycles}})
    return out


def params_bin_size(eng):
    return eng.systems[0]["bin_size"]


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported: this build runs the "
              f"engine on one GPU (sharding windows across GPUs is a design, not built).", file=sys.stderr)
        sys.exit(2)
    if a.workload in AT_WORKLOADS:
        if a.converge > 0 and a.workload == "config3":
            converge_at(a)
        elif a.converge > 0:
            converge_batch(a)
        else:
            batched(a)
        return
    from rewl_simulator_amd import configs
    from rewl_simulator_amd.engine import Engine

    systems, params = configs.config2()
    params = dict(params, sweeps_per_check=a.sweeps_per_step, sweeps_per_exchange=a.sweeps_per_step,
                  walkers_per_group=a.walkers_per_group, producers=a.producers, sweep_kernel=a.sweep_kernel,
                  average_mode=a.average_mode, observables=a.observables)
    if a.one_over_t:
        from rewl_simulator_amd import abi
        params["schedule"] = abi.SCHED_ONE_OVER_T
    eng = Engine(systems, params)
    eng.adjust_to_range()
    W = eng.num_walkers
    N = systems[0]["L"] ** 2
    per_check = params["sweeps_per_check"] // params["sweeps_per_exchange"]

    def step():
        for _ in range(per_check):
            eng.sweeps(params["sweeps_per_exchange"])
            eng.exchange()
        eng.check()

    if a.converge > 0:
        converge(eng, params, a, systems)
        return
    for _ in range(a.warmup):
        step()
    eng.synchronize()
    st0 = eng.stats()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    eng.synchronize()
    t1 = time.perf_counter()
    st1 = eng.stats()
    elapsed = t1 - t0
    attempts = st1["attempts"] - st0["attempts"]
    rate = attempts / elapsed
    launches = st1["sweep_launches"] - st0["sweep_launches"]
    kms = st1["kernel_ms_sweep"] - st0["kernel_ms_sweep"]
    avg_launch_s = kms / 1e3 / max(launches, 1)
    att_per_launch = attempts / max(launches, 1)
    achieved = att_per_launch * BYTES_PER_ATTEMPT / avg_launch_s / 1e9
    lnf = min(w["lnf"] for w in (eng.window_info(k) for k in range(eng.num_windows)))
    default_launch = a.sweep_kernel in (0, 3) and a.producers == 0 and a.walkers_per_group == 0 and not a.observables
    traffic = MEASURED_TRAFFIC.get(params["sweeps_per_check"]) if default_launch else None

    peak = lds_peak()
    clock_ghz = 2.4
    walkers_per_wave = W / max(1, -(-W // 33))
    cycles = avg_launch_s * clock_ghz * 1e9 / (att_per_launch / W)
    line = {
        "metric": "spin-flip attempts/s; time to ln f=1e-6 (learned warm start vs plain)",
        "value": rate,
        "unit": "spin-flip attempts/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 energies, f64 ln g",
        "data": "synthetic (all-up start, Philox4x32-10 draws)",
        "config": {
            "workload": "2D Ising REWL, L=64 periodic, 2049 bins of width 8, 32 windows (75% overlap) x 64 "
                        "walkers, flatness 0.8, ln f halving to 1e-8",
            "walkers": W,
            "sweeps_per_step": params["sweeps_per_check"],
            "sweeps_per_exchange": params["sweeps_per_exchange"],
            "attempts_per_step": attempts // a.steps,
            "parallelism": "one lane per walker, 1 GPU",
            "producers": a.producers,
            "sweep_kernel": a.sweep_kernel,
            "observables": a.observables,
        },
        "roofline": {
            # the histograms (all 32 of the 32.25 algorithmic bytes per
            # attempt) live in LDS: the bound is the LDS bandwidth measured on
            # this device; HBM carries only the lattice stream
            "bound": "lds",
            "achieved": achieved,
            "peak": peak,
            "unit": "GB/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "peak_source": "rewl_selfcheck(2): 16-byte conflict-free LDS reads on every CU of this GPU",
            "kernel": "k_sweep_tuned<Ising, L=64> (row-unrolled sweep)" if a.sweep_kernel in (0, 2, 3, 4)
                      else "k_sweep_fast<Ising> (row kernel)",
            "avg_launch_ms": avg_launch_s * 1e3,
            "attempts_per_launch": att_per_launch,
            "bytes_per_attempt": BYTES_PER_ATTEMPT,
            "kernel_attempts_per_s": att_per_launch / avg_launch_s,
            "hbm": {"peak": HBM_PEAK_GBS, "measured_GBs": traffic / avg_launch_s / 1e9 if traffic else None,
                    "frac": traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS if traffic else None},
            # one walker is one sequential chain: what bounds this workload is
            # the walker wavefront's time per attempt, not a bandwidth
            "latency": latency_model(cycles, clock_ghz, W, walkers_per_wave, default_launch),
        },
        "state": {"min_lnf_after": lnf, "flat_events": st1["flat_events"],
                  "exchanges_accepted": st1["exchanges_accepted"]},
    }
    eng.close()
    if a.saturate > 0:
        line["saturating_batch"] = saturating_batch(systems, params, a.saturate, peak)
    if a.cadence_sweeps > 0:
        line["reference_cadence"] = reference_cadence(systems, params, a.cadence_sweeps)
    if a.time_to_lnf > 0:
        line["time_to_lnf"] = time_to_lnf_leg(systems, params, a)
    if a.warm_start_config2 and a.time_to_lnf > 0:
        w = config2_warm_leg(systems, params, a)
        plain = line.get("time_to_lnf")
        if plain and plain["converged"] and w["converged"]:
            w["factor_wall"] = plain["value"] / (w["value"] + w["init_seconds"])
            w["factor_sweeps"] = plain["sweeps"] / max(w["sweeps"], 1)
        line["warm_start_config2"] = w
    if a.warm_start_lnf > 0:
        line["warm_start"] = warm_start_leg(a)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baselines(a, systems, params, rate)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
