"""Writes tests/golden/ref/: result files of the reference REWL_Simulator itself.

Each case builds the reference from its sources with oracle/ref_build.sh
(MPICH under /opt/conda, DIFFERENT_SEEDS off so every walker is seeded with 1),
runs it with `mpiexec -n <ranks>` in oracle/_ref/fixtures/runs/<case>/work and keeps
what the run wrote:

  <case>.microcanonical.txt   the reference's microcanonical_observables file,
                              verbatim (E/N, ln g/N, per-bin observables)
  <case>.self_averaged.txt    its self_averaged_observables file, header and
                              every 1000th temperature row (50 of 50 000)
  <case>.nonlinear.txt        its nonlinear_observables file, the same rows
  cases.json                  the parameters of every case, the run's wall
                              time and the names the reference gave its files

These are outputs of the reference program (data), not its sources.  The
reference is only present in the build container, so the fixtures are made
there and committed; tests/test_reference_fixtures.py checks the oracle
against them on any machine.

    python tests/golden/make_ref_fixtures.py [case ...]
"""
import glob
import json
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "ref")
REF_OUT = os.path.join(ROOT, "oracle", "_ref", "fixtures")
MPIEXEC = "/opt/conda/bin/mpiexec"

# name -> build settings (oracle/ref_build.sh) and MPI ranks.  Ranks =
# windows x replicas_per_window (rewl_simulation.hpp:72-80).
CASES = {
    # the reference's defaults throughout (1/t, all-visited flatness, exchange
    # every sweep, final ln f 1e-7, one walker per window, CORRELATION_LENGTHS)
    "ising_L8_w4r1": dict(model="ISING2D", L=8, ranks=4, replicas=1, final="1e-7"),
    # two replicas per window: partner shuffles, count-weighted averaging
    "ising_L8_w4r2": dict(model="ISING2D", L=8, ranks=8, replicas=2, final="1e-5"),
    # four replicas per window: the MPI_Allreduce order of four doubles
    "ising_L6_w2r4": dict(model="ISING2D", L=6, ranks=8, replicas=4, final="1e-4"),
    # halving with the classic (max - min)/min flatness
    "ising_L6_w2r2_halving": dict(model="ISING2D", L=6, ranks=4, replicas=2, final="1e-3", one_over_t="OFF",
                                  flatness="0.3"),
    "at_L6_w4r2": dict(model="ASHKIN_TELLER2D", L=6, ranks=8, replicas=2, final="1e-5"),
    # random fields with generate_random_field's clock seed fixed
    "rfim_L6_w2r2": dict(model="ISING2D", L=6, ranks=4, replicas=2, final="1e-4", disorder="RFIM",
                         field_h="0.6", field_seed="12345"),
    "rfat_L6_w2r2": dict(model="ASHKIN_TELLER2D", L=6, ranks=4, replicas=2, final="1e-4",
                         disorder="RFAT_BAXTER", field_h="0.8", field_seed="777"),
}


def build(name, c):
    env = dict(os.environ, REF_OUT_DIR=REF_OUT, REF_NAME=name, REF_MODEL=c["model"], REF_L=str(c["L"]), REF_REPLICAS=str(c["replicas"]),
               REF_FINAL=c["final"], REF_ONE_OVER_T=c.get("one_over_t", "ON"), REF_FLATNESS=c.get("flatness", "0.3"),
               REF_DISORDER=c.get("disorder", "none"))
    if "field_h" in c:
        env["REF_FIELD_H"] = c["field_h"]
        env["REF_FIELD_SEED"] = c["field_seed"]
    subprocess.run(["bash", os.path.join(ROOT, "oracle", "ref_build.sh")], env=env, check=True)
    return os.path.join(REF_OUT, "rewl_ref_" + name)


def subsample(src, dst, every=1000):
    with open(src) as f, open(dst, "w") as g:
        k = 0
        for line in f:
            if line.startswith("#"):
                g.write(line)
                continue
            if k % every == 0:
                g.write(line)
            k += 1


def run(name, c, exe):
    rundir = os.path.join(REF_OUT, "runs", name)
    shutil.rmtree(rundir, ignore_errors=True)
    os.makedirs(os.path.join(rundir, "work"))
    t = time.time()
    with open(os.path.join(rundir, "stdout.txt"), "w") as log:
        subprocess.run([MPIEXEC, "-n", str(c["ranks"]), exe], cwd=os.path.join(rundir, "work"), stdout=log,
                       stderr=subprocess.STDOUT, check=True, timeout=3600)
    wall = time.time() - t
    files = glob.glob(os.path.join(rundir, "REWL_Simulation_Data", "*", "*", "*.txt"))
    by = {os.path.basename(f).split("-")[0]: f for f in files}
    os.makedirs(OUT, exist_ok=True)
    shutil.copyfile(by["microcanonical_observables"], os.path.join(OUT, name + ".microcanonical.txt"))
    subsample(by["self_averaged_observables"], os.path.join(OUT, name + ".self_averaged.txt"))
    subsample(by["nonlinear_observables"], os.path.join(OUT, name + ".nonlinear.txt"))
    model_dir = os.path.basename(os.path.dirname(os.path.dirname(by["microcanonical_observables"])))
    return wall, dict(model_dir=model_dir, **{k: os.path.basename(v) for k, v in by.items()})


def main(names):
    manifest_path = os.path.join(OUT, "cases.json")
    manifest = json.load(open(manifest_path)) if os.path.exists(manifest_path) else {}
    for name in names or list(CASES):
        c = CASES[name]
        exe = build(name, c)
        wall, files = run(name, c, exe)
        manifest[name] = dict(c, reference_wall_s=round(wall, 2), files=files)
        print(f"{name}: {wall:.1f} s", flush=True)
        with open(manifest_path, "w") as f:
            json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
