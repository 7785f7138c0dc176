"""Build the HIP engine (gfx950) and the CPU oracle in-tree.

    python -m rewl_simulator_amd.build            # all
    python -m rewl_simulator_amd.build engine     # HIP engine only
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
ENGINE_SO = os.path.join(PKG, "librewl_hip.so")
HOST_DIR = os.path.join(PKG, "host")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the row-unrolled sweep kernel unrolls L attempts (with their rare exact
    # path) per row; the default cap on pragma-requested unrolling is too low
    "-mllvm", "-pragma-unroll-threshold=200000",
    "-Wall",
    "-Wno-unused-result",
]


def build_engine(verbose=False):
    src = os.path.join(CSRC, "engine.hip")
    tmp = ENGINE_SO + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, ENGINE_SO)
    return ENGINE_SO


def build_host(verbose=False):
    """The C++ host program over rewl_host.hpp, for both model plug-ins."""
    outs = []
    for name, defs in (("rewl_run", []), ("rewl_run_at", ["-DREWL_MODEL_AT"])):
        out = os.path.join(HOST_DIR, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", *defs, "-I", os.path.join(ROOT, "include"),
               "-I", HOST_DIR, os.path.join(HOST_DIR, "rewl_run.cpp"), "-o", out, "-L", PKG, "-lrewl_hip",
               "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        outs.append(out)
    return outs


def build_oracle(verbose=False):
    cmd = ["make", "-C", ORACLE_DIR, "-B", "liboracle.so"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=None if verbose else subprocess.DEVNULL)
    return ORACLE_SO


REFERENCE_SRC = "/root/reference/src"
REF_PROGRAM = os.path.join(ORACLE_DIR, "_ref", "rewl_ref_ising_L64")


# The reference program builds timed as CPU baselines: its defaults, and the
# L = 64 Ising workload's own cadence (an exchange every 1000 sweeps) without
# the correlator (CORRELATION_LENGTHS off; the reference still updates its
# magnetisation moments at every attempt).
REFERENCE_BUILDS = {
    "ising_L64": dict(REF_MODEL="ISING2D", REF_L="64"),
    "ising_L64_plain": dict(REF_MODEL="ISING2D", REF_L="64", REF_CORR="OFF", REF_SPE="1000"),
}


def build_reference(verbose=False):
    """The reference program itself for the L = 64 Ising workload (every
    walker seeded with 1), the CPU baseline of kind "reference": compiled from
    its sources by oracle/ref_build.sh where they exist (the build container);
    elsewhere the binaries built there are used."""
    if not os.path.isdir(REFERENCE_SRC):
        return None
    cmd = ["bash", os.path.join(ORACLE_DIR, "ref_build.sh")]
    for name, opts in REFERENCE_BUILDS.items():
        env = dict(os.environ, REF_NAME=name, **opts)
        if verbose:
            print(" ".join(f"{k}={v}" for k, v in dict(REF_NAME=name, **opts).items()) + " " + " ".join(cmd),
                  flush=True)
        subprocess.run(cmd, env=env, check=True, stdout=None if verbose else subprocess.DEVNULL)
    return REF_PROGRAM


def main(argv):
    what = argv[1:] or ["engine", "host", "oracle", "reference"]
    if "engine" in what:
        print(build_engine(verbose=True))
    if "host" in what:
        print(build_host(verbose=True))
    if "oracle" in what:
        print(build_oracle(verbose=True))
    if "reference" in what:
        print(build_reference(verbose=True))


if __name__ == "__main__":
    main(sys.argv)
