"""The C-ABI library: it loads, exports every function the headers under
include/ declare, its ctypes mirror has the C struct layouts, and it refuses
to run without a GPU instead of falling back to the CPU.  No compute calls
here."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from rewl_simulator_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rewl_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = abi.load_engine()
    names = declared_functions()
    assert len(names) >= 36
    assert "rewl_mlp_train" in names
    for n in names:
        assert hasattr(lib, n), n
        assert n in abi.SIGNATURES, f"{n} missing from the ctypes mirror"


def test_struct_layouts_match_c():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "rewl_c.h"
#include "rewl_mlp.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("rewl_system_desc %zu\n", sizeof(rewl_system_desc));
  printf("rewl_params %zu\n", sizeof(rewl_params));
  printf("rewl_window_info %zu\n", sizeof(rewl_window_info));
  printf("rewl_stats %zu\n", sizeof(rewl_stats));
  printf("rewl_mlp_train_params %zu\n", sizeof(rewl_mlp_train_params));
  printf("REWL_MLP_NUM_PARAMS %d\n", REWL_MLP_NUM_PARAMS);
  F(rewl_system_desc, field) F(rewl_system_desc, seed) F(rewl_system_desc, walker_seeds)
  F(rewl_params, lnf_initial) F(rewl_params, exchange) F(rewl_params, producers) F(rewl_params, adjust_mode) F(rewl_params, average_mode) F(rewl_params, sweep_kernel) F(rewl_params, retire)
  F(rewl_window_info, lnf) F(rewl_window_info, one_over_t)
  F(rewl_stats, kernel_ms_sweep) F(rewl_stats, sweep_launches)
  F(rewl_mlp_train_params, lr) F(rewl_mlp_train_params, seed) F(rewl_mlp_train_params, warmup)
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        cpath = os.path.join(d, "probe.c")
        open(cpath, "w").write(src)
        exe = os.path.join(d, "probe")
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), cpath, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    assert int(got["rewl_system_desc"]) == C.sizeof(abi.SystemDesc)
    assert int(got["rewl_params"]) == C.sizeof(abi.Params)
    assert int(got["rewl_window_info"]) == C.sizeof(abi.WindowInfo)
    assert int(got["rewl_stats"]) == C.sizeof(abi.Stats)
    assert int(got["rewl_mlp_train_params"]) == C.sizeof(abi.MlpTrainParams)
    assert int(got["REWL_MLP_NUM_PARAMS"]) == abi.MLP_NUM_PARAMS
    pairs = {"rewl_system_desc": abi.SystemDesc, "rewl_params": abi.Params,
             "rewl_window_info": abi.WindowInfo, "rewl_stats": abi.Stats,
             "rewl_mlp_train_params": abi.MlpTrainParams}
    for key, val in got.items():
        if "." in key:
            t, m = key.split(".")
            assert getattr(pairs[t], m).offset == int(val), key


def test_engine_refuses_to_run_without_gpu():
    """On a machine without a GPU the engine must fail loudly, never compute on
    the CPU."""
    from rewl_simulator_amd import configs
    from rewl_simulator_amd.engine import Engine, device_count
    if device_count() > 0:
        pytest.skip("a GPU is visible")
    s, p = configs.config1()
    with pytest.raises(abi.REWLError) as e:
        Engine(s, p)
    assert e.value.code == abi.REWL_ERR_NO_DEVICE


def test_mlp_refuses_to_run_without_gpu():
    from rewl_simulator_amd.engine import device_count
    from rewl_simulator_amd.mlp import MLP
    if device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(abi.REWLError) as e:
        MLP(seed=1)
    assert e.value.code == abi.REWL_ERR_NO_DEVICE
