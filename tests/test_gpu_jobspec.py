"""A TOML configuration run through jobspec.run (the reference's main: thermalize, measure,
analyse, write outputs) on the GPU equals the oracle's driver on the same schedule and seed."""
import json
import os
import subprocess

import numpy as np
import pytest

import dqmc_amd as D
import jobspec as J

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "_build", "refcpu")

CONFIG = """
[Model]
    type = "RepulsiveHubbard"
    [Model.Params]
    hopping_t = 1.0
    onsite_u = 4.0
    chemical_potential = -0.2
[Lattice]
    type = "Square"
    cell = [ 4, 4 ]
    momentum = "MPoint"
    momentum_list = "KstarsAll"
[CheckerBoard]
    whether_or_not = false
[MonteCarlo]
    beta = 3.0
    time_size = 40
    stabilization_pace = 10
[Measure]
    sweeps_warmup = 10
    bin_num = 3
    bin_size = 6
    sweeps_between_bins = 2
    observables = [ "all" ]
"""


def test_config_run_matches_oracle_driver(tmp_path):
    print("[gpu-jobspec] 4x4 config, one chain", flush=True)
    spec = J.load(text=CONFIG)
    ctx = D.Context(0)
    res = J.run(spec, ctx, chains=1, seed=42, out_dir=str(tmp_path))
    orc = subprocess.run([ORACLE, "--side", "4", "--slices", "40", "--beta", "3.0", "--pace", "10", "--model", "0",
                          "--cb", "0", "--U", "4.0", "--mu", "-0.2", "--seed", "42", "--warm", "10", "--bins", "3",
                          "--bin-size", "6", "--between", "2", "--dynamic", "1"],
                         check=True, capture_output=True, text=True, timeout=600).stdout
    ref = json.loads(orc.strip().splitlines()[-1])
    pairs = [("filling_number", "filling"), ("double_occupancy", "double_occupancy"),
             ("kinetic_energy", "kinetic_energy"), ("momentum_distribution", "momentum_distribution"),
             ("local_spin_corr", "local_spin_corr"), ("spin_density_structure_factor", "sdw"),
             ("charge_density_structure_factor", "cdw"), ("s_wave_pairing_corr", "swave")]
    for ours, theirs in pairs:
        assert abs(res[ours][0] - ref[theirs][0]) < 1e-9 * max(1.0, abs(ref[theirs][0])), ours
        assert abs(res[ours][1] - ref[theirs][1]) < 1e-8 * max(1.0, abs(ref[theirs][1])) + 1e-12, ours
    for name in ["superfluid_stiffness", "greens_functions", "density_of_states", "dynamic_spin_susceptibility"]:
        m = np.ravel(res[name][0])
        rm = np.array(ref[name][0])
        if name == "greens_functions":     # the oracle prints (k, t) at k + n_k t
            rm = rm.reshape(40, 6).T
        rm = np.ravel(rm)
        assert m.shape == rm.shape and np.max(np.abs(m - rm)) < 1e-8, name
    # the reference's output files
    for f in ["fields.out", "kstars.out", "tgrids.out", "swave.out", "swave.bins.out", "dos.out", "dos.bins.out",
              "greens.out", "greens.bins.out", "dss.out", "dss.bins.out"]:
        assert (tmp_path / f).exists(), f
    assert [int(x) for x in open(tmp_path / "greens.bins.out").readline().split()] == [3, 6, 40]
    J.report(res)
    ctx.close()


def test_config_with_many_chains_and_a_field_file(tmp_path):
    # bins shared by chains (the GPU's ranks), every chain starting from one field file
    print("[gpu-jobspec] 4 chains from a field file", flush=True)
    spec = J.load(text=CONFIG.replace("bin_num = 3", "bin_num = 8"))
    f = np.where(np.random.default_rng(1).random((40, 16)) < 0.5, 1, -1).astype(np.int8)
    D.write_fields(str(tmp_path / "in.out"), f)
    ctx = D.Context(0)
    res = J.run(spec, ctx, chains=4, seed=7, fields=D.read_fields(str(tmp_path / "in.out"), 40, 16))
    assert res["bins"] == 8
    assert res["equaltime_sign"][0] > 0.0 and res["dynamic_sign"][0] > 0.0
    assert res["greens_functions"][0].shape == (6, 40)
    ctx.close()
