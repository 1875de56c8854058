"""The C++ facade (dqmc-framework_amd/facade/dqmc_framework.hpp) driven the way
the reference's main drives its modules (set_seed, initial_modules, random
fields, initial_dqmc, Dqmc::thermalize / measure / analyse) runs the same
Markov chain as the oracle's driver with the same seed and schedule, so the
binned observables agree to rounding."""
import json
import os
import subprocess

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "dqmc-framework_amd", "lib", "reference_flow")
ORACLE = os.path.join(ROOT, "oracle", "_build", "refcpu")

NAMES = [("filling_number", "filling"), ("double_occupancy", "double_occupancy"),
         ("kinetic_energy", "kinetic_energy"), ("momentum_distribution", "momentum_distribution"),
         ("local_spin_corr", "local_spin_corr"), ("spin_density_structure_factor", "sdw"),
         ("charge_density_structure_factor", "cdw"), ("s_wave_pairing_corr", "swave")]


@pytest.mark.parametrize("cfg", [
    # side L beta pace model cb U mu seed warm bins bin_size between
    (4, 80, 4.0, 10, 0, 0, 4.0, 0.0, 42, 20, 4, 10, 2),
    (4, 40, 2.0, 10, 1, 1, 4.0, -0.5, 7, 10, 3, 6, 0),
])
def test_facade_driver_matches_oracle_driver(cfg):
    print(f"[gpu-facade] {cfg}", flush=True)
    side, L, beta, pace, model, cb, U, mu, seed, warm, bins, bin_size, between = cfg
    args = [str(x) for x in cfg]
    out = subprocess.run([DRIVER] + args, check=True, capture_output=True, text=True, timeout=600).stdout
    got = json.loads(out.strip().splitlines()[-1])
    orc = subprocess.run([ORACLE, "--side", str(side), "--slices", str(L), "--beta", str(beta), "--pace", str(pace),
                          "--model", str(model), "--cb", str(cb), "--U", str(U), "--mu", str(mu),
                          "--seed", str(seed), "--warm", str(warm), "--bins", str(bins),
                          "--bin-size", str(bin_size), "--between", str(between)],
                         check=True, capture_output=True, text=True, timeout=600).stdout
    ref = json.loads(orc.strip().splitlines()[-1])
    for ours, theirs in NAMES:
        m, e = got[ours]
        rm, re_ = ref[theirs]
        assert abs(m - rm) < 1e-9 * max(1.0, abs(rm)), (ours, m, rm)
        assert abs(e - re_) < 1e-8 * max(1.0, abs(re_)) + 1e-12, (ours, e, re_)


DYN_NAMES = ["dynamic_sign", "superfluid_stiffness", "greens_functions", "density_of_states",
             "dynamic_spin_susceptibility"]


@pytest.mark.parametrize("cfg", [
    # side L beta pace model cb U mu seed warm bins bin_size between
    (4, 40, 3.0, 10, 0, 0, 4.0, -0.2, 42, 10, 3, 6, 2),
    (4, 24, 2.0, 6, 1, 1, 4.0, -0.5, 7, 6, 3, 4, 0),
])
def test_facade_dynamic_driver_matches_oracle_driver(cfg, tmp_path):
    # the measuring forth-and-back starts with the time-displaced sweep when the handler
    # measures dynamic observables (src/dqmc.cpp:48-70); equal-time ones come from the
    # backward sweeps only
    print(f"[gpu-facade] dynamic {cfg}", flush=True)
    side, L, beta, pace, model, cb, U, mu, seed, warm, bins, bin_size, between = cfg
    args = [str(x) for x in cfg] + ["1", str(tmp_path)]
    out = subprocess.run([DRIVER] + args, check=True, capture_output=True, text=True, timeout=600).stdout
    got = json.loads(out.strip().splitlines()[-1])
    orc = subprocess.run([ORACLE, "--side", str(side), "--slices", str(L), "--beta", str(beta), "--pace", str(pace),
                          "--model", str(model), "--cb", str(cb), "--U", str(U), "--mu", str(mu),
                          "--seed", str(seed), "--warm", str(warm), "--bins", str(bins),
                          "--bin-size", str(bin_size), "--between", str(between), "--dynamic", "1"],
                         check=True, capture_output=True, text=True, timeout=600).stdout
    ref = json.loads(orc.strip().splitlines()[-1])
    for ours, theirs in NAMES:
        m, e = got[ours]
        rm, re_ = ref[theirs]
        assert abs(m - rm) < 1e-9 * max(1.0, abs(rm)), (ours, m, rm)
    for name in DYN_NAMES:
        m, e = (np.array(x) for x in got[name])
        rm, re_ = (np.array(x) for x in ref[name])
        assert m.shape == rm.shape, name
        assert np.max(np.abs(m - rm)) < 1e-8, (name, np.max(np.abs(m - rm)))
        assert np.max(np.abs(e - re_)) < 1e-7, (name, np.max(np.abs(e - re_)))
    # the reference's output files (src/dqmc_main.cpp:270-360), written by the facade's DqmcIO
    import dqmc_amd as D
    f = D.read_fields(str(tmp_path / "fields.out"), L, side * side)
    assert set(np.unique(f)) <= {-1, 1}
    sw = [float(x) for x in open(tmp_path / "swave.out").read().split()]
    assert abs(sw[0] - got["s_wave_pairing_corr"][0]) < 1e-9 * max(1.0, abs(sw[0]))
    lines = open(tmp_path / "dos.out").read().splitlines()
    assert int(lines[0]) == L and len(lines) == L + 1
    dos = np.array([float(x.split()[1]) for x in lines[1:]])
    assert np.max(np.abs(dos - np.array(got["density_of_states"][0]))) < 1e-9
    head = open(tmp_path / "greens.out").readline().split()
    nk = (side // 2 + 1) * (side // 2 + 2) // 2
    assert [int(x) for x in head] == [nk, L]
    head = open(tmp_path / "dss.bins.out").readline().split()
    assert [int(x) for x in head] == [bins, L]
    assert open(tmp_path / "tgrids.out").readline().split()[0] == str(L)


BLOCKS = os.path.join(ROOT, "dqmc-framework_amd", "lib", "building_blocks")


@pytest.mark.parametrize("cfg", [
    # side L beta pace model cb U mu seed
    (4, 40, 4.0, 10, 0, 0, 4.0, -0.2, 42),
    (6, 30, 3.0, 7, 1, 1, 4.0, -0.5, 7),
])
def test_facade_building_blocks_match_oracle(cfg):
    # ModelBase::mult_B_from_left, SvdStack::push and NumericalStable::compute_equaltime_greens
    # of the facade, used as DqmcWalker::initial_svd_stacks / initial_greens_functions use them
    # (src/dqmc_walker.cpp:109-147), against the oracle's B(t) and its SVD-stack G(0)
    print(f"[gpu-facade] building blocks {cfg}", flush=True)
    side, L, beta, pace, model, cb, U, mu, seed = cfg
    out = subprocess.run([BLOCKS] + [str(x) for x in cfg], check=True, capture_output=True, text=True,
                         timeout=300).stdout
    got = json.loads(out.strip().splitlines()[-1])
    n = side * side
    orc = O.Chain(side, L, beta, pace=pace, model=model, checkerboard=bool(cb), U=U, mu=mu)
    orc.seed(seed)
    orc.fields_random()
    assert np.array_equal(np.array(got["fields"]).reshape(L, n), orc.fields())
    eye = np.asfortranarray(np.eye(n))
    for key, t in (("B1", 1), ("B0", 0)):
        want = orc.mult(0, t, 1, eye)
        assert np.max(np.abs(np.array(got[key]).reshape(n, n, order="F") - want)) < 1e-13 * np.max(np.abs(want))
    assert got["stack_length"] == (L + pace - 1) // pace
    orc.init()
    g0 = orc.green(1)
    for key in ("g0_stack", "g0_engine"):
        err = np.max(np.abs(np.array(got[key]).reshape(n, n, order="F") - g0))
        print(f"[gpu-facade] {key}: max |dG| {err:.2e}", flush=True)
        assert err < 1e-10, key
