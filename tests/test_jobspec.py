"""TOML job specifications with the reference's keys, defaults and checks
(DqmcInitializer::parse_toml_config, src/dqmc_initializer.cpp:27-338), on the CPU."""
import os

import numpy as np
import pytest

import _oracle as O
import dqmc_amd as D
import jobspec as J

HERE = os.path.dirname(os.path.abspath(__file__))


def test_example_config():
    s = J.load(os.path.join(HERE, "golden", "example_config.toml"))
    assert s.model == D.ATTRACTIVE and s.onsite_u == 4.0 and s.hopping_t == 1.0 and s.chemical_potential == 0.0
    assert s.side == 8 and not s.checkerboard
    assert (s.beta, s.time_size, s.pace) == (8.0, 160, 10)
    assert (s.sweeps_warmup, s.bin_num, s.bin_size, s.sweeps_between_bins) == (512, 20, 100, 20)
    assert s.observables == J.ALL and s.is_equaltime and s.is_dynamic
    lat = O.lattice(8)
    assert s.momentum == lat["points"][2]                      # MPoint
    assert s.momentum_list == list(range(15))                  # KstarsAll
    p = s.params(seed=3)
    assert (p.side, p.time_size, p.model, p.momentum_index, p.seed) == (8, 160, D.ATTRACTIVE, 14, 3)
    # split over 10 processes as the reference's example run was
    assert J.load(os.path.join(HERE, "golden", "example_config.toml"), world_size=10).bins_per_proc == 2
    assert J.load(os.path.join(HERE, "golden", "example_config.toml"), world_size=3).bins_per_proc == 7


def test_defaults_and_keywords():
    s = J.load(text='[Lattice]\ncell = [4, 4]\nmomentum = "XPoint"\nmomentum_list = "DeltaLine"\n'
                    '[Measure]\nobservables = ["none"]\n[MonteCarlo]\ntime_size = 50.7\n')
    assert s.model == D.REPULSIVE and (s.hopping_t, s.onsite_u, s.chemical_potential) == (1.0, 4.0, 0.0)
    assert s.beta == 4.0 and s.time_size == 50 and s.pace == 10        # time_size truncated
    assert (s.sweeps_warmup, s.bin_num, s.bin_size, s.sweeps_between_bins) == (512, 20, 100, 20)
    assert s.observables == [] and not s.is_equaltime and not s.is_dynamic
    assert s.momentum == O.lattice(4)["points"][1]


@pytest.mark.parametrize("text,msg", [
    ('[Model]\ntype = "Ising"', "undefined model type"),
    ('[Lattice]\ncell = [4]', "cell"),
    ('[Lattice]\ncell = [4, 0]', "cell"),
    ('[Lattice]\ncell = [4, 6]\n[Measure]\nobservables = ["all"]', "equal sides"),
    ('[Lattice]\ntype = "Cubic"\ncell = [4, 4, 4]', "not supported"),
    ('[Lattice]\ncell = [4, 4]\nmomentum = "MPoint"\nmomentum_list = "KstarsAll"', "undefined observables"),
    ('[Lattice]\ncell = [4, 4]\nmomentum = "YPoint"\nmomentum_list = "KstarsAll"\n[Measure]\nobservables = ["all"]',
     "undefined momentum"),
    ('[Lattice]\ncell = [4, 4]\nmomentum = "MPoint"\nmomentum_list = "Loop"\n[Measure]\nobservables = ["all"]',
     "undefined momentum list"),
    ('[Lattice]\ncell = [4, 4]\nmomentum = "MPoint"\nmomentum_list = "KstarsAll"\n[Measure]\nobservables = ["x"]',
     "undefined observable"),
])
def test_config_errors(text, msg):
    with pytest.raises(J.ConfigError, match=msg):
        J.load(text=text)


@pytest.mark.parametrize("side", [4, 6, 8, 12])
def test_momentum_lines_follow_the_brillouin_zone(side):
    # pinned by geometry: k-stars of each line/point against the lattice's k-star table
    k = O.lattice(side)["kstars"]
    pi = np.pi
    g, x, m = (J.momentum_index(side, n) for n in ("GammaPoint", "XPoint", "MPoint"))
    assert np.allclose(k[g], [0, 0]) and np.allclose(k[x], [pi, 0]) and np.allclose(k[m], [pi, pi])
    steps = 2 * pi / side * np.arange(side // 2 + 1)
    delta = k[J.momentum_list(side, "DeltaLine")]
    assert np.allclose(delta[:, 0], steps) and np.allclose(delta[:, 1], 0)
    z = k[J.momentum_list(side, "ZLine")]
    assert np.allclose(z[:, 0], pi) and np.allclose(z[:, 1], steps)
    sig = k[J.momentum_list(side, "SigmaLine")]
    assert np.allclose(sig[:, 0], steps) and np.allclose(sig[:, 1], steps)
    loop = J.momentum_list(side, "Gamma2X2M2GammaLoop")
    h = side // 2
    assert loop == (J.momentum_list(side, "DeltaLine")[:h] + J.momentum_list(side, "ZLine")[:h]
                    + J.momentum_list(side, "SigmaLine")[h:0:-1])
    assert J.momentum_list(side, "KstarsAll") == list(range(len(k)))


def test_serving_job_from_spec():
    s = J.load(os.path.join(HERE, "golden", "example_config.toml"))
    j = J.job(s, seed=5)
    assert j["warmup"] == 512 and j["bin_size"] == 100 and j["measure"] == 2000
    assert j["params"].seed == 5 and j["params"].side == 8


def test_serving_jobs_from_config_files(tmp_path):
    import serve
    p = tmp_path / "small.toml"
    p.write_text('[Model]\ntype = "AttractiveHubbard"\n[Model.Params]\nonsite_u = 3.0\n'
                 '[Lattice]\ncell = [6, 6]\nmomentum = "MPoint"\nmomentum_list = "KstarsAll"\n'
                 '[MonteCarlo]\nbeta = 2.0\ntime_size = 20\n'
                 '[Measure]\nsweeps_warmup = 11\nbin_num = 3\nbin_size = 7\nobservables = ["double_occupancy"]\n')
    jobs = serve.jobs_from_configs([os.path.join(HERE, "golden", "example_config.toml"), str(p)])
    assert [j["id"] for j in jobs] == [0, 1]
    j = jobs[1]
    assert (j["warmup"], j["measure"], j["bin_size"]) == (10, 18, 6)     # whole forth-and-back iterations
    assert j["params"].side == 6 and j["params"].time_size == 20 and j["params"].model == D.ATTRACTIVE
    assert j["params"].onsite_u == 3.0 and j["params"].seed == 1
