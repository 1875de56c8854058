"""Pins of the oracle's random stream and lattice tables.

* random stream: against golden streams written by libstdc++ itself
  (tests/golden/gen_rng.cpp), i.e. the generator the reference draws from
  (src/random.cpp:6-10, src/dqmc_walker.cpp:184, src/model/repulsive_hubbard.cpp:92-97);
* lattice: against the reference's example k-star table (example/output/kstars.out)
  and closed forms of the square lattice (src/lattice/square.cpp).
"""
import json
import math
import os

import numpy as np
import pytest

import _oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bits(hexstr, n):
    out = np.zeros(len(hexstr) * 4, np.uint8)
    for k, ch in enumerate(hexstr):
        v = int(ch, 16)
        for b in range(4):
            out[4 * k + b] = (v >> b) & 1
    return out[:n]


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLD, "rng_golden.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("seed", [0, 1, 42, 255, 12345, -7])
def test_bernoulli_half_stream(golden, seed):
    g = golden[str(seed)]
    ref = _bits(g["p05_bits"], 32768)
    got = O.bernoulli_stream(seed, 32768, 0.5)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("seed", [0, 42, -7])
def test_bernoulli_variable_p(golden, seed):
    ref = _bits(golden[str(seed)]["pvar_bits"], 20000)
    # the oracle draws with a fixed p per call: replay the same p schedule draw by draw
    got = np.zeros(20000, np.uint8)
    c = O.canonical_stream(seed, 20000)
    for k in range(20000):
        got[k] = 1 if c[k] < (k % 101) / 100.0 else 0
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("seed", [0, 1, 42, 255])
def test_canonical_bits(golden, seed):
    ref = [float.fromhex(x) for x in golden[str(seed)]["canonical"]]
    got = O.canonical_stream(seed, 64)
    assert [x.hex() for x in got] == [x.hex() for x in ref]


def test_seed_zero_equals_seed_one(golden):
    assert golden["0"]["raw"] == golden["1"]["raw"]
    assert golden["0"]["raw"][0] == 16807


def test_kstars_match_example_table():
    lat = O.lattice(8)
    with open(os.path.join(GOLD, "example_kstars_L8.out")) as f:
        lines = f.read().split("\n")
    nk = int(lines[0].split()[0])
    assert nk == 15 == lat["kstars"].shape[0]
    for k in range(nk):
        idx, kx, ky = lines[1 + k].split()
        assert int(idx) == k
        assert abs(float(kx) - lat["kstars"][k, 0]) < 1e-9
        assert abs(float(ky) - lat["kstars"][k, 1]) < 1e-9
    gamma, xp, mp = lat["points"]
    assert (gamma, xp, mp) == (0, 10, 14)
    assert np.allclose(lat["kstars"][mp], [math.pi, math.pi])
    assert np.allclose(lat["kstars"][xp], [math.pi, 0.0])


@pytest.mark.parametrize("side", [2, 3, 4, 6, 12])
def test_lattice_closed_forms(side):
    lat = O.lattice(side)
    n = side * side
    nk = (side // 2 + 1) * (side // 2 + 2) // 2
    assert lat["kstars"].shape == (nk, 2)
    x = np.arange(n) % side
    y = np.arange(n) // side
    nn = lat["nn"]
    assert np.array_equal(nn[:, 0], (x + 1) % side + side * y)
    assert np.array_equal(nn[:, 1], x + side * ((y + 1) % side))
    assert np.array_equal(nn[:, 2], (x - 1) % side + side * y)
    assert np.array_equal(nn[:, 3], x + side * ((y - 1) % side))
    disp = lat["disp"]
    for i in range(n):
        dx = (x - x[i]) % side
        dy = (y - y[i]) % side
        assert np.array_equal(disp[i], dx + side * dy)
    # hopping: periodic adjacency; for side 2 each bond is counted twice (+= on the same entry)
    H = lat["hopping"]
    A = np.zeros((n, n))
    for i in range(n):
        for d in (0, 1):
            A[i, nn[i, d]] += 1
            A[nn[i, d], i] += 1
    assert np.array_equal(H, A)
    assert np.allclose(H, H.T)
    if side > 2:
        assert set(np.unique(H)) <= {0.0, 1.0}
        assert np.all(H.sum(axis=0) == 4)
    else:
        # side 2: +x and -x neighbours coincide, so each bond is added twice
        assert set(np.unique(H)) == {0.0, 2.0}
        assert np.all(H.sum(axis=0) == 4)
    # Fourier factor cos(-(x kx + y ky))
    ks = lat["kstars"]
    expect = np.cos(-(np.outer(x, ks[:, 0]) + np.outer(y, ks[:, 1])))
    assert np.allclose(lat["fourier"], expect, atol=1e-13)
