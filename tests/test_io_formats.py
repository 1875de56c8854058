"""The reference's on-disk formats (QuantumMonteCarlo::DqmcIO, include/dqmc_io.h:267-629) as
written and read by the library's host code (no GPU needed): byte-for-byte against the files of
the reference's example run (example/output/*.out; the small ones are kept under golden/, the
large ones are pinned by their SHA-256, computed from the reference's files).
"""
import hashlib
import os

import numpy as np
import pytest

import dqmc_amd as D

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")

# SHA-256 of example/output/fields.out and dos.bins.out
FIELDS_SHA256 = "30af7efe14eff6aa45e52ca535de0c7caad358a8562ea744da9b9450c2a90e92"
DOS_BINS_SHA256 = "ff24bbe712cd6b69ce7df2d78e65bb37e9e5c4a3f592dc3024734014839e5ca1"
# SHA-256 of example/output/greens.out with every line reduced to its first four columns
# (indices, mean, error) joined by single spaces: the relative-error column cannot be
# recomputed exactly from the 10-decimal mean and error
GREENS_4COL_SHA256 = "8cade0c51021b58246e9dd09f000bf8d6ca8c26f86af5c60c37b68c03fabbfcc"


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


@pytest.fixture(scope="module")
def ex():
    return np.load(os.path.join(G, "example_observables_L160_N64.npz"))


def test_fields_file_byte_identical_and_round_trip(tmp_path):
    f = np.load(os.path.join(G, "example_fields_L160_N64.npy"))
    p = str(tmp_path / "fields.out")
    D.write_fields(p, f)
    assert sha(p) == FIELDS_SHA256
    assert np.array_equal(D.read_fields(p, 160, 64), f)


def test_fields_file_errors(tmp_path):
    f = np.load(os.path.join(G, "example_fields_L160_N64.npy"))
    p = str(tmp_path / "fields.out")
    D.write_fields(p, f)
    with pytest.raises(D.DqmcError):          # inconsistent time or space size
        D.read_fields(p, 80, 64)
    with pytest.raises(D.DqmcError):
        D.read_fields(str(tmp_path / "missing.out"), 160, 64)
    bad = tmp_path / "bad.out"
    bad.write_text("2 2\n0 0 1.0\n1 1 0.5\n")   # a field that is not +-1
    with pytest.raises(D.DqmcError):
        D.read_fields(str(bad), 2, 2)
    part = tmp_path / "part.out"
    part.write_text("2 2\n0 1 -1.0\n")          # entries left out keep their value
    assert np.array_equal(D.read_fields(str(part), 2, 2), np.array([[1, -1], [1, 1]], np.int8))


def test_kstars_and_tgrids_byte_identical(tmp_path):
    p = str(tmp_path / "kstars.out")
    D.write_kstars(p, 8)
    assert open(p).read() == open(os.path.join(G, "example_kstars_L8.out")).read()
    p = str(tmp_path / "tgrids.out")
    D.write_tgrids(p, 160, 8.0)
    assert open(p).read() == open(os.path.join(G, "example_tgrids.out")).read()


def test_scalar_observable_and_bins_byte_identical(tmp_path):
    ref = open(os.path.join(G, "example_swave.out")).read()
    m, e, _ = (float(x) for x in ref.split())
    p = str(tmp_path / "swave.out")
    D.write_observable(p, m, e)
    assert open(p).read() == ref
    ref_bins = open(os.path.join(G, "example_swave.bins.out")).read()
    bins = np.array([float(line.split()[1]) for line in ref_bins.splitlines()[1:]])
    p = str(tmp_path / "swave.bins.out")
    D.write_bins(p, bins)
    assert open(p).read() == ref_bins


@pytest.mark.parametrize("name", ["dos", "dss"])
def test_vector_observable_file(tmp_path, ex, name):
    p = str(tmp_path / f"{name}.out")
    D.write_observable(p, ex[name][0], ex[name][1])
    ours = open(p).read().splitlines()
    ref = open(os.path.join(G, f"example_{name}.out")).read().splitlines()
    assert len(ours) == len(ref) and ours[0] == ref[0]
    for a, b in zip(ours[1:], ref[1:]):
        assert a[:60] == b[:60]                           # index, mean, error: identical text
        # relative error: recomputed from the 10-decimal mean and error, so within the
        # propagated half-unit of their last digit (plus its own)
        _, m, e, r = (float(x) for x in b.split())
        bound = abs(r) * (0.5e-10 / abs(e) + 0.5e-10 / abs(m)) + 1e-10
        assert abs(float(a.split()[3]) - r) <= bound


def test_vector_bins_byte_identical(tmp_path, ex):
    p = str(tmp_path / "dos.bins.out")
    D.write_bins(p, ex["dos_bins"])
    assert sha(p) == DOS_BINS_SHA256


def test_matrix_observable_file(tmp_path, ex):
    p = str(tmp_path / "greens.out")
    D.write_observable(p, ex["greens"][0], ex["greens"][1])
    text = "\n".join(" ".join(line.split()[:4]) for line in open(p).read().splitlines())
    assert hashlib.sha256(text.encode()).hexdigest() == GREENS_4COL_SHA256


def test_matrix_bins_layout(tmp_path):
    # bin-major, then row, then column (include/dqmc_io.h:404-420)
    b = np.arange(2 * 2 * 3, dtype=np.float64).reshape(2, 2, 3) / 8.0
    p = str(tmp_path / "m.bins.out")
    D.write_bins(p, b)
    lines = open(p).read().splitlines()
    assert lines[0] == f"{2:20d}{2:20d}{3:20d}"
    rows = [tuple(float(x) for x in line.split()) for line in lines[1:]]
    assert rows == [(k, i, j, b[k, i, j]) for k in range(2) for i in range(2) for j in range(3)]
    assert len(lines[1]) == 80
