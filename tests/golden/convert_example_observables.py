"""Pack the statistical results of the reference's example run (example/config.toml: 8 x 8
attractive Hubbard, U = 4, mu = 0, beta = 8, 160 slices, dense, pace 10; 20 bins of 50
forth-and-back sweeps on 10 ranks) into example_observables_L160_N64.npz:

  dos, dss        (2, 160)      mean and error of D(tau), DSS(tau)      (output/dos.out, dss.out)
  greens          (2, 15, 160)  mean and error of G(k,tau) per k-star   (output/greens.out)
  scalar_names, scalars (n, 2)  equal-time observables and rho_s       (log.out)
  dos_bins        (20, 160)     the bins of D(tau)                      (output/dos.bins.out)

The text formats are those written by include/dqmc_io.h:267-429 (a header with the sizes, then
one line per entry: indices, mean, error, relative error).  Run from the repository root with the
reference mounted:
  python tests/golden/convert_example_observables.py /root/reference/example
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

LOG_NAMES = {"Filling number": "filling_number", "Double occupation": "double_occupancy",
             "Kinetic energy": "kinetic_energy", "Local spin correlation": "local_spin_corr",
             "Momentum distribution": "momentum_distribution", "SDW order parameter": "spin_density_structure_factor",
             "CDW order parameter": "charge_density_structure_factor",
             "S-wave pairing correlation": "s_wave_pairing_corr", "Superfluid stiffness": "superfluid_stiffness"}


def read_vector(path):
    with open(path) as f:
        n = int(f.readline().split()[0])
        out = np.zeros((2, n))
        for line in f:
            p = line.split()
            if len(p) == 4:
                out[0, int(p[0])], out[1, int(p[0])] = float(p[1]), float(p[2])
    return out


def read_matrix(path):
    with open(path) as f:
        nk, nt = (int(x) for x in f.readline().split()[:2])
        out = np.zeros((2, nk, nt))
        for line in f:
            p = line.split()
            if len(p) == 5:
                k, t = int(p[0]), int(p[1])
                out[0, k, t], out[1, k, t] = float(p[2]), float(p[3])
    return out


def read_vector_bins(path):
    with open(path) as f:
        nb, n = (int(x) for x in f.readline().split()[:2])
        out = np.zeros((nb, n))
        for line in f:
            p = line.split()
            if len(p) == 3:
                out[int(p[0]), int(p[1])] = float(p[2])
    return out


def read_log(path):
    names, vals = [], []
    pat = re.compile(r"^\s*(.+?)\s+->\s+(\S+)\s+pm\s+(\S+)")
    with open(path) as f:
        for line in f:
            m = pat.match(line)
            if m and m.group(1).strip() in LOG_NAMES:
                names.append(LOG_NAMES[m.group(1).strip()])
                vals.append((float(m.group(2)), float(m.group(3))))
    return np.array(names), np.array(vals)


if __name__ == "__main__":
    ex = sys.argv[1]
    names, vals = read_log(os.path.join(ex, "log.out"))
    np.savez_compressed(os.path.join(HERE, "example_observables_L160_N64.npz"),
                        dos=read_vector(os.path.join(ex, "output", "dos.out")),
                        dss=read_vector(os.path.join(ex, "output", "dss.out")),
                        greens=read_matrix(os.path.join(ex, "output", "greens.out")),
                        scalar_names=names, scalars=vals,
                        dos_bins=read_vector_bins(os.path.join(ex, "output", "dos.bins.out")))
    print(list(names))
