"""Pack the reference's example HS-field snapshot (example/output/fields.out,
text format: header "L N", then one "t i s" line per site, written by
include/dqmc_io.h:479-534) into example_fields_L160_N64.npy (int8, shape L x N).

Run from the repository root with the reference mounted:
  python tests/golden/convert_example_fields.py /root/reference/example/output/fields.out
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_fields(path):
    with open(path) as f:
        head = f.readline().split()
        L, N = int(head[0]), int(head[1])
        arr = np.zeros((L, N), np.int8)
        for line in f:
            parts = line.split()
            if len(parts) != 3:
                continue
            t, i, s = int(parts[0]), int(parts[1]), float(parts[2])
            arr[t, i] = 1 if s > 0 else -1
    return arr


if __name__ == "__main__":
    a = read_fields(sys.argv[1])
    np.save(os.path.join(HERE, "example_fields_L160_N64.npy"), a)
    print(a.shape, int((a > 0).sum()), int((a < 0).sum()))
