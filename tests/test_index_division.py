"""The engine's staging loops divide by wave-uniform sizes with a multiply-shift
(q = floor(e m / 2^32), m = ceil(2^32 / d), struct FastDiv in csrc/dqmc_kernels.hip)
instead of the integer divide sequence.  Exhaustive check on the CPU that it is
floor division for every divisor the engine uses (1..448) and every index below
2^18 (the staging indices stay below N^2 <= 448^2), and that the plaquette corners it
yields are the ones of the modular form (CheckerBoard::Square, src/checkerboard/square.cpp:28-50)."""
import numpy as np


def fastdiv(e, d):
    m = ((1 << 32) + d - 1) // d
    return ((e.astype(np.uint64) * np.uint64(m)) >> np.uint64(32)).astype(np.int64)


def test_multiply_shift_is_floor_division():
    e = np.arange(1 << 18, dtype=np.int64)
    for d in range(1, 449):
        assert np.array_equal(fastdiv(e, d), e // d), d


def test_plaquette_corners():
    for side in range(2, 22, 2):
        h = side // 2
        p = np.arange(h * h, dtype=np.int64)
        for start in (0, 1):
            py = fastdiv(p, h)
            px = p - py * h
            x, y = start + 2 * px, start + 2 * py
            x1 = np.where(x + 1 == side, 0, x + 1)
            y1 = np.where(y + 1 == side, 0, y + 1)
            fast = np.stack([x + side * y, x1 + side * y, x + side * y1, x1 + side * y1])
            xm, ym = start + 2 * (p % h), start + 2 * (p // h)
            ref = np.stack([xm + side * ym, (xm + 1) % side + side * ym, xm + side * ((ym + 1) % side),
                            (xm + 1) % side + side * ((ym + 1) % side)])
            assert np.array_equal(fast, ref), (side, start)
