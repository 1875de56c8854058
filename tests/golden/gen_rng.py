"""Regenerate tests/golden/rng_golden.json from gen_rng.cpp (g++ + libstdc++).

The golden file pins the oracle's restatement of std::default_random_engine
(minstd_rand0), std::generate_canonical and std::bernoulli_distribution, and
through it the device generator.  Run: python tests/golden/gen_rng.py
"""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

def main():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gen_rng")
        subprocess.check_call(["g++", "-O2", "-std=c++17", os.path.join(HERE, "gen_rng.cpp"), "-o", exe])
        out = subprocess.check_output([exe]).decode()
    with open(os.path.join(HERE, "rng_golden.json"), "w") as f:
        f.write(out)

if __name__ == "__main__":
    main()
