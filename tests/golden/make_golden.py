"""Writes tests/golden/sobol_ref.npz: Sobol points from the reference's own
generator header (c++/keldy/qrng_details/digitalseq_b2g.hpp), compiled by
oracle/Makefile (`make -C oracle ref`) into oracle/_ref/sobol_ref and driven
through the keldy::sobol set-up (qrng.hpp:69-109), plus the first outputs of
std::mt19937_64.  Run in the development container (needs the reference).
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

# (dim, state_seed, do_shift, do_scramble, rng_seed_shift, n_points)
CASES = [(1, 0, 0, 0, 0, 64), (3, 0, 0, 0, 0, 256), (3, 0, 1, 1, 0, 256), (6, 0, 1, 1, 7, 256),
         (8, 0, 1, 1, 0, 256), (8, 5, 1, 0, 3, 128), (8, 0, 0, 1, 11, 128), (10, 100, 1, 1, 15, 128),
         (12, 0, 1, 1, 2, 64), (32, 0, 1, 1, 1, 32), (2, 2147483600, 0, 0, 0, 48)]


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = {}
    for i, c in enumerate(CASES):
        out[f"case{i}"] = np.array(c, dtype=np.int64)
        out[f"pts{i}"] = O.ref_sobol_points(*c)
    for seed in (0, 5489):
        r = subprocess.run([O.REF_SOBOL, "mt64", str(seed), "16"], check=True, capture_output=True, text=True)
        out[f"mt64_{seed}"] = np.array([int(x) for x in r.stdout.split()], dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "sobol_ref.npz"), **out)
    print("wrote", os.path.join(HERE, "sobol_ref.npz"), "with", len(CASES), "cases")


if __name__ == "__main__":
    main()
