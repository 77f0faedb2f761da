"""Timing worker of the CPU baseline.  TEST INFRASTRUCTURE ONLY: run by the cpu_baseline
leg of bench.py, as one of several processes that time the oracle side by side.

    python -m oracle.rate SAMPLES.json SEED BUDGET_S

SAMPLES.json holds ``{"classes": [[quartet, ...], ...]}`` with a quartet as four shells
``[l, [x, y, z], [exps], [coeffs]]``.  The worker visits the classes in the given order
with an equal share of BUDGET_S each, times ``oracle.assembly.contraction_block`` on
quartets drawn at random (seeded by SEED) from the class until its share is spent (at
least one quartet), and prints one JSON line: per class [seconds, quartets].
"""

import json
import sys
import time

import numpy as np


class Shell:
    """A contracted Cartesian shell with the attributes the oracle reads."""

    def __init__(self, l, coord, exps, coeffs):
        self.angmom = int(l)
        self.coord = np.asarray(coord, dtype=np.float64)
        self.exps = np.asarray(exps, dtype=np.float64)
        self.coeffs = np.asarray(coeffs, dtype=np.float64)[:, None]
        self.angmom_components_cart = np.array(
            [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)],
            dtype=np.int64).reshape(-1, 3)


def time_classes(classes, seed, budget_s):
    from .assembly import contraction_block

    rng = np.random.default_rng(seed)
    share = budget_s / max(1, len(classes))
    out = []
    for quartets in classes:
        spent, n = 0.0, 0
        while n == 0 or spent < share:
            shells = [Shell(*s) for s in quartets[rng.integers(0, len(quartets))]]
            t0 = time.perf_counter()
            contraction_block(*shells)
            spent += time.perf_counter() - t0
            n += 1
        out.append([spent, n])
    return out


def main(argv):
    path, seed, budget = argv[1], int(argv[2]), float(argv[3])
    with open(path) as f:
        classes = json.load(f)["classes"]
    print(json.dumps(time_classes(classes, seed, budget)), flush=True)


if __name__ == "__main__":
    main(sys.argv)
