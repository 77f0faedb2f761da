"""Seeded synthetic molecules for the benchmark workloads.

Each generator returns plain shell lists (``GeneralizedContractionShell``)
built from uniform-random atom positions with a minimum separation, random
shell compositions per atom, exponents log-uniform in a range and
coefficients uniform in [0.1, 1] rescaled so every contracted function is
normalised (``norm_cont`` = 1).
"""

import numpy as np

from .contractions import GeneralizedContractionShell


def _positions(rng, n_atoms, side, min_sep):
    pts = []
    tries = 0
    while len(pts) < n_atoms:
        p = rng.uniform(0.0, side, 3)
        tries += 1
        if all(np.linalg.norm(p - q) >= min_sep for q in pts):
            pts.append(p)
        if tries > 100000:
            raise RuntimeError("could not place atoms with the requested separation")
    return np.array(pts)


def _shell(rng, l, coord, k, emin, emax, coord_type, icenter):
    exps = np.exp(rng.uniform(np.log(emin), np.log(emax), k))
    coeffs = rng.uniform(0.1, 1.0, k)
    sh = GeneralizedContractionShell(int(l), coord.copy(), coeffs, exps, coord_type,
                                     icenter=icenter)
    sh.coeffs = sh.coeffs * sh.norm_cont[0, 0]   # normalise the contraction
    sh.assign_norm_cont()
    return sh


def dense_small(seed=1):
    """4 atoms in a 4 Bohr cube (min separation 1.2); s,s,p,d Cartesian per atom, K=3,
    exponents log-uniform in [0.1, 50]; N = 44."""
    rng = np.random.default_rng(seed)
    pos = _positions(rng, 4, 4.0, 1.2)
    basis = []
    for ia, c in enumerate(pos):
        for l in (0, 0, 1, 2):
            basis.append(_shell(rng, l, c, 3, 0.1, 50.0, "cartesian", ia))
    return basis


def batch_molecule(seed, n_atoms=None, lmax=2, kmax=6):
    """One molecule of the 64-molecule batch: 4-24 atoms in a cube of side
    3 n^(1/3) Bohr (min separation 1.2); per atom 1-3 s, 0-2 p, 0-1 d shells;
    K uniform in 1..6; exponents log-uniform in [0.05, 1e3]."""
    rng = np.random.default_rng(seed)
    if n_atoms is None:
        n_atoms = int(rng.integers(4, 25))
    pos = _positions(rng, n_atoms, 3.0 * n_atoms ** (1.0 / 3.0), 1.2)
    basis = []
    for ia, c in enumerate(pos):
        ls = [0] * int(rng.integers(1, 4)) + [1] * int(rng.integers(0, 3)) + \
             [2] * int(rng.integers(0, 2))
        for l in ls:
            if l > lmax:
                continue
            basis.append(_shell(rng, l, c, int(rng.integers(1, kmax + 1)), 0.05, 1e3,
                                "cartesian", ia))
    return basis


def batch64(base_seed=1000, count=64):
    return [batch_molecule(base_seed + m) for m in range(count)]


def large_spdf(seed=3):
    """16 atoms in a 10 Bohr cube (min separation 1.4); s,p,d,f Cartesian per atom, K=4,
    exponents log-uniform in [0.08, 2e3]; N = 320."""
    rng = np.random.default_rng(seed)
    pos = _positions(rng, 16, 10.0, 1.4)
    return [_shell(rng, l, c, 4, 0.08, 2e3, "cartesian", ia) for ia, c in enumerate(pos)
            for l in (0, 1, 2, 3)]


def high_l_pure(seed=5):
    """6 atoms in a 6 Bohr cube; s,p,d,f,g pure shells per atom, K=8, exponents
    log-uniform in [0.02, 1e5]; N = 150."""
    rng = np.random.default_rng(seed)
    pos = _positions(rng, 6, 6.0, 1.2)
    return [_shell(rng, l, c, 8, 0.02, 1e5, "spherical", ia) for ia, c in enumerate(pos)
            for l in (0, 1, 2, 3, 4)]


def serving_request(seed):
    """A serving request: 2-30 atoms built as batch_molecule, l <= 2, K <= 6."""
    rng = np.random.default_rng(seed)
    return batch_molecule(seed, n_atoms=int(rng.integers(2, 31)))
