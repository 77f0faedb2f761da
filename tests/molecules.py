"""Seeded synthetic molecules shared by the tests (shapes of the benchmark configs)."""
import numpy as np

from gbasis_amd.contractions import GeneralizedContractionShell


def random_basis(seed, shell_ls, n_atoms=None, box=3.0, kmin=1, kmax=3, emin=0.1, emax=10.0,
                 coord_type="cartesian", min_sep=0.0):
    """Shells with the given l's spread over n_atoms random centres."""
    rng = np.random.default_rng(seed)
    if n_atoms is None:
        n_atoms = len(shell_ls)
    centres = []
    while len(centres) < n_atoms:
        c = rng.uniform(0, box, 3)
        if all(np.linalg.norm(c - d) >= min_sep for d in centres):
            centres.append(c)
    basis = []
    for i, l in enumerate(shell_ls):
        k = int(rng.integers(kmin, kmax + 1))
        exps = np.exp(rng.uniform(np.log(emin), np.log(emax), k))
        coeffs = rng.uniform(0.1, 1.0, k)
        basis.append(GeneralizedContractionShell(int(l), centres[i % n_atoms].copy(), coeffs, exps,
                                                 coord_type, icenter=i % n_atoms))
    return basis
