"""Write tests/golden/reference_vectors.npz from the reference itself.

Run once, in a container that holds the reference at /root/reference (it never
travels with this repository):

    python tests/golden/make_reference_vectors.py

It pins what only the reference's own functions fix: its Schwarz bound
matrices (gbasis/integrals/_schwarz_screening.py:62-90), from which the screened
quartet lists follow, and its FP64 ERIs (gbasis/integrals/electron_repulsion.py:415-504)
on the dense s,s,p,d workload (4 atoms, N = 44, seed 1): aggregates of the whole
array and a seeded sample of entries.  The shells are the ones
gbasis_amd.workloads generates, rebuilt as the reference's shell objects from
the same numbers.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, "/root/reference")

from gbasis.contractions import GeneralizedContractionShell as RefShell  # noqa: E402
from gbasis.integrals._schwarz_screening import compute_schwarz_bounds  # noqa: E402
from gbasis.integrals._two_elec_int_improved import (  # noqa: E402
    compute_two_electron_integrals_os_hgp)
from gbasis.integrals.boys_functions import get_boys_function  # noqa: E402
from gbasis.integrals.electron_repulsion import (  # noqa: E402
    electron_repulsion_integral_improved)

from gbasis_amd.workloads import batch_molecule, dense_small  # noqa: E402


def _ref_basis(basis):
    return [RefShell(s.angmom, np.array(s.coord, dtype=float), np.array(s.coeffs, dtype=float),
                     np.array(s.exps, dtype=float), s.coord_type) for s in basis]


def main():
    out = {}
    boys = get_boys_function("coulomb")
    t0 = time.time()
    basis = dense_small(1)
    ref = _ref_basis(basis)
    out["dense_small_norm_cont"] = np.concatenate([r.norm_cont.ravel() for r in ref])
    out["dense_small_Q"] = compute_schwarz_bounds(ref, boys, compute_two_electron_integrals_os_hgp)
    eri = electron_repulsion_integral_improved(ref, notation="chemist")
    n = eri.shape[0]
    out["dense_small_n"] = np.array(n)
    out["dense_small_sum"] = np.array(eri.sum())
    out["dense_small_sumsq"] = np.array((eri * eri).sum())
    out["dense_small_sumabs"] = np.array(np.abs(eri).sum())
    rng = np.random.default_rng(20260101)
    idx = rng.integers(0, n, size=(4000, 4))
    out["dense_small_idx"] = idx.astype(np.int16)
    out["dense_small_val"] = eri[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]
    print(f"dense_small: N={n} in {time.time() - t0:.1f} s", flush=True)
    for seed in (1000, 1001):
        t0 = time.time()
        ref = _ref_basis(batch_molecule(seed))
        out[f"batch{seed}_Q"] = compute_schwarz_bounds(ref, boys,
                                                       compute_two_electron_integrals_os_hgp)
        print(f"batch molecule {seed}: {len(ref)} shells in {time.time() - t0:.1f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **out)


if __name__ == "__main__":
    main()
