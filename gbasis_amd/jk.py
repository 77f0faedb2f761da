"""Coulomb (J) and exchange (K) matrices digested on the device straight from
the packed ERIs, without the dense N^4 array.

    J_mn = sum_ls (mn|ls) D_ls        K_ml = sum_ns (mn|ls) D_ns

The reference has no such entry point: its users build J and K (or their
energies) from ``electron_repulsion_integral``'s dense array, as
notebooks/tutorial/JCP_paper.ipynb (cell 28) does for ethylene in the MO basis
(sum_ij 2 (ii|jj) and sum_ij (ij|ij) over occupied orbitals).  For a
closed-shell density D = 2 C_occ C_occ^T those energies are tr(D J) / 2 and
tr(D K) / 4, which ``coulomb_exchange_energies`` returns.
"""

import numpy as np
import torch

from .electron_repulsion import electron_repulsion_integral_packed
from .engine import get_engine


def coulomb_exchange_packed(packed, nfunc, density, device="cuda"):
    """(J, K) as device float64 tensors from one molecule's packed ERIs (device)."""
    D = np.asarray(density, dtype=np.float64) if not torch.is_tensor(density) else density
    if not torch.is_tensor(D):
        tol = 1e-12 * max(1.0, float(np.abs(D).max())) if D.size else 0.0
        if not np.allclose(D, D.T, rtol=0, atol=tol):
            raise ValueError("density must be symmetric")
    return get_engine(device).jk(packed, nfunc, D)


def coulomb_exchange(basis, density, coord_types=None, schwarz_threshold=0.0, precision="fp64",
                     fp32_cutoff=1e-2, device="cuda"):
    """(J, K) as NumPy arrays for a basis (list of shells) and a symmetric density
    in its function order; the ERIs stay packed on the device."""
    packed, nfunc = electron_repulsion_integral_packed(basis, coord_types, schwarz_threshold,
                                                       precision, fp32_cutoff, device)
    J, K = coulomb_exchange_packed(packed, nfunc, density, device)
    return J.cpu().numpy(), K.cpu().numpy()


def coulomb_exchange_energies(J, K, density):
    """(E_J, E_K) = (tr(D J) / 2, tr(D K) / 4)."""
    D = np.asarray(density, dtype=np.float64)
    return float(np.sum(D * np.asarray(J)) / 2.0), float(np.sum(D * np.asarray(K)) / 4.0)
