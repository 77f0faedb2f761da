"""Schwarz screening with device-computed bounds: the interface of
gbasis/integrals/_schwarz_screening.py (compute_schwarz_bound_shell_pair :14-59,
compute_schwarz_bounds :62-90, shell_pair_significant :93-128,
SchwarzScreener :131-218).

Q_ij = sqrt(max |(ij|ij)|) over the shell pair's Cartesian components, without
``norm_cont`` (as the reference), is evaluated for every shell pair at once by
the device kernels' Schwarz mode.  The ``boys_func`` and
``compute_integral_func`` arguments of the reference are accepted for
signature compatibility and not used: the bound is the Coulomb one, which is
what the reference's screener uses (gbasis/integrals/electron_repulsion.py:475).
"""

import numpy as np

from .engine import DeviceTables, get_engine
from .tables import BatchTables, MoleculeTables


def compute_schwarz_bounds(contractions, boys_func=None, compute_integral_func=None,
                           device="cuda"):
    """Q (n_shells, n_shells), symmetric, in the order of ``contractions``."""
    contractions = list(contractions)
    if not contractions:
        return np.zeros((0, 0))
    mol = MoleculeTables(contractions, ["cartesian"] * len(contractions))
    dev = DeviceTables(BatchTables([mol]), device)
    q = get_engine(device).schwarz(dev).cpu().numpy()[:, 0]
    return mol.schwarz_orig(q)


def compute_schwarz_bound_shell_pair(boys_func, cont_one, cont_two, compute_integral_func=None,
                                     device="cuda"):
    """sqrt(max |(ab|ab)|) of one shell pair."""
    return float(compute_schwarz_bounds([cont_one, cont_two], device=device)[0, 1])


def shell_pair_significant(cont_one, cont_two, threshold=1e-12):
    """False iff every primitive pair's Gaussian-product factor exp(-ab/(a+b) |AB|^2)
    is at most ``threshold`` (coincident centres are always significant)."""
    r2 = float(np.sum((np.asarray(cont_one.coord) - np.asarray(cont_two.coord)) ** 2))
    if r2 < 1e-10:
        return True
    a = np.asarray(cont_one.exps, dtype=np.float64)[:, None]
    b = np.asarray(cont_two.exps, dtype=np.float64)[None, :]
    return bool(np.any(np.exp(-a * b / (a + b) * r2) > threshold))


class SchwarzScreener:
    """Schwarz screening of shell quartets: |(ij|kl)| <= Q_ij Q_kl, with counters.

    ``bounds`` (keyword, an extension) supplies a precomputed Q instead of the
    device evaluation."""

    def __init__(self, contractions, boys_func=None, compute_integral_func=None, threshold=1e-12,
                 bounds=None):
        self.threshold = threshold
        self.n_screened = 0
        self.n_computed = 0
        self.bounds = (np.asarray(bounds, dtype=np.float64) if bounds is not None
                       else compute_schwarz_bounds(contractions))

    def is_significant(self, i, j, k, l_shell):
        if self.bounds[i, j] * self.bounds[k, l_shell] < self.threshold:
            self.n_screened += 1
            return False
        self.n_computed += 1
        return True

    def get_statistics(self):
        total = self.n_screened + self.n_computed
        return {"n_screened": self.n_screened, "n_computed": self.n_computed, "total": total,
                "percent_screened": 100.0 * self.n_screened / total if total else 0.0,
                "speedup_factor": total / max(self.n_computed, 1)}

    def reset_counters(self):
        self.n_screened = 0
        self.n_computed = 0
