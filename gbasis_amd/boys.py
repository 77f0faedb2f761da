"""Host-side Boys-function table for the device kernels.

The device evaluates F_m(T) (the reference's ``boys_function_standard``,
gbasis/integrals/boys_functions.py:26-63) from a table of F_n(k h) by an
8-term Taylor expansion plus the downward recurrence, whose e^{-T} is the
tabulated e^{-k h} times a short Taylor series of e^{k h - T}.  Row k of the
table is [e^{-k h}, F_0(k h), ..., F_{nord-2}(k h)].  It is computed here in
extended precision (x87 80-bit ``np.longdouble``) from the series

    F_n(T) = e^{-T} sum_{i>=0} (2T)^i / ((2n+1)(2n+3)...(2n+2i+1)),

whose terms are all positive, so the sum carries no cancellation.
"""

from functools import lru_cache

import numpy as np

BOYS_H = 0.1       # grid spacing
BOYS_TMAX = 40.0   # last grid point; beyond it the asymptotic branch is used
BOYS_NORD = 33     # e^{-kh}, then orders 0..31 -> total angular momentum up to 24 (8 Taylor terms)


@lru_cache(maxsize=None)
def boys_table(h=BOYS_H, tmax=BOYS_TMAX, nord=BOYS_NORD):
    """float64 array (nt, nord): column 0 e^{-k h}, column 1 + n F_n(k h)."""
    nt = int(round(tmax / h)) + 1
    T = np.arange(nt, dtype=np.longdouble) * np.longdouble(h)
    out = np.empty((nt, nord), dtype=np.longdouble)
    out[:, 0] = np.exp(-T)
    nterms = 400
    for n in range(nord - 1):
        term = np.ones(nt, dtype=np.longdouble) / np.longdouble(2 * n + 1)
        total = term.copy()
        for i in range(1, nterms):
            term = term * (2 * T) / np.longdouble(2 * n + 2 * i + 1)
            total += term
        out[:, n + 1] = total * np.exp(-T)
    tab = out.astype(np.float64)
    tab.setflags(write=False)
    return tab
