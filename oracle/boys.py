"""Boys function F_m(T) for the oracle.  TEST INFRASTRUCTURE ONLY.

``boys_hyp1f1`` is the reference's own definition
(gbasis/integrals/boys_functions.py:26-63, ``boys_function_standard``):

    F_m(T) = 1F1(m + 1/2; m + 3/2; -T) / (2m + 1)

evaluated with ``scipy.special.hyp1f1`` (scipy is an unpinned dependency of the
reference, ``pyproject.toml:37``; this container has 1.15.3).

``boys_mpmath`` is the same function at 50 significant digits and is what the
tests hold the other two against.
"""

import numpy as np
from scipy.special import hyp1f1


def boys_hyp1f1(orders, T):
    """F_m(T) broadcast over ``orders`` and ``T`` (float64)."""
    orders = np.asarray(orders, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    return hyp1f1(orders + 0.5, orders + 1.5, -T) / (2.0 * orders + 1.0)


def boys_mpmath(m, T, dps=50):
    """F_m(T) in mpmath at ``dps`` digits, returned as a Python float."""
    import mpmath

    with mpmath.workdps(dps):
        m = mpmath.mpf(m)
        T = mpmath.mpf(T)
        return float(mpmath.hyp1f1(m + 0.5, m + 1.5, -T) / (2 * m + 1))


def boys_longdouble(orders, T):
    """F_m(T) in x87 extended precision (``np.longdouble``), for tests.

    T below 60: the positive series F_m(T) = e^{-T} sum_i (2T)^i / ((2m+1)...(2m+2i+1));
    above: F_0 = sqrt(pi/T)/2 (erfc(sqrt 60) ~ 1e-28) and the upward recurrence
    F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T), stable for T > m.
    """
    orders, T = np.broadcast_arrays(np.asarray(orders), np.asarray(T, dtype=np.longdouble))
    m = orders.astype(np.longdouble)
    T = T.astype(np.longdouble)
    out = np.empty(T.shape, dtype=np.longdouble)
    pi = np.longdouble("3.14159265358979323846264338327950288")
    low = T < 60
    if low.any():
        t = T[low]
        mm = m[low]
        term = 1 / (2 * mm + 1)
        total = term.copy()
        for i in range(1, 2000):
            term = term * 2 * t / (2 * mm + 2 * i + 1)
            total += term
            if np.all(term < total * np.longdouble(1e-22)):
                break
        out[low] = total * np.exp(-t)
    high = ~low
    if high.any():
        t = T[high]
        mi = orders[high].astype(np.int64)
        f = np.sqrt(pi / t) / 2
        e = np.exp(-t)
        res = np.where(mi == 0, f, 0)
        for k in range(int(mi.max()) if mi.size else 0):
            f = ((2 * k + 1) * f - e) / (2 * t)
            res = np.where(mi == k + 1, f, res)
        out[high] = res
    return out
