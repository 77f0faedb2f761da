"""Boys functions evaluated on the device: the interface of
gbasis/integrals/boys_functions.py (boys_function_standard :26-63,
boys_function_all_orders :66-99, boys_function_recursion :102-142,
boys_function_erf :145-182, boys_function_erfc :185-223,
get_boys_function :290-345).

Values come from the device kernel the ERI kernels use (tabulated Taylor
expansion, downward recursion, asymptotic branch beyond the table), within
2.2e-15 relative of mpmath for orders up to 16 and T up to 1e5.  Orders up to
``get_engine().boys_max_order`` (24) are served; higher orders raise
ValueError (there is no host fallback).  The reference's mpmath evaluation
(boys_function_mpmath :226-280) is a validation tool, not a device path; the
tests use the oracle's mpmath Boys function in its place.
"""

import numpy as np

from .engine import get_engine


def boys_function_standard(orders, weighted_dist, device="cuda"):
    """F_m(T) = 1F1(m + 1/2; m + 3/2; -T) / (2m + 1), broadcast over ``orders`` and
    ``weighted_dist``."""
    orders = np.asarray(orders)
    T = np.asarray(weighted_dist, dtype=np.float64)
    o = np.broadcast_to(orders, np.broadcast_shapes(orders.shape, T.shape))
    if o.size == 0:
        return np.zeros(o.shape)
    if np.any(o < 0) or np.any(o != np.floor(o)):
        raise ValueError("Boys orders must be non-negative integers")
    if np.any(T < 0):
        raise ValueError("Boys arguments must be non-negative")
    eng = get_engine(device)
    mmax = int(o.max())
    if mmax > eng.boys_max_order:
        raise ValueError(f"Boys order {mmax} exceeds the device table (at most "
                         f"{eng.boys_max_order})")
    # each element of weighted_dist once, with all orders; then pick (argument, order)
    tf = np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
    F = eng.boys(None, tf, mmax).cpu().numpy()
    which = np.broadcast_to(np.arange(tf.size).reshape(T.shape), o.shape)
    return F[which, o.astype(np.int64)]


boys_function_standard.operator = ("coulomb", None)


def boys_function_all_orders(m_max, weighted_dist, device="cuda"):
    """F_m(T) for m = 0..m_max, shape (m_max + 1,) + T.shape (boys_functions.py:66-99).
    One device pass gives every order (downward recursion from the top order)."""
    T = np.asarray(weighted_dist, dtype=np.float64)
    m_max = int(m_max)
    if m_max < 0:
        raise ValueError("m_max must be non-negative")
    if np.any(T < 0):
        raise ValueError("Boys arguments must be non-negative")
    if T.size == 0:
        return np.zeros((m_max + 1,) + T.shape)
    eng = get_engine(device)
    if m_max > eng.boys_max_order:
        raise ValueError(f"Boys order {m_max} exceeds the device table (at most "
                         f"{eng.boys_max_order})")
    F = eng.boys(None, np.ascontiguousarray(T).reshape(-1), m_max).cpu().numpy()
    return np.ascontiguousarray(F.T).reshape((m_max + 1,) + T.shape)


# the reference's downward-recursion variant (boys_functions.py:102-142) computes the
# same values; the device evaluation recurs downwards already
boys_function_recursion = boys_function_all_orders


def boys_function_erf(orders, weighted_dist, rho, omega, device="cuda"):
    """s^(m+1/2) F_m(s T), s = omega^2 / (omega^2 + rho): erf(omega r)/r."""
    orders = np.asarray(orders)
    s = omega ** 2 / (omega ** 2 + np.asarray(rho, dtype=np.float64))
    return s ** (orders + 0.5) * boys_function_standard(orders, s * np.asarray(weighted_dist),
                                                        device)


def boys_function_erfc(orders, weighted_dist, rho, omega, device="cuda"):
    """F_m(T) - s^(m+1/2) F_m(s T): erfc(omega r)/r."""
    return (boys_function_standard(orders, weighted_dist, device)
            - boys_function_erf(orders, weighted_dist, rho, omega, device))


def get_boys_function(potential="coulomb", omega=None):
    """A callable boys_func(orders, weighted_dist, rho=None) for "coulomb" / "standard" /
    "1/r", "erf" / "erf_coulomb" or "erfc" / "erfc_coulomb" (case-insensitive), with the
    reference factory's errors: unknown potential, missing omega, missing rho.

    The callable carries ``operator`` = (potential, omega): assigned as the
    ``boys_func`` class attribute of the ERI classes (electron_repulsion.py:84,303
    of the reference) it selects the operator the device kernels evaluate."""
    potential = potential.lower()
    if potential in ("coulomb", "standard", "1/r"):
        def boys_func(orders, weighted_dist, rho=None):
            return boys_function_standard(orders, weighted_dist)
        boys_func.operator = ("coulomb", None)
        return boys_func
    kinds = {"erf": ("erf", boys_function_erf), "erf_coulomb": ("erf", boys_function_erf),
             "erfc": ("erfc", boys_function_erfc), "erfc_coulomb": ("erfc", boys_function_erfc)}
    if potential not in kinds:
        raise ValueError(f"Unknown potential type: {potential}. Valid options: coulomb, erf, erfc")
    name, fn = kinds[potential]
    if omega is None:
        raise ValueError(f"omega parameter is required for {name}-attenuated potential")

    def boys_func(orders, weighted_dist, rho=None):
        if rho is None:
            raise ValueError(f"rho parameter is required for {name}-attenuated potential")
        return fn(orders, weighted_dist, rho, omega)
    boys_func.operator = (name, float(omega))
    return boys_func
