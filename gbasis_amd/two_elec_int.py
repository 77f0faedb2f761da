"""Function-level interface of the two ERI engines, evaluated on the device:
compute_two_electron_integrals_os_hgp (gbasis/integrals/_two_elec_int_improved.py:591-732),
_compute_two_elec_integrals (gbasis/integrals/_two_elec_int.py:149-814) and
_compute_two_elec_integrals_angmom_zero (:8-145).

Each takes one quartet's raw shell data (centre, angular momentum, component
order, exponents, coefficients (K, M)) and returns the contracted chemists'
integrals (L_a, L_b, L_c, L_d, M_a, M_b, M_c, M_d) with the primitive and
component normalisation and no contraction normalisation, as the reference
engines do.  Any angular-momentum order is accepted (the reference callers
order the shells first; the device path needs no ordering).

``boys_func`` selects the operator: a callable from
``gbasis_amd.boys_functions`` (``boys_function_standard`` or one made by
``get_boys_function``) carries the operator it stands for; other callables
cannot run inside the device kernels and raise NotImplementedError.
``primitive_threshold`` > 0 (primitive-quartet screening, which the
reference's public API never sets) is not supported and raises.
"""

import numpy as np

from .utils import default_cart_components, default_sph_components
from .electron_repulsion import _quartet_block


class _QuartetShell:
    """One centre of a quartet, with ``norm_cont`` = 1."""

    def __init__(self, coord, angmom, components, exps, coeffs):
        self.coord = np.asarray(coord, dtype=np.float64).reshape(3)
        self.angmom = int(angmom)
        self.angmom_components_cart = np.asarray(components).reshape(-1, 3)
        if not np.all(self.angmom_components_cart.sum(axis=1) == self.angmom):
            raise ValueError("angular momentum components must sum to the angular momentum")
        self.angmom_components_sph = default_sph_components(self.angmom)
        self.exps = np.asarray(exps, dtype=np.float64).reshape(-1)
        c = np.asarray(coeffs, dtype=np.float64)
        self.coeffs = c[:, None] if c.ndim == 1 else c
        if self.coeffs.shape[0] != self.exps.shape[0]:
            raise ValueError("coefficients must have one row per exponent")
        self.norm_cont = np.ones((self.coeffs.shape[1], len(self.angmom_components_cart)))
        self.coord_type = "cartesian"


def _operator(boys_func):
    tag = getattr(boys_func, "operator", None)
    if tag is None:
        raise NotImplementedError(
            "`boys_func` must be gbasis_amd.boys_functions.boys_function_standard or come from "
            "get_boys_function: the device kernels evaluate the Boys function themselves")
    return tag


def compute_two_electron_integrals_os_hgp(boys_func,
                                          coord_a, angmom_a, angmom_components_a, exps_a, coeffs_a,
                                          coord_b, angmom_b, angmom_components_b, exps_b, coeffs_b,
                                          coord_c, angmom_c, angmom_components_c, exps_c, coeffs_c,
                                          coord_d, angmom_d, angmom_components_d, exps_d, coeffs_d,
                                          primitive_threshold=0.0):
    """(L_a, L_b, L_c, L_d, M_a, M_b, M_c, M_d) chemists' integrals (ab|cd)."""
    if primitive_threshold > 0:
        raise NotImplementedError("primitive-quartet screening (primitive_threshold > 0) is not "
                                  "supported; the device kernels keep every primitive quartet")
    potential, omega = _operator(boys_func)
    shells = [_QuartetShell(*a) for a in (
        (coord_a, angmom_a, angmom_components_a, exps_a, coeffs_a),
        (coord_b, angmom_b, angmom_components_b, exps_b, coeffs_b),
        (coord_c, angmom_c, angmom_components_c, exps_c, coeffs_c),
        (coord_d, angmom_d, angmom_components_d, exps_d, coeffs_d))]
    blk = _quartet_block(shells, potential, omega)
    return np.ascontiguousarray(np.transpose(blk, (1, 3, 5, 7, 0, 2, 4, 6)))


def _compute_two_elec_integrals(boys_func,
                                coord_a, angmom_a, angmom_components_a, exps_a, coeffs_a,
                                coord_b, angmom_b, angmom_components_b, exps_b, coeffs_b,
                                coord_c, angmom_c, angmom_components_c, exps_c, coeffs_c,
                                coord_d, angmom_d, angmom_components_d, exps_d, coeffs_d):
    """The first engine's entry point (same integrals, same layout).  Four s shells
    raise ValueError, as the reference's does (tests/test_two_elec_int.py:1466-1491):
    they belong to ``_compute_two_elec_integrals_angmom_zero``."""
    if angmom_a == angmom_b == angmom_c == angmom_d == 0:
        raise ValueError("four s shells: use _compute_two_elec_integrals_angmom_zero")
    return compute_two_electron_integrals_os_hgp(
        boys_func, coord_a, angmom_a, angmom_components_a, exps_a, coeffs_a,
        coord_b, angmom_b, angmom_components_b, exps_b, coeffs_b,
        coord_c, angmom_c, angmom_components_c, exps_c, coeffs_c,
        coord_d, angmom_d, angmom_components_d, exps_d, coeffs_d)


def _compute_two_elec_integrals_angmom_zero(boys_func, coord_a, exps_a, coeffs_a, coord_b, exps_b,
                                            coeffs_b, coord_c, exps_c, coeffs_c, coord_d, exps_d,
                                            coeffs_d):
    """(1, 1, 1, 1, M_a, M_b, M_c, M_d) for four s shells."""
    s = default_cart_components(0)
    return compute_two_electron_integrals_os_hgp(
        boys_func, coord_a, 0, s, exps_a, coeffs_a, coord_b, 0, s, exps_b, coeffs_b,
        coord_c, 0, s, exps_c, coeffs_c, coord_d, 0, s, exps_d, coeffs_d)
