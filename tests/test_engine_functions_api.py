"""Argument handling of the function-level engine interface
(gbasis/integrals/_two_elec_int_improved.py:591-732, _two_elec_int.py:8-814); the
values are checked on the GPU (test_gpu_parity.py::test_engine_functions_vs_oracle)."""
import numpy as np
import pytest

from gbasis_amd.boys_functions import boys_function_standard
from gbasis_amd.two_elec_int import compute_two_electron_integrals_os_hgp
from gbasis_amd.utils import default_cart_components


def _quartet(l=1):
    a = (np.zeros(3), l, default_cart_components(l), np.ones(2), np.ones((2, 1)))
    return [x for _ in range(4) for x in a]


def test_untagged_boys_function_is_rejected():
    with pytest.raises(NotImplementedError):
        compute_two_electron_integrals_os_hgp(lambda o, t, rho=None: t, *_quartet())


def test_primitive_screening_is_rejected():
    with pytest.raises(NotImplementedError):
        compute_two_electron_integrals_os_hgp(boys_function_standard, *_quartet(),
                                              primitive_threshold=1e-12)


def test_inconsistent_components_and_coefficients():
    q = _quartet()
    q[2] = default_cart_components(2)          # components of d on a p centre
    with pytest.raises(ValueError):
        compute_two_electron_integrals_os_hgp(boys_function_standard, *q)
    q = _quartet()
    q[4] = np.ones((3, 1))                      # three coefficients, two exponents
    with pytest.raises(ValueError):
        compute_two_electron_integrals_os_hgp(boys_function_standard, *q)


def test_first_engine_rejects_four_s_shells():
    """As the reference's test_compute_two_elec_integrals_prim: the general function of
    the first engine refuses four s shells (they go to the closed-form function)."""
    from gbasis_amd.two_elec_int import _compute_two_elec_integrals

    with pytest.raises(ValueError):
        _compute_two_elec_integrals(boys_function_standard, *_quartet(0))
