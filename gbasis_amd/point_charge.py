"""``PointChargeIntegral.boys_func`` of gbasis/integrals/point_charge.py:69-118, the
Boys function the first ERI class uses (electron_repulsion.py:84).  Only this
member is on the ERI path; the point-charge (one-electron) integrals themselves
are not part of this library.
"""

from .boys_functions import boys_function_standard


class PointChargeIntegral:
    """Holder of the Coulomb Boys function, evaluated on the device; assignable as the
    ``boys_func`` class attribute of the ERI classes (it carries the Coulomb
    operator tag)."""

    boys_func = staticmethod(boys_function_standard)
