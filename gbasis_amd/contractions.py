"""Shell data model, duck-type compatible with the reference's shell class.

``GeneralizedContractionShell`` follows ``gbasis/contractions.py:8-572``: same
constructor ``(angmom, coord, coeffs, exps, coord_type, icenter=None)``, same
validation errors, same component orders and the same ``norm_cont`` meaning
(shape ``(M, L_cart)``: the factor that makes every contracted Cartesian
function unit-normalised).  The reference obtains ``norm_cont`` from its
overlap integral code (``contractions.py:503-524``); here it is computed in
closed form.  For component-normalised primitives on one centre the overlap of
two primitives of the same component does not depend on the component:

    S_pq = (2 sqrt(a_p a_q) / (a_p + a_q)) ** (l + 3/2)

so ``norm_cont[m, :] = (sum_pq c_pm c_qm S_pq) ** -1/2``.

The ERI host code never relies on this class: it reads the duck-typed
attributes (``coord``, ``angmom``, ``exps``, ``coeffs``, ``coord_type``,
``norm_cont``, ``angmom_components_cart``, ``angmom_components_sph``) of
whatever shell objects it is given, so subclasses that override orders or
norms (as HORTON/iodata/pyscf wrappers do in the reference) are honoured.
"""

from numbers import Integral

import numpy as np

from .utils import default_cart_components, default_sph_components, factorial2


class GeneralizedContractionShell:
    """A shell of contracted Cartesian/pure Gaussians sharing one centre and l."""

    def __init__(self, angmom, coord, coeffs, exps, coord_type, icenter=None):
        self.angmom = angmom
        self.coord = coord
        self.coeffs = coeffs
        self.exps = exps
        self.assign_norm_cont()
        self.coord_type = coord_type
        self.icenter = icenter

    # --- validated attributes (gbasis/contractions.py:152-358) -----------------
    @property
    def icenter(self):
        return self._icenter

    @icenter.setter
    def icenter(self, icenter):
        if isinstance(icenter, Integral):
            self._icenter = int(icenter)
        elif icenter is None:
            self._icenter = None
        else:
            raise TypeError(f"Center should be of `numbers.Integral` type. Got {type(icenter)}")

    @property
    def coord(self):
        return self._coord

    @coord.setter
    def coord(self, coord):
        if not (isinstance(coord, np.ndarray) and coord.size == 3):
            raise TypeError("Coordinate must be given as a `numpy` array with 3 elements.")
        if coord.dtype == int:      # integer coordinates are taken as floats
            coord = coord.astype(float)
        if coord.dtype != float:
            raise TypeError("The data type of the coordinate must be int or float.")
        self._coord = coord

    @property
    def angmom(self):
        return self._angmom

    @angmom.setter
    def angmom(self, angmom):
        if not isinstance(angmom, int):
            raise TypeError("Angular momentum must be given as an integer.")
        if angmom < 0:
            raise ValueError("Angular momentum must be a nonnegative integer.")
        self._angmom = angmom

    @property
    def exps(self):
        return self._exps

    @exps.setter
    def exps(self, exps):
        if not (isinstance(exps, np.ndarray) and exps.dtype == float):
            raise TypeError("Exponents must be given as a `numpy` array of floats.")
        if hasattr(self, "_coeffs") and self.coeffs.shape[0] != exps.shape[0]:
            raise ValueError("Exponents array must have the same number of primitives as the "
                             "coefficients array.")
        self._exps = exps

    @property
    def coeffs(self):
        return self._coeffs

    @coeffs.setter
    def coeffs(self, coeffs):
        if not (isinstance(coeffs, np.ndarray) and coeffs.dtype == float):
            raise TypeError("Contraction coefficients must be a `numpy` array of floats.")
        if coeffs.ndim not in (1, 2):
            raise ValueError("Contraction coefficients must be a one- or two-dimensional "
                             "`numpy` array.")
        if hasattr(self, "_exps") and coeffs.shape[0] != self.exps.shape[0]:
            raise ValueError("Coefficients array must have the same number of primitives as the "
                             "exponents array.")
        self._coeffs = coeffs[:, np.newaxis] if coeffs.ndim == 1 else coeffs

    @property
    def coord_type(self):
        return self._coord_type

    @coord_type.setter
    def coord_type(self, coord_type):
        if not isinstance(coord_type, str):
            raise TypeError("Coordinate type must be given as a string.")
        table = {"c": "cartesian", "cartesian": "cartesian", "p": "spherical",
                 "spherical": "spherical"}
        if coord_type not in table:
            raise ValueError("`coord_type` is incorrectly specified. It must be either 'c' or "
                             "'cartesian' for Cartesian coordinates, or 'p' or 'spherical' for "
                             "spherical coordinates.")
        self._coord_type = table[coord_type]

    # --- derived quantities --------------------------------------------------
    @property
    def angmom_components_cart(self):
        return default_cart_components(self.angmom)

    @property
    def angmom_components_sph(self):
        return default_sph_components(self.angmom)

    @property
    def num_cart(self):
        return (self.angmom + 1) * (self.angmom + 2) // 2

    @property
    def num_sph(self):
        return 1 + 2 * self.angmom

    @property
    def num_seg_cont(self):
        return self.coeffs.shape[1]

    @property
    def norm_prim_cart(self):
        """Primitive normalisation per (component, primitive), contractions.py:434-461."""
        e = self.exps[np.newaxis, :]
        comps = self.angmom_components_cart
        f2 = np.prod(factorial2(2 * comps - 1), axis=1)[:, np.newaxis]
        return (2 * e / np.pi) ** 0.75 * (4 * e) ** (self.angmom / 2) / np.sqrt(f2)

    def assign_norm_cont(self):
        """Closed-form contraction norms; same values as contractions.py:503-524."""
        a = self.exps
        s = (2.0 * np.sqrt(np.outer(a, a)) / np.add.outer(a, a)) ** (self.angmom + 1.5)
        self_ovl = np.einsum("pm,pq,qm->m", self.coeffs, s, self.coeffs)
        ncart = (self.angmom + 1) * (self.angmom + 2) // 2
        self.norm_cont = np.repeat((self_ovl ** -0.5)[:, np.newaxis], ncart, axis=1)
