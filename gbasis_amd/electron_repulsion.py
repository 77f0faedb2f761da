"""Drop-in electron-repulsion-integral API backed by the HIP kernels.

Same names, arguments, notation handling and error behaviour as
gbasis/integrals/electron_repulsion.py:

* ``electron_repulsion_integral(basis, transform=None, notation="physicist")``
  (:214-281) -- no screening;
* ``electron_repulsion_integral_improved(basis, transform=None,
  notation="physicist", schwarz_threshold=0.0)`` (:415-504) -- Schwarz
  screening of shell quartets with Q_ij Q_kl < threshold, Q built from the
  un-normalised (ij|ij) blocks (gbasis/integrals/_schwarz_screening.py);
* ``ElectronRepulsionIntegral`` / ``ElectronRepulsionIntegralImproved`` with
  ``construct_array_cartesian / _spherical / _mix / _lincomb``
  (gbasis/base_four_symm.py:160-616).

Both engines of the reference compute the same integrals (their test suite
holds them equal to rtol 1e-10); here both run the same device kernels and
differ only in screening.  Additive API: ``electron_repulsion_integral_packed``
returns the canonical 8-fold packed array on the device, optionally with the
FP32-bulk mixed precision.
"""

import numpy as np

from .base_four_symm import BaseFourIndexSymmetric
from .boys_functions import get_boys_function
from .contractions import GeneralizedContractionShell
from .point_charge import PointChargeIntegral
from .engine import DeviceTables, get_engine
from .tables import BatchTables, MoleculeTables

_NOTATIONS = ("physicist", "chemist")
_REQUIRED = ("coord", "angmom", "exps", "coeffs", "norm_cont", "angmom_components_cart",
             "angmom_components_sph")


def _check_basis(basis):
    for shell in basis:
        if not isinstance(shell, GeneralizedContractionShell) and not all(
                hasattr(shell, a) for a in _REQUIRED):
            raise TypeError("`basis` must hold `GeneralizedContractionShell` instances.")


def _coord_types(basis, coord_type=None):
    if coord_type is None:
        return [s.coord_type for s in basis]
    if isinstance(coord_type, str):
        if coord_type not in ("cartesian", "spherical", "c", "p"):
            raise TypeError("`coord_type` must be a list/tuple of the strings 'cartesian' or "
                            "'spherical'")
        return [coord_type] * len(basis)
    if not isinstance(coord_type, (list, tuple)):
        raise TypeError("`coord_type` must be a list/tuple of the strings 'cartesian' or "
                        "'spherical'")
    if len(coord_type) == 1:
        return list(coord_type) * len(basis)
    if len(coord_type) != len(basis):
        raise ValueError("`coord_types` must have the same length as the given contractions.")
    for ct in coord_type:
        if ct not in ("cartesian", "spherical", "c", "p"):
            raise ValueError("Each entry of `coord_types` must be one of 'cartesian' or "
                             "'spherical'.")
    return list(coord_type)


def electron_repulsion_integral_packed(basis, coord_types=None, schwarz_threshold=0.0,
                                       precision="fp64", fp32_cutoff=1e-2, device="cuda",
                                       potential="coulomb", omega=None):
    """Canonical 8-fold packed ERIs (float64 device tensor) of one basis.

    ``precision="mixed"`` runs shell quartets whose contracted [e0|f0] bound
    P_ij P_kl is below ``fp32_cutoff`` with an FP32 recursion bulk (Boys,
    prefactors, accumulation and the horizontal recursion stay FP64).
    ``potential`` "erf" / "erfc" with ``omega`` evaluates the range-separated
    operators erf(omega r)/r and erfc(omega r)/r, the Boys functions of
    gbasis/integrals/boys_functions.py:145-223 (FP64 only; Schwarz screening
    keeps the Coulomb bounds, as the reference's screener,
    gbasis/integrals/electron_repulsion.py:475).
    Returns ``(packed, nfunc)``.
    """
    if precision not in ("fp64", "mixed"):
        raise ValueError("`precision` must be one of 'fp64' or 'mixed'")
    if potential not in ("coulomb", "erf", "erfc"):
        raise ValueError("`potential` must be one of 'coulomb', 'erf' or 'erfc'")
    if potential != "coulomb":
        if omega is None:
            raise ValueError(f"omega parameter is required for {potential}-attenuated potential")
        if precision != "fp64":
            raise ValueError("range-separated potentials are evaluated in FP64 only")
    _check_basis(basis)
    basis = list(basis)
    ctypes_ = _coord_types(basis, coord_types)
    eng = get_engine(device)
    mol = MoleculeTables(basis, ctypes_)
    dev = DeviceTables(BatchTables([mol]), device, potential=potential, omega=omega)
    need_q = schwarz_threshold > 0 or precision == "mixed"
    bound = eng.pair_bounds(dev, eng.schwarz(dev)) if need_q else None
    plans = eng.plan_device(dev, bound, schwarz_threshold,
                            fp32_cutoff if precision == "mixed" else 0.0, stats=False)
    packed = eng.run(dev, plans)
    return packed, mol.nfunc


def _dense(basis, coord_types, notation, schwarz_threshold, transform, device="cuda",
           potential="coulomb", omega=None):
    if notation not in _NOTATIONS:
        raise ValueError("`notation` must be one of 'physicist' or 'chemist'")
    packed, nfunc = electron_repulsion_integral_packed(basis, coord_types, schwarz_threshold,
                                                       device=device, potential=potential,
                                                       omega=omega)
    eng = get_engine(device)
    if transform is None:
        dense = eng.unpack_dense(packed, nfunc, notation)
        out = dense.cpu().numpy()
        return out
    # sum_ijkl T_ai T_bj T_ck T_dl (ij|kl) (gbasis/base_four_symm.py:611-615) by the
    # library's two-pass transform of the packed array (no dense AO array)
    arr = eng.mo_transform(packed, nfunc, transform, notation)
    del packed
    return arr.cpu().numpy()


def electron_repulsion_integral(basis, transform=None, notation="physicist"):
    """Dense ERIs of ``basis`` (electron_repulsion.py:214-281)."""
    if notation not in _NOTATIONS:
        raise ValueError("`notation` must be one of 'physicist' or 'chemist'")
    _check_basis(basis)
    return _dense(basis, None, notation, 0.0, transform)


def electron_repulsion_integral_improved(basis, transform=None, notation="physicist",
                                         schwarz_threshold=0.0, potential="coulomb", omega=None):
    """Dense ERIs with Schwarz screening (electron_repulsion.py:415-504).  The keyword
    arguments ``potential`` ("coulomb", "erf", "erfc") and ``omega`` select the
    operator, as a ``boys_func = get_boys_function(potential, omega)`` override of
    the reference's ElectronRepulsionIntegralImproved does
    (boys_functions.py:290-345)."""
    if notation not in _NOTATIONS:
        raise ValueError("`notation` must be one of 'physicist' or 'chemist'")
    _check_basis(basis)
    return _dense(basis, None, notation, schwarz_threshold, transform, potential=potential,
                  omega=omega)


class ElectronRepulsionIntegral(BaseFourIndexSymmetric):
    """Class interface of gbasis/integrals/electron_repulsion.py:20-211 (chemists' notation).

    Every quartet is evaluated on the device.  A subclass that supplies its own
    ``construct_array_contraction`` gets the reference's per-quartet assembly
    (``BaseFourIndexSymmetric``) around it instead."""

    schwarz_threshold = 0.0
    potential = "coulomb"   # or "erf" / "erfc" with omega (ElectronRepulsionIntegralImproved)
    omega = None
    # the reference's Boys hook (electron_repulsion.py:84,303): a callable from
    # gbasis_amd.boys_functions; its ``operator`` tag selects what the device evaluates
    boys_func = PointChargeIntegral.boys_func

    @classmethod
    def _operator(cls):
        """(potential, omega): a non-Coulomb ``boys_func`` wins over the ``potential``
        and ``omega`` attributes.  Any other callable cannot run inside the device
        kernels and is rejected."""
        tag = getattr(cls.boys_func, "operator", None)
        if tag is None:
            raise NotImplementedError(
                "`boys_func` must come from gbasis_amd.boys_functions.get_boys_function: the "
                "device kernels evaluate the Boys function themselves")
        if tag[0] != "coulomb":
            return tag
        return cls.potential, cls.omega

    def __init__(self, contractions):
        contractions = tuple(contractions)
        _check_basis(contractions)
        self._axes_contractions = (contractions,)

    def _plugin(self):
        """True when a subclass replaces the device ``construct_array_contraction``."""
        for klass in type(self).__mro__:
            if "construct_array_contraction" in klass.__dict__:
                return klass.__dict__["construct_array_contraction"] not in _DEVICE_CONTRACTION
        return False

    def _build(self, coord_types, **kwargs):
        if self._plugin():
            return self._assemble([ct in ("spherical", "p") for ct in coord_types], kwargs)
        thr = kwargs.get("schwarz_threshold", self.schwarz_threshold)
        pot, om = self._operator()
        return _dense(self.contractions, coord_types, "chemist", thr, None, potential=pot,
                      omega=om)

    def construct_array_cartesian(self, **kwargs):
        return self._build(["cartesian"] * len(self.contractions), **kwargs)

    def construct_array_spherical(self, **kwargs):
        return self._build(["spherical"] * len(self.contractions), **kwargs)

    def construct_array_mix(self, coord_types, **kwargs):
        if not isinstance(coord_types, (list, tuple)):
            raise TypeError("`coord_types` must be a list or a tuple.")
        return self._build(_coord_types(self.contractions, coord_types), **kwargs)

    def construct_array_lincomb(self, transform, coord_type, **kwargs):
        if self._plugin():
            return BaseFourIndexSymmetric.construct_array_lincomb(self, transform, coord_type,
                                                                  **kwargs)
        thr = kwargs.get("schwarz_threshold", self.schwarz_threshold)
        pot, om = self._operator()
        return _dense(self.contractions, _coord_types(self.contractions, coord_type), "chemist",
                      thr, transform, potential=pot, omega=om)

    @classmethod
    def construct_array_contraction(cls, cont_one, cont_two, cont_three, cont_four, **kwargs):
        """Cartesian block (M_1, L_1, M_2, L_2, M_3, L_3, M_4, L_4) of one shell quartet in
        chemists' notation, in each shell's Cartesian component order, without
        ``norm_cont`` (electron_repulsion.py:87-211, :306-412).

        The quartet is evaluated on the device as the four shells' dense array, from
        which the block is cut.  As in the reference, this engine takes no screening
        keywords (``ElectronRepulsionIntegralImproved`` does).
        """
        conts = (cont_one, cont_two, cont_three, cont_four)
        names = ("cont_one", "cont_two", "cont_three", "cont_four")
        for name, c in zip(names, conts):
            if not isinstance(c, GeneralizedContractionShell):
                raise TypeError(f"`{name}` must be a `GeneralizedContractionShell` instance.")
        pot, om = cls._operator()
        return _quartet_block([_UnitNormShell(c) for c in conts], pot, om)


def _quartet_block(shells, potential="coulomb", omega=None):
    """(M1, L1, M2, L2, M3, L3, M4, L4) chemists' block of four shells whose
    ``norm_cont`` is 1, cut from their dense array evaluated on the device."""
    dense = _dense(shells, ["cartesian"] * 4, "chemist", 0.0, None, potential=potential,
                   omega=omega)
    sizes = [s.coeffs.shape[1] * len(s.angmom_components_cart) for s in shells]
    off = np.concatenate([[0], np.cumsum(sizes)])
    blk = dense[off[0]:off[1], off[1]:off[2], off[2]:off[3], off[3]:off[4]]
    shape = []
    for s in shells:
        shape += [s.coeffs.shape[1], len(s.angmom_components_cart)]
    return np.ascontiguousarray(blk).reshape(shape)


class _UnitNormShell:
    """A shell seen with ``norm_cont`` = 1 (the contraction block of the reference
    excludes it)."""

    def __init__(self, shell):
        self._s = shell
        self.norm_cont = np.ones((shell.coeffs.shape[1], len(shell.angmom_components_cart)))

    def __getattr__(self, name):
        return getattr(self._s, name)


class ElectronRepulsionIntegralImproved(ElectronRepulsionIntegral):
    """Class interface of gbasis/integrals/electron_repulsion.py:284-412."""

    boys_func = get_boys_function("coulomb")

    @classmethod
    def construct_array_contraction(cls, cont_one, cont_two, cont_three, cont_four, **kwargs):
        """Cartesian block of one shell quartet, as the base class, with the Schwarz test of
        electron_repulsion.py:341-361: given both ``screener`` (a SchwarzScreener) and
        ``contraction_index_map`` ({id(shell): shell index}), a quartet whose four shells
        are all in the map and that ``screener.is_significant(i, j, k, l)`` rejects is
        returned as a zero block without any evaluation; ``is_significant`` updates the
        screener's counters as in the reference."""
        conts = (cont_one, cont_two, cont_three, cont_four)
        names = ("cont_one", "cont_two", "cont_three", "cont_four")
        for name, c in zip(names, conts):
            if not isinstance(c, GeneralizedContractionShell):
                raise TypeError(f"`{name}` must be a `GeneralizedContractionShell` instance.")
        screener = kwargs.get("screener")
        index_map = kwargs.get("contraction_index_map")
        if screener is not None and index_map is not None:
            ids = [index_map.get(id(c)) for c in conts]
            if all(x is not None for x in ids) and not screener.is_significant(*ids):
                shape = []
                for c in conts:
                    shape += [c.coeffs.shape[1], len(c.angmom_components_cart)]
                return np.zeros(shape)
        return ElectronRepulsionIntegral.construct_array_contraction.__func__(
            cls, cont_one, cont_two, cont_three, cont_four)


# the device implementations of the per-quartet hook (a subclass replacing them runs the
# host per-quartet assembly instead)
_DEVICE_CONTRACTION = (ElectronRepulsionIntegral.__dict__["construct_array_contraction"],
                       ElectronRepulsionIntegralImproved.__dict__["construct_array_contraction"])
