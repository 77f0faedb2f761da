"""The per-quartet plugin point of gbasis/base_four_symm.py:13-616.

``BaseFourIndexSymmetric`` assembles a four-index array with 8-fold
permutational symmetry from a subclass's ``construct_array_contraction``
(cont1, cont2, cont3, cont4, **kwargs) -> (M1, L1, M2, L2, M3, L3, M4, L4, ...),
called once per unique shell quartet (i <= j, k <= l, pair(ij) <= pair(kl),
base_four_symm.py:189-191), scaled by each shell's ``norm_cont``, transformed
to pure functions per shell where asked (``generate_transformation``, "left"),
and scattered into the eight permutations of the quartet.

The scatter goes straight into one preallocated array, in the order of
base_four_symm.py:216-226, so that where two permutations name the same block
(diagonal quartets) the later one is kept, as in the reference: a
non-symmetric plugin block reads back as in the reference
(tests/test_base_four_symm.py:681-730 expects ijkl->lkji for one shell).

The ERI classes of ``gbasis_amd.electron_repulsion`` derive from this class
and evaluate every quartet on the device; this host loop runs only for a
subclass that supplies its own ``construct_array_contraction``.
"""

import abc
import itertools

import numpy as np

from .spherical import generate_transformation

_REQUIRED = ("coord", "angmom", "exps", "coeffs", "norm_cont", "angmom_components_cart")

# destination (as indices into (i, j, k, l)) and the block axes that land on
# its four axes, in the reference's assignment order
_SCATTER = (((0, 1, 2, 3), (0, 1, 2, 3)),
            ((0, 1, 3, 2), (0, 1, 3, 2)),
            ((1, 0, 2, 3), (1, 0, 2, 3)),
            ((1, 0, 3, 2), (1, 0, 3, 2)),
            ((2, 3, 0, 1), (2, 3, 0, 1)),
            ((3, 2, 0, 1), (3, 2, 0, 1)),
            ((2, 3, 1, 0), (2, 3, 1, 0)),
            ((3, 2, 1, 0), (3, 2, 1, 0)))


def _is_shell(obj):
    from .contractions import GeneralizedContractionShell
    return isinstance(obj, GeneralizedContractionShell) or all(hasattr(obj, a) for a in _REQUIRED)


class BaseFourIndexSymmetric(abc.ABC):
    """Four-index array over one set of contractions (base_four_symm.py:13-92)."""

    def __init__(self, contractions):
        if not isinstance(contractions, (list, tuple)) or not all(
                _is_shell(c) for c in contractions):
            raise TypeError("Contractions must be given as a list or tuple of "
                            "`GeneralizedContractionShell` instances.")
        if len(contractions) == 0:
            raise ValueError("At least one `GeneralizedContractionShell` must be given.")
        self._axes_contractions = (tuple(contractions),)

    @property
    def contractions(self):
        return self._axes_contractions[0]

    @abc.abstractmethod
    def construct_array_contraction(self, cont1, cont2, cont3, cont4, **kwargs):
        """(M1, L1, M2, L2, M3, L3, M4, L4, ...) block of one shell quartet, in each
        shell's Cartesian component order, without ``norm_cont``."""

    # -- assembly -------------------------------------------------------------------
    def _assemble(self, spherical, kwargs):
        """Dense array; ``spherical[s]`` says whether shell s is transformed to pure
        functions."""
        shells = self.contractions
        transforms = [generate_transformation(c.angmom, c.angmom_components_cart,
                                              c.angmom_components_sph, "left") if sph else None
                      for c, sph in zip(shells, spherical)]
        sizes = [c.coeffs.shape[1] * (t.shape[0] if t is not None
                                      else len(c.angmom_components_cart))
                 for c, t in zip(shells, transforms)]
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        out = None
        pairs = list(itertools.combinations_with_replacement(range(len(shells)), 2))
        for p, (i, j) in enumerate(pairs):
            for k, l in pairs[p:]:
                quartet = (i, j, k, l)
                block = self.construct_array_contraction(*(shells[s] for s in quartet), **kwargs)
                block = np.asarray(block)
                extra = block.shape[8:]
                for a, s in enumerate(quartet):
                    shape = [1] * block.ndim
                    shape[2 * a:2 * a + 2] = block.shape[2 * a:2 * a + 2]
                    block = block * np.asarray(shells[s].norm_cont).reshape(shape)
                for a, s in enumerate(quartet):
                    if transforms[s] is not None:
                        block = np.moveaxis(np.tensordot(transforms[s], block, (1, 2 * a + 1)),
                                            0, 2 * a + 1)
                block = block.reshape(tuple(block.shape[2 * a] * block.shape[2 * a + 1]
                                            for a in range(4)) + extra)
                if out is None:
                    out = np.zeros((int(off[-1]),) * 4 + extra, dtype=block.dtype)
                tail = tuple(range(4, block.ndim))
                for dest, axes in _SCATTER:
                    idx = tuple(slice(off[quartet[d]], off[quartet[d] + 1]) for d in dest)
                    out[idx] = block.transpose(axes + tail)
        return out

    def construct_array_cartesian(self, **kwargs):
        """(K_cart,)*4 array (base_four_symm.py:160-244)."""
        return self._assemble([False] * len(self.contractions), kwargs)

    def construct_array_spherical(self, **kwargs):
        """(K_sph,)*4 array (base_four_symm.py:246-389)."""
        return self._assemble([True] * len(self.contractions), kwargs)

    def construct_array_mix(self, coord_types, **kwargs):
        """(K_cont,)*4 array with a coordinate type per shell (base_four_symm.py:391-557)."""
        if not isinstance(coord_types, (list, tuple)):
            raise TypeError("`coord_types` must be a list or a tuple.")
        if not all(ct in ("cartesian", "spherical") for ct in coord_types):
            raise ValueError("Each entry of `coord_types` must be one of 'cartesian' or "
                             "'spherical'.")
        if len(coord_types) != len(self.contractions):
            raise ValueError("`coord_types` must have the same number of entries as the number "
                             "of `GeneralizedContractionShell` in the instance.")
        return self._assemble([ct == "spherical" for ct in coord_types], kwargs)

    def construct_array_lincomb(self, transform, coord_type, **kwargs):
        """sum T_ai T_bj T_ck T_dl X_ijkl over the Cartesian, pure or mixed array X
        (base_four_symm.py:559-616); ``coord_type`` is a list or tuple, as there."""
        if all(ct == "cartesian" for ct in coord_type):
            array = self.construct_array_cartesian(**kwargs)
        elif all(ct == "spherical" for ct in coord_type):
            array = self.construct_array_spherical(**kwargs)
        elif isinstance(coord_type, (list, tuple)):
            array = self.construct_array_mix(coord_type, **kwargs)
        else:
            raise TypeError("`coord_type` must be a list/tuple of the strings 'cartesian' or "
                            "'spherical'")
        for ax in range(4):
            array = np.moveaxis(np.tensordot(transform, array, (1, ax)), 0, ax)
        return array
