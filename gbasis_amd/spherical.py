"""Cartesian -> pure (real solid harmonic) transformation matrices.

Same convention as ``gbasis/spherical.py:229-340`` (``generate_transformation``):
real regular solid harmonics in the Helgaker form (Molecular Electronic-Structure
Theory, eqs. 6.4.47-6.4.50)

    S_lm = N_lm sum_{t,u,v} C^{lm}_{tuv} x^{2t+|m|-2(u+v)} y^{2(u+v)} z^{l-2t-|m|}
    C^{lm}_{tuv} = (-1)^{t+v-v_m} 4^{-t} binom(l,t) binom(l-t,|m|+t) binom(t,u) binom(|m|,2v)
    N_lm = sqrt(2 (l+|m|)! (l-|m|)! / 2^{delta_m0}) / (2^{|m|} l!)

with ``v_m = 0`` for cosine-type (m >= 0) and ``1/2`` for sine-type (m < 0)
functions; the matrix column of each Cartesian component is then scaled by
``sqrt(prod (2a_i-1)!!) / sqrt((2l-1)!!)`` because the Cartesian functions it
multiplies are normalised per component.  Labels are ``c{m}``/``s{m}`` with an
optional leading ``-`` that flips the sign of that row.

The matrices are host-side constants: they are uploaded once per distinct
(l, Cartesian order, pure order) and applied on the device.
"""

from functools import lru_cache
from math import comb, factorial, sqrt

import numpy as np

from .utils import factorial2


def _solid_harmonic_terms(l, m):
    """Monomials {(ax, ay, az): coefficient} of the real solid harmonic S_lm."""
    am = abs(m)
    vm2 = 1 if m < 0 else 0  # 2 * v_m
    norm = sqrt(2.0 * factorial(l + am) * factorial(l - am) / (2.0 if m == 0 else 1.0)) / (
        2.0 ** am * factorial(l))
    terms = {}
    for t in range((l - am) // 2 + 1):
        for u in range(t + 1):
            # v runs over v_m, v_m + 1, ... with 2v <= |m|; work with w = 2v.
            for w in range(vm2, am + 1, 2):
                sign_exp = t + (w - vm2) // 2
                coeff = ((-1.0) ** sign_exp) * (0.25 ** t) * comb(l, t) * comb(l - t, am + t) \
                    * comb(t, u) * comb(am, w)
                if coeff == 0.0:
                    continue
                ax = 2 * t + am - 2 * u - w
                ay = 2 * u + w
                az = l - 2 * t - am
                key = (ax, ay, az)
                terms[key] = terms.get(key, 0.0) + coeff * norm
    return terms


def _parse_label(label, l):
    sign = 1.0
    if label.startswith("-"):
        sign = -1.0
        label = label[1:]
    if len(label) < 2 or label[0] not in "cs" or not label[1:].isdigit():
        raise ValueError(f"bad pure-function label {label!r}")
    m = int(label[1:])
    if m > l or (label[0] == "s" and m == 0):
        raise ValueError(f"bad pure-function label {label!r} for l={l}")
    return sign, (m if label[0] == "c" else -m)


@lru_cache(maxsize=None)
def _transformation_cached(l, cart_key, sph_key):
    cart = np.array(cart_key, dtype=np.int64).reshape(-1, 3)
    position = {tuple(c): i for i, c in enumerate(cart.tolist())}
    mat = np.zeros((2 * l + 1, cart.shape[0]))
    for row, label in enumerate(sph_key):
        sign, m = _parse_label(label, l)
        for comp, coeff in _solid_harmonic_terms(l, m).items():
            mat[row, position[comp]] = sign * coeff
    colscale = np.sqrt(np.prod(factorial2(2 * cart - 1), axis=1)) / sqrt(factorial2(2 * l - 1))
    mat *= colscale[np.newaxis, :]
    mat.setflags(write=False)
    return mat


def generate_transformation(angmom, cartesian_order, spherical_order, apply_from="left"):
    """Matrix taking Cartesian components (in ``cartesian_order``) to pure ones.

    Returns shape ``(2l+1, n_cart)`` for ``apply_from="left"`` and its transpose
    for ``"right"`` (``gbasis/spherical.py:229-340``).
    """
    if not isinstance(angmom, int) or angmom < 0:
        raise TypeError("Angular momentum must be a non-negative integer.")
    cart = np.asarray(cartesian_order)
    ncart = (angmom + 1) * (angmom + 2) // 2
    if cart.shape != (ncart, 3) or np.any(cart.sum(axis=1) != angmom):
        raise ValueError("Cartesian order does not match the angular momentum.")
    sph = tuple(spherical_order)
    if len(sph) != 2 * angmom + 1 or len({s.lstrip("-") for s in sph}) != 2 * angmom + 1:
        raise ValueError("Spherical order must hold 2l+1 distinct pure-function labels.")
    if apply_from not in ("left", "right"):
        raise ValueError("`apply_from` must be one of 'left' or 'right'.")
    mat = _transformation_cached(angmom, tuple(map(tuple, cart.tolist())), sph)
    return mat.copy() if apply_from == "left" else mat.T.copy()
