"""Small numeric helpers shared by the host-side table builders.

Mirrors ``gbasis/utils.py:7-22`` (``factorial2``): the double factorial with
``n!! = 1`` for every ``n <= 0`` (so ``(-1)!! = 1``, the convention every
Cartesian normalisation in the reference relies on).
"""

import numpy as np


def factorial2(n):
    """Double factorial n!! of an integer or integer array; 1 for n <= 0.

    Computed exactly with Python integers and returned as float64 (the values
    needed here, up to (2*8-1)!!, are far below 2**53).
    """
    arr = np.asarray(n)
    flat = arr.reshape(-1)
    out = np.empty(flat.shape, dtype=np.float64)
    for idx, v in enumerate(flat.tolist()):
        v = int(v)
        acc = 1
        while v > 1:
            acc *= v
            v -= 2
        out[idx] = float(acc)
    if arr.ndim == 0:
        return float(out[0])
    return out.reshape(arr.shape)


def ncart(l):
    """Number of Cartesian components (l+1)(l+2)/2 of angular momentum l."""
    return (l + 1) * (l + 2) // 2


def nsph(l):
    """Number of pure (spherical) components 2l+1 of angular momentum l."""
    return 2 * l + 1


def default_cart_components(l):
    """Cartesian exponents in the reference's default order.

    ``gbasis/contractions.py:379-385``: x descending, then y descending, z fills
    the rest, e.g. l=2 -> xx, xy, xz, yy, yz, zz.
    """
    return np.array([(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)],
                    dtype=np.int64).reshape(-1, 3)


def default_sph_components(l):
    """Pure-function labels in the reference's default order.

    ``gbasis/contractions.py:426-432``: ``s_l .. s_1, c_0 .. c_l``, except l=1
    which is ``c1, s1, c0`` (that is x, y, z).
    """
    if l == 1:
        return ("c1", "s1", "c0")
    return tuple([f"s{m}" for m in range(l, 0, -1)] + [f"c{m}" for m in range(l + 1)])
