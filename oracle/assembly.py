"""Four-index assembly and Schwarz screening for the oracle.  TEST INFRASTRUCTURE ONLY.

Restates:
  * gbasis/integrals/electron_repulsion.py:305-412 — per shell quartet: Schwarz
    test on the original indices, canonical reorder (l_a>=l_b, l_c>=l_d,
    l_a>=l_c, applied in that order), engine call, un-swap, transpose to
    (M1,L1,M2,L2,M3,L3,M4,L4).
  * gbasis/base_four_symm.py:160-389 — unique quartet loop (i<=j, k<=l,
    pair(ij)<=pair(kl)), norm_cont scaling, per-shell Cartesian->pure
    transform, 8-fold placement into the dense tensor.
  * gbasis/integrals/electron_repulsion.py:214-281,415-504 — notation handling.
  * gbasis/integrals/_schwarz_screening.py:14-90,184-191 — Q_ij from the
    engine on (ij|ij) in the given orientation without norm_cont; a quartet is
    skipped iff Q_ij * Q_kl < threshold.

The Cartesian->pure matrices come from ``gbasis_amd.spherical`` (host-side
constants, pinned separately by the HORTON cartpure tables).
"""

import itertools as it

import numpy as np

from gbasis_amd.spherical import generate_transformation
from .boys import boys_longdouble
from .os_hgp import os_hgp_quartet


def _engine(s1, s2, s3, s4, exact=False, potential="coulomb", omega=None):
    if exact == "quad":
        if potential != "coulomb":
            raise ValueError("the binary128 engine evaluates the Coulomb operator only")
        from .hp import hp_engine

        return hp_engine(s1, s2, s3, s4, prec="quad", mode=1, tau=0.0)
    if exact:
        return os_hgp_quartet(
            s1.coord, s1.angmom, s1.angmom_components_cart, s1.exps, s1.coeffs,
            s2.coord, s2.angmom, s2.angmom_components_cart, s2.exps, s2.coeffs,
            s3.coord, s3.angmom, s3.angmom_components_cart, s3.exps, s3.coeffs,
            s4.coord, s4.angmom, s4.angmom_components_cart, s4.exps, s4.coeffs,
            boys=boys_longdouble, dtype=np.longdouble, ket="vrr", potential=potential,
            omega=omega).astype(np.float64)
    return os_hgp_quartet(s1.coord, s1.angmom, s1.angmom_components_cart, s1.exps, s1.coeffs,
                          s2.coord, s2.angmom, s2.angmom_components_cart, s2.exps, s2.coeffs,
                          s3.coord, s3.angmom, s3.angmom_components_cart, s3.exps, s3.coeffs,
                          s4.coord, s4.angmom, s4.angmom_components_cart, s4.exps, s4.coeffs,
                          potential=potential, omega=omega)


def schwarz_bounds(basis, exact=False):
    """Q (n_sh, n_sh): sqrt(max |(ij|ij)|), un-normalised, original orientation."""
    n = len(basis)
    Q = np.zeros((n, n))
    for i in range(n):
        for j in range(i, n):
            blk = _engine(basis[i], basis[j], basis[i], basis[j], exact)
            Q[i, j] = Q[j, i] = np.sqrt(np.max(np.abs(blk)))
    return Q


def contraction_block(s1, s2, s3, s4, exact=False, potential="coulomb", omega=None):
    """(M1,L1,M2,L2,M3,L3,M4,L4) Cartesian block with the reference's l-reordering.

    ``exact="quad"`` evaluates them in IEEE binary128 with every primitive pair built on
    the centre of its larger exponent (oracle/eri_hp.c), the accuracy reference for
    shells up to g with exponents up to 1e5, where the horizontal recursion of the
    reference's orientation amplifies x87 rounding beyond 1e-12.

    ``exact=True`` evaluates the integrals in x87 extended precision with an
    extended-precision Boys function and the bounded-coefficient Obara-Saika build
    of the ket centre instead of the electron transfer: the transfer relation
    multiplies by zeta/eta and, for exponent ranges like 0.06..685, loses all
    digits in FP64 (errors of O(1) on values of O(0.1)) and ~1e-3 even in x87
    precision, so it cannot serve as the accuracy reference.
    """
    bra = s1.angmom < s2.angmom
    if bra:
        s1, s2 = s2, s1
    ket = s3.angmom < s4.angmom
    if ket:
        s3, s4 = s4, s3
    bk = s1.angmom < s3.angmom
    if bk:
        s1, s3 = s3, s1
        s2, s4 = s4, s2
    out = _engine(s1, s2, s3, s4, exact, potential, omega)
    if bk:
        out = np.transpose(out, (2, 3, 0, 1, 6, 7, 4, 5))
    if ket:
        out = np.swapaxes(np.swapaxes(out, 2, 3), 6, 7)
    if bra:
        out = np.swapaxes(np.swapaxes(out, 0, 1), 4, 5)
    return np.transpose(out, (4, 0, 5, 1, 6, 2, 7, 3))


def _shell_out_matrix(shell, coord_type):
    """Matrix applied to a shell's Cartesian index after norm_cont (identity if Cartesian)."""
    if coord_type == "spherical":
        return generate_transformation(shell.angmom, shell.angmom_components_cart,
                                       shell.angmom_components_sph, "left")
    return np.eye(len(shell.angmom_components_cart))


def normalized_block(basis, coord_types, i, j, k, l, Q=None, threshold=0.0, exact=False,
                     potential="coulomb", omega=None):
    """Final (n1, n2, n3, n4) block (segmented contractions flattened m-major).

    Screened quartets (Q given and Q_ij Q_kl < threshold) are zero blocks.
    """
    shells = [basis[i], basis[j], basis[k], basis[l]]
    mats = [_shell_out_matrix(s, coord_types[x]) for s, x in zip(shells, (i, j, k, l))]
    if Q is not None and threshold > 0 and Q[i, j] * Q[k, l] < threshold:
        sizes = [s.coeffs.shape[1] * m.shape[0] for s, m in zip(shells, mats)]
        return np.zeros(sizes)
    blk = contraction_block(*shells, exact=exact, potential=potential, omega=omega)
    for axis, s in enumerate(shells):
        nc = np.asarray(s.norm_cont)
        shape = [1] * 8
        shape[2 * axis] = nc.shape[0]
        shape[2 * axis + 1] = nc.shape[1]
        blk = blk * nc.reshape(shape)
    # apply the Cartesian->pure matrices on the component axes 1,3,5,7
    blk = np.einsum("aA,bB,cC,dD,wAxByCzD->wa xb yc zd".replace(" ", ""), mats[0], mats[1],
                    mats[2], mats[3], blk, optimize=True)
    sh = blk.shape
    return blk.reshape(sh[0] * sh[1], sh[2] * sh[3], sh[4] * sh[5], sh[6] * sh[7])


def basis_offsets(basis, coord_types):
    sizes = []
    for s, ct in zip(basis, coord_types):
        ncomp = (2 * s.angmom + 1) if ct == "spherical" else len(s.angmom_components_cart)
        sizes.append(s.coeffs.shape[1] * ncomp)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return off


def dense_eri(basis, notation="physicist", schwarz_threshold=0.0, coord_types=None, exact=False,
              potential="coulomb", omega=None):
    """Dense (N,N,N,N) ERI tensor as the reference assembles it.  ``potential`` "erf" /
    "erfc" with ``omega``: the range-separated operators (Schwarz bounds stay Coulomb,
    gbasis/integrals/electron_repulsion.py:475)."""
    if notation not in ("physicist", "chemist"):
        raise ValueError("`notation` must be one of 'physicist' or 'chemist'")
    basis = list(basis)
    if coord_types is None:
        coord_types = [s.coord_type for s in basis]
    off = basis_offsets(basis, coord_types)
    n = int(off[-1])
    Q = schwarz_bounds(basis) if schwarz_threshold > 0 else None
    out = np.zeros((n, n, n, n))
    pairs = list(it.combinations_with_replacement(range(len(basis)), 2))
    for p, (i, j) in enumerate(pairs):
        for (k, l) in pairs[p:]:
            blk = normalized_block(basis, coord_types, i, j, k, l, Q, schwarz_threshold, exact,
                                   potential, omega)
            si, sj, sk, sl = (slice(off[x], off[x + 1]) for x in (i, j, k, l))
            out[si, sj, sk, sl] = blk
            out[si, sj, sl, sk] = blk.transpose(0, 1, 3, 2)
            out[sj, si, sk, sl] = blk.transpose(1, 0, 2, 3)
            out[sj, si, sl, sk] = blk.transpose(1, 0, 3, 2)
            out[sk, sl, si, sj] = blk.transpose(2, 3, 0, 1)
            out[sl, sk, si, sj] = blk.transpose(3, 2, 0, 1)
            out[sk, sl, sj, si] = blk.transpose(2, 3, 1, 0)
            out[sl, sk, sj, si] = blk.transpose(3, 2, 1, 0)
    if notation == "physicist":
        out = np.transpose(out, (0, 2, 1, 3))
    return out


def packed_index(i, j, k, l):
    """Canonical 8-fold slot of function quartet (i j|k l)."""
    i, j = (i, j) if i >= j else (j, i)
    k, l = (k, l) if k >= l else (l, k)
    ij = i * (i + 1) // 2 + j
    kl = k * (k + 1) // 2 + l
    if ij < kl:
        ij, kl = kl, ij
    return ij * (ij + 1) // 2 + kl


def pack_dense_chemist(dense):
    """Canonical packed vector from a dense chemist tensor."""
    n = dense.shape[0]
    npair = n * (n + 1) // 2
    out = np.empty(npair * (npair + 1) // 2)
    ii, jj = np.tril_indices(n)
    pair_i = ii
    pair_j = jj
    order = np.argsort(pair_i * (pair_i + 1) // 2 + pair_j)
    pair_i, pair_j = pair_i[order], pair_j[order]
    for ij in range(npair):
        i, j = pair_i[ij], pair_j[ij]
        out[ij * (ij + 1) // 2: ij * (ij + 1) // 2 + ij + 1] = dense[i, j, pair_i[:ij + 1],
                                                                     pair_j[:ij + 1]]
    return out
