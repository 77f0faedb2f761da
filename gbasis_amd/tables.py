"""Host-side shell, shell-pair and shell-quartet tables.

This is the host half of the ERI path.  It reads the duck-typed shell
attributes the reference reads (gbasis/base_four_symm.py:160-389,
gbasis/integrals/electron_repulsion.py:305-412): ``coord``, ``angmom``,
``exps``, ``coeffs`` (K, M), ``norm_cont`` (M, L_cart),
``angmom_components_cart``, ``angmom_components_sph`` and the per-shell
coordinate type, and produces flat arrays for the device:

* expanded shells: a generalised contraction with M columns becomes M
  consecutive segmented shells, which reproduces the reference's m-major
  function order (base_four_symm.py:207-213);
* per-shell output maps: the kernels compute Cartesian components in the
  canonical (reference default) order; each shell maps them to its own
  component order, scales by ``norm_cont`` and, for pure shells with l >= 2,
  applies the dense Cartesian->pure matrix of ``generate_transformation``;
* shell pairs (bra orientation l_a >= l_b) with their primitive-pair records
  {zeta, 1/(2 zeta), P, P-A, K', 0}, K' = c_a c_b N_a N_b exp(-ab/zeta |AB|^2)
  sqrt(2) pi^(5/4) / zeta with N = (2a/pi)^(3/4) (4a)^(l/2)
  (_two_elec_int_improved.py:44-84, :684-690);
* screened, class-bucketed shell-quartet lists (the reference's unique
  quartet loop, base_four_symm.py:189-194, with the Schwarz test of
  _schwarz_screening.py:184-191).

Packed output: every molecule owns a contiguous slice of canonical slots
``ijkl = ij (ij+1)/2 + kl`` with ``ij = i (i+1)/2 + j``, i >= j, k >= l,
ij >= kl over its output functions.
"""

import os
from dataclasses import dataclass, field

import numpy as np

from .programs import comps
from .spherical import generate_transformation
from .utils import factorial2

SQRT2_PI_5_4 = np.sqrt(2.0) * np.pi ** 1.25


def f2norm_table(lmax=8):
    """1/sqrt(prod (2a_i-1)!!) for the canonical components of l = 0..lmax."""
    vals = []
    for l in range(lmax + 1):
        c = np.array(comps(l))
        vals.append(1.0 / np.sqrt(np.prod(factorial2(2 * c - 1), axis=1)))
    return np.concatenate(vals)


@dataclass
class ExpandedShell:
    orig: int           # index of the shell in the caller's basis
    l: int
    coord: np.ndarray
    exps: np.ndarray
    coeffs: np.ndarray  # (K,)
    kind: int           # 0 = scaled permutation, 1 = dense Cartesian->pure matrix
    nout: int
    comp_map: np.ndarray    # (ncart,) int: output position of each canonical component
    comp_scale: np.ndarray  # (ncart,) float
    out_mat: np.ndarray = field(default=None)  # (nout, ncart) for kind 1
    func_off: int = 0


def _is_spherical(coord_type):
    return coord_type in ("spherical", "p")


def expand_shells(basis, coord_types):
    """Expand generalised contractions and build per-shell output maps."""
    out = []
    func = 0
    for idx, (shell, ctype) in enumerate(zip(basis, coord_types)):
        l = int(shell.angmom)
        cart = np.asarray(shell.angmom_components_cart).reshape(-1, 3)
        ncart = (l + 1) * (l + 2) // 2
        if cart.shape[0] != ncart:
            raise ValueError(f"shell {idx}: {cart.shape[0]} Cartesian components for l={l}")
        pos_of = {tuple(c): i for i, c in enumerate(cart.tolist())}
        perm = np.array([pos_of[c] for c in comps(l)], dtype=np.int64)  # canonical -> shell order
        coeffs = np.asarray(shell.coeffs, dtype=np.float64)
        if coeffs.ndim == 1:
            coeffs = coeffs[:, None]
        norm_cont = np.asarray(shell.norm_cont, dtype=np.float64).reshape(coeffs.shape[1], ncart)
        exps = np.asarray(shell.exps, dtype=np.float64)
        coord = np.asarray(shell.coord, dtype=np.float64).reshape(3)
        sph = _is_spherical(ctype)
        if sph:
            tmat = generate_transformation(l, cart, tuple(shell.angmom_components_sph), "left")
        for m in range(coeffs.shape[1]):
            nc = norm_cont[m]
            if not sph:
                es = ExpandedShell(idx, l, coord, exps, coeffs[:, m].copy(), 0, ncart,
                                   perm.astype(np.int32), nc[perm].copy())
            else:
                dense = tmat[:, perm] * nc[perm][None, :]  # (nout, ncart canonical)
                nz = np.count_nonzero(dense, axis=0)
                if l <= 1 and np.all(nz == 1):
                    rows = np.argmax(dense != 0, axis=0)
                    es = ExpandedShell(idx, l, coord, exps, coeffs[:, m].copy(), 0, 2 * l + 1,
                                       rows.astype(np.int32),
                                       dense[rows, np.arange(ncart)].copy())
                else:
                    es = ExpandedShell(idx, l, coord, exps, coeffs[:, m].copy(), 1, 2 * l + 1,
                                       np.arange(ncart, dtype=np.int32), np.ones(ncart),
                                       np.ascontiguousarray(dense))
            es.func_off = func
            func += es.nout
            out.append(es)
    return out, func


# Split threshold of second_centre_group: a primitive pair whose horizontal recursion
# from the first centre would multiply rounding by more than this is built on the
# second centre.  With 4096, plain FP64 (the kernels' operation order) stays within
# 7.2e-14 max(1, |value|) of binary128 on the 293 sampled quartets of the dense
# s,s,p,d, the s,p,d,f and the s-g workloads, where the first-centre build alone is off
# by up to 5.2e-4 (23 s-g quartets above 1e-12), and within 1.1e-13 on 32275 random
# quartets of 16 batch molecules (s, p, d shells, exponents up to 1e3), where it is off by
# up to 8.5e-12 (512: 1.0e-13; 32768: 9.1e-13).  It splits 0.9 % of the batch's shell
# pairs, 21 % of the s,p,d,f and 32 % of the s-g molecule's.
SPLIT_AMPLIFICATION = float(os.environ.get("GBX_SPLIT_AMPLIFICATION", "4096"))


def second_centre_group(l_second, e_first, e_second, ab):
    """True for the primitive pairs of a shell pair (first, second) that are built on
    the second centre.

    The Head-Gordon-Pople transfer [a, b+1_i| = [a+1_i, b| + AB_i [a, b| expands the
    second shell's polynomial about the first centre.  When the product Gaussian of a
    primitive pair sits near the second centre (tight exponent there, diffuse on the
    first), the terms it sums exceed their result by about r^(l_second),
    r = (|AB| + max(|PA|, s)) / max(|PB|, s), s = (2 zeta)^(-1/2) the width of the
    product Gaussian, and the horizontal recursion after contraction
    (gbasis/integrals/_two_elec_int_improved.py:364-588) passes the rounding of its
    inputs on multiplied by that factor: for g shells with exponents up to 1e5 it
    reaches 1e-4 Eh in FP64.  Such pairs are built on the second centre and transferred
    to the first by the mirrored recursion, where r is at most about 2.  Pairs are moved
    only when the second exponent is the larger, the second shell has l > 0, and
    r^(l_second) exceeds SPLIT_AMPLIFICATION.
    """
    l_second = np.asarray(l_second)
    e_first = np.asarray(e_first, dtype=np.float64)
    e_second = np.asarray(e_second, dtype=np.float64)
    ab = np.asarray(ab, dtype=np.float64)
    zeta = e_first + e_second
    s = 1.0 / np.sqrt(2.0 * zeta)
    pa = e_second / zeta * ab
    pb = e_first / zeta * ab
    r = (ab + np.maximum(pa, s)) / np.maximum(pb, s)
    return (l_second > 0) & (e_second > e_first) & (r ** l_second > SPLIT_AMPLIFICATION)


def prim_norm(exps, l):
    return (2.0 * exps / np.pi) ** 0.75 * (4.0 * exps) ** (l / 2.0)


class MoleculeTables:
    """Shell, pair and primitive-pair arrays of one molecule (host, NumPy)."""

    def __init__(self, basis, coord_types=None):
        basis = list(basis)
        if coord_types is None:
            coord_types = [s.coord_type for s in basis]
        self.n_orig = len(basis)
        self.shells, self.nfunc = expand_shells(basis, coord_types)
        ns = len(self.shells)
        self.nshell = ns
        # ---- shell pairs I >= J over expanded shells, ordered by (l_a, l_b, primitive
        # pairs): runs of equal pair classes make the device queue builder's
        # candidates (p >= q) share bucket keys within a wavefront ----
        ii, jj = np.tril_indices(ns)
        ls = np.array([s.l for s in self.shells])
        kp = np.array([s.exps.size for s in self.shells])
        hi = np.maximum(ls[ii], ls[jj])
        lo = np.minimum(ls[ii], ls[jj])
        order = np.lexsort((jj, ii, kp[ii] * kp[jj], lo, hi))
        ii, jj = ii[order], jj[order]
        self.npair = ii.size
        swap = ls[ii] < ls[jj]
        sa = np.where(swap, jj, ii)
        sb = np.where(swap, ii, jj)
        self.pair_shell_a = sa.astype(np.int32)
        self.pair_shell_b = sb.astype(np.int32)
        self.pair_la = ls[sa]
        self.pair_lb = ls[sb]
        self.pair_orig = (np.array([s.orig for s in self.shells])[ii],
                          np.array([s.orig for s in self.shells])[jj])
        nprim = np.array([s.exps.size for s in self.shells])
        self.pair_nprim = (nprim[sa] * nprim[sb]).astype(np.int32)
        self.pair_prim_off = np.concatenate([[0], np.cumsum(self.pair_nprim)[:-1]]).astype(
            np.int64)
        coords = np.array([s.coord for s in self.shells])
        self.pair_geom = np.zeros((self.npair, 4))
        self.pair_geom[:, :3] = coords[sa] - coords[sb]
        # ---- primitive-pair records, all pairs at once: record r of pair p is the
        # primitive combination (ia, ib) = divmod(r - first record of p, K_b) ----
        nprim_s = nprim.astype(np.int64)
        sh_off = np.concatenate([[0], np.cumsum(nprim_s)[:-1]])
        e_all = np.concatenate([s.exps for s in self.shells])
        cn_all = np.concatenate([s.coeffs * prim_norm(s.exps, s.l) for s in self.shells])
        npp = self.pair_nprim.astype(np.int64)
        rp = np.repeat(np.arange(self.npair), npp)
        t = np.arange(rp.size) - self.pair_prim_off[rp]
        kb = nprim_s[sb][rp]
        ga = sh_off[sa][rp] + t // kb
        gb = sh_off[sb][rp] + t % kb
        ea, eb = e_all[ga], e_all[gb]
        A, B = coords[sa][rp], coords[sb][rp]
        zeta = ea + eb
        P = (ea[:, None] * A + eb[:, None] * B) / zeta[:, None]
        d = coords[sa] - coords[sb]
        ab2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])[rp]
        K = cn_all[ga] * cn_all[gb] * np.exp(-(ea * eb / zeta) * ab2) * SQRT2_PI_5_4 / zeta
        recs = np.zeros((rp.size, 10))
        recs[:, 0] = zeta
        recs[:, 1] = 0.5 / zeta
        recs[:, 2:5] = P
        recs[:, 5:8] = P - A
        recs[:, 8] = K
        # ---- build centre of each primitive pair (second_centre_group): the records
        # of a pair are ordered with the first-centre group first; the second-centre
        # group is also stored with P - B in place of P - A, for the sub-pair (b, a) ----
        grp = second_centre_group(self.pair_lb[rp], ea, eb, np.sqrt(ab2))
        order = np.lexsort((np.arange(rp.size), grp, rp))
        recs = recs[order]
        grp = grp[order]
        self.prim_pairs = recs
        self.pair_nB = np.bincount(rp[grp], minlength=self.npair).astype(np.int32)
        self.pair_nA = (self.pair_nprim - self.pair_nB).astype(np.int32)
        sub = recs[grp].copy()
        Bs = coords[sb][rp][order][grp]
        sub[:, 5:8] = sub[:, 2:5] - Bs
        self.sub_prims = sub
        self.pair_subB_off = np.concatenate([[0], np.cumsum(self.pair_nB)[:-1]]).astype(np.int64)
        self.n_slots = _n_slots(self.nfunc)

    # ------------------------------------------------------------------
    def shell_arrays(self):
        """shell_info (ns, 8) int32, comp_map, comp_scale, out_mat (flat)."""
        info = np.zeros((self.nshell, 8), dtype=np.int32)
        cmap, cscale, mats = [], [], []
        moff = 0
        coff = 0
        for i, s in enumerate(self.shells):
            info[i, 0] = s.func_off
            info[i, 1] = s.l
            info[i, 2] = s.kind
            info[i, 3] = coff
            info[i, 5] = s.nout
            cmap.append(s.comp_map)
            cscale.append(s.comp_scale)
            coff += s.comp_map.size
            if s.kind == 1:
                info[i, 4] = moff
                mats.append(s.out_mat.ravel())
                moff += s.out_mat.size
        return (info, np.concatenate(cmap).astype(np.int32), np.concatenate(cscale),
                np.concatenate(mats) if mats else np.zeros(1))

    def schwarz_orig(self, q_pair):
        """Per original-shell-pair bound (max over the expanded pairs), (n_orig, n_orig)."""
        Q = np.zeros((self.n_orig, self.n_orig))
        oi, oj = self.pair_orig
        np.maximum.at(Q, (oi, oj), q_pair)
        Q = np.maximum(Q, Q.T)
        return Q

    def quartets(self, q_pair=None, threshold=0.0):
        """Unique, unscreened shell quartets as (bra pair, ket pair) in canonical orientation.

        Returns (bra, ket, cls_a, cls_b, cls_c, cls_d) int arrays.  With a threshold > 0,
        quartet (ij|kl) is dropped iff Q_ij * Q_kl < threshold where Q is the bound of the
        caller's (original) shells, as SchwarzScreener.is_significant.
        """
        P = self.npair
        p, q = np.tril_indices(P)
        if threshold > 0:
            Qo = self.schwarz_orig(q_pair)
            oi, oj = self.pair_orig
            qp = Qo[oi, oj]
            keep = qp[p] * qp[q] >= threshold
            p, q = p[keep], q[keep]
        la, lb = self.pair_la, self.pair_lb
        key_p = (la[p] + lb[p]) * 16 + la[p]
        key_q = (la[q] + lb[q]) * 16 + la[q]
        sw = key_p < key_q
        bra = np.where(sw, q, p).astype(np.int32)
        ket = np.where(sw, p, q).astype(np.int32)
        return bra, ket


MAX_PRIM_BIN = 15


def prim_bin(nprimq):
    """Primitive-count bin of the device work-queue key: ceil(log2(n)), at most 15 (4 bits
    of the key; the last bin holds every count above 2^14, the kernels read the true
    counts from the pair table)."""
    n = np.maximum(np.asarray(nprimq, dtype=np.int64), 1)
    b = np.zeros(n.shape, dtype=np.int64)
    while True:
        grow = (b < MAX_PRIM_BIN) & ((np.int64(1) << b) < n)
        if not grow.any():
            return b
        b += grow


def _n_slots(nfunc):
    npair = nfunc * (nfunc + 1) // 2
    return npair * (npair + 1) // 2


NUM_PHASES = 4
NUM_BUCKETS = NUM_PHASES * 4096 * 2 * 2 * 16
STORE_PHASES = (0,)   # phases whose launches store their packed slots; the later ones add


def queue_key(phase, cls, dense, fp32, nprimq):
    """Bucket key of the device work queue (gbx_queue.hip):
    ((((phase * 4096 + cls) * 2 + dense) * 2 + fp32) * 16 + prim_bin(nprimq)),
    cls = ((la * 8 + lb) * 8 + lc) * 8 + ld in the sub-quartet's own orientation."""
    return ((((np.asarray(phase, dtype=np.int64) * 4096 + cls) * 2 + dense) * 2 + fp32) * 16
            + prim_bin(nprimq))


def decode_head(head):
    """(phase, cls tuple, dense, fp32) of key >> 4."""
    fp32 = head & 1
    dense = (head >> 1) & 1
    c = head >> 2
    cls = c & 4095
    return c >> 12, ((cls >> 9) & 7, (cls >> 6) & 7, (cls >> 3) & 7, cls & 7), bool(dense), fp32


class BatchTables:
    """Tables of several molecules concatenated into one device-resident set.

    Pair entries 0 .. npair-1 are the shell pairs of the molecules (bra orientation
    l_a >= l_b) with all their primitive pairs: the Schwarz pass and the work-queue
    candidates use them.  A pair some of whose primitive pairs are built on the
    second centre (second_centre_group) has two more entries after all of those:
    its first-centre part (a, b) and its second-centre part (b, a), whose records
    hold P - B.  ``pair_sub[p]`` = (first-centre entry, second-centre entry or -1);
    the first-centre entry of an unsplit pair is the pair itself.  A shell quartet
    (p|q) is the sum of up to four sub-quartets, one per phase:
    (A_p|A_q), (B_p|A_q), (A_p|B_q), (B_p|B_q); phase 0 always runs and stores,
    the others add into the same packed slots.
    """

    def __init__(self, molecules):
        self.mols = molecules
        shell_off = pair_off = prim_off = 0
        slot_off = 0
        infos, cmaps, cscales, mats = [], [], [], []
        pinfo, pgeom, pbase, prims = [], [], [], []
        self.mol_pair_off, self.mol_slot_off = [], []
        coff = moff = 0
        split_rows = []   # per molecule: (global pair ids, nA, nB, sub record offsets)
        for m in molecules:
            info, cmap, cscale, mat = m.shell_arrays()
            info = info.copy()
            info[:, 3] += coff
            info[:, 4] += moff
            infos.append(info)
            cmaps.append(cmap)
            cscales.append(cscale)
            mats.append(mat)
            coff += cmap.size
            moff += mat.size
            pi = np.zeros((m.npair, 4), dtype=np.int64)
            pi[:, 0] = m.pair_shell_a + shell_off
            pi[:, 1] = m.pair_shell_b + shell_off
            pi[:, 2] = m.pair_prim_off + prim_off
            pi[:, 3] = m.pair_nprim
            pinfo.append(pi)
            pgeom.append(m.pair_geom)
            pbase.append(np.full(m.npair, slot_off, dtype=np.uint64))
            prims.append(m.prim_pairs)
            self.mol_pair_off.append(pair_off)
            self.mol_slot_off.append(slot_off)
            shell_off += m.nshell
            pair_off += m.npair
            prim_off += m.prim_pairs.shape[0]
            slot_off += m.n_slots
        # second-centre records after all first-centre ones
        sub_off = prim_off
        for mi, m in enumerate(molecules):
            split = np.flatnonzero(m.pair_nB > 0)
            split_rows.append((split + self.mol_pair_off[mi], split, sub_off))
            prims.append(m.sub_prims)
            sub_off += m.sub_prims.shape[0]
        npair = pair_off
        pinfo = np.concatenate(pinfo)
        pgeom = np.concatenate(pgeom)
        pbase = np.concatenate(pbase)
        pair_la = np.concatenate([m.pair_la for m in molecules])
        pair_lb = np.concatenate([m.pair_lb for m in molecules])
        pair_sub = np.stack([np.arange(npair), np.full(npair, -1)], axis=1).astype(np.int64)
        ext_info, ext_geom, ext_base, ext_la, ext_lb = [], [], [], [], []
        nxt = npair
        for mi, (gids, local, soff) in enumerate(split_rows):
            m = molecules[mi]
            n = gids.size
            if n == 0:
                continue
            a_ids = nxt + 2 * np.arange(n)
            pair_sub[gids, 0] = a_ids
            pair_sub[gids, 1] = a_ids + 1
            ia = pinfo[gids].copy()
            ia[:, 3] = m.pair_nA[local]
            ib = np.zeros_like(ia)
            ib[:, 0] = pinfo[gids, 1]
            ib[:, 1] = pinfo[gids, 0]
            ib[:, 2] = soff + m.pair_subB_off[local]
            ib[:, 3] = m.pair_nB[local]
            ext_info.append(np.stack([ia, ib], axis=1).reshape(-1, 4))
            ga = pgeom[gids]
            gb = ga.copy()
            gb[:, :3] = -gb[:, :3]
            ext_geom.append(np.stack([ga, gb], axis=1).reshape(-1, 4))
            ext_base.append(np.repeat(pbase[gids], 2))
            ext_la.append(np.stack([pair_la[gids], pair_lb[gids]], axis=1).reshape(-1))
            ext_lb.append(np.stack([pair_lb[gids], pair_la[gids]], axis=1).reshape(-1))
            nxt += 2 * n
        if ext_info:
            pinfo = np.concatenate([pinfo] + ext_info)
            pgeom = np.concatenate([pgeom] + ext_geom)
            pbase = np.concatenate([pbase] + ext_base)
            pair_la = np.concatenate([pair_la] + ext_la)
            pair_lb = np.concatenate([pair_lb] + ext_lb)
        if sub_off >= 2 ** 31 or nxt >= 2 ** 31:
            raise ValueError("batch too large for 32-bit pair / primitive indices")
        self.shell_info = np.concatenate(infos)
        self.comp_map = np.concatenate(cmaps).astype(np.int32)
        self.comp_scale = np.concatenate(cscales)
        self.out_mat = np.concatenate(mats)
        self.pair_info = pinfo.astype(np.int32)
        self.pair_geom = pgeom
        self.pair_out_base = pbase
        self.prim_pairs = np.concatenate(prims)
        self.n_slots = slot_off
        self.npair = npair            # shell pairs (Schwarz pass, queue candidates)
        self.npair_all = nxt          # plus the sub-pair entries
        self.pair_sub = pair_sub.astype(np.int32)
        # shell pair of every entry (a sub-pair entry maps to the pair it splits)
        entry_pair = np.arange(nxt, dtype=np.int64)
        split_ids = np.flatnonzero(pair_sub[:, 1] >= 0)
        entry_pair[pair_sub[split_ids, 0]] = split_ids
        entry_pair[pair_sub[split_ids, 1]] = split_ids
        self.entry_pair = entry_pair
        self.pair_la = pair_la
        self.pair_lb = pair_lb
        kind = self.shell_info[:, 2]
        # pairs touching a pure shell with l >= 2 need the dense output transform
        self.pair_dense = (kind[self.pair_info[:, 0]] == 1) | (kind[self.pair_info[:, 1]] == 1)
        # per pair entry, key for the device queue builder: la | lb << 3 | dense << 6 |
        # nprim << 8 (la, lb in the entry's own orientation)
        self.pair_key = (self.pair_la.astype(np.int64) | (self.pair_lb.astype(np.int64) << 3)
                         | (self.pair_dense.astype(np.int64) << 6)
                         | (self.pair_info[:, 3].astype(np.int64) << 8)).astype(np.int32)
        nout = self.shell_info[:, 5].astype(np.int64)
        sa, sb = self.pair_info[:, 0], self.pair_info[:, 1]
        self.pair_npf = np.where(sa == sb, nout[sa] * (nout[sa] + 1) // 2,
                                 nout[sa] * nout[sb]).astype(np.int32)
        npairs = np.array([m.npair for m in molecules], dtype=np.int64)
        self.mol_cand_off = np.concatenate([[0], np.cumsum(npairs * (npairs + 1) // 2)])
        self.n_cand = int(self.mol_cand_off[-1])
        # generalised contractions: the Schwarz bound of an expanded pair is the max over
        # the pairs of the same original shells, which the host folds in
        self.generalized = any(m.nshell != m.n_orig for m in molecules)
        self._capacity = None

    def sub_quartets(self, bra, ket):
        """Sub-quartets of shell quartets (bra, ket) (pair entries < npair, oriented as
        the work queue orients them): a list of (phase, bra entries, ket entries)."""
        bra = np.asarray(bra, dtype=np.int64)
        ket = np.asarray(ket, dtype=np.int64)
        ab, bb = self.pair_sub[bra, 0], self.pair_sub[bra, 1]
        ak, bk = self.pair_sub[ket, 0], self.pair_sub[ket, 1]
        out = [(0, ab, ak)]
        for phase, (x, y) in ((1, (bb, ak)), (2, (ab, bk)), (3, (bb, bk))):
            sel = (x >= 0) & (y >= 0)
            if sel.any():
                out.append((phase, x[sel], y[sel]))
        return out

    def bucket_capacity(self):
        """Sub-quartets per bucket key (precision bit 0) over all unscreened candidates:
        an upper bound of every bucket the device queue builder can fill, computed from
        the histogram of the pairs' sub-pair keys in each molecule (pairs of equal
        weight (la + lb, la) are counted in both bra/ket orders, since the device orders
        them by pair index)."""
        if self._capacity is not None:
            return self._capacity
        cap = np.zeros(NUM_BUCKETS, dtype=np.int64)
        for mi, m in enumerate(self.mols):
            po = self.mol_pair_off[mi]
            g = np.arange(po, po + m.npair)
            ka = self.pair_key[self.pair_sub[g, 0]].astype(np.int64)
            sb_ = self.pair_sub[g, 1]
            kb_ = np.where(sb_ >= 0, self.pair_key[np.maximum(sb_, 0)].astype(np.int64), -1)
            tup, counts = np.unique(np.stack([ka, kb_], axis=1), axis=0, return_counts=True)
            counts = counts.astype(np.int64)
            ki, kj = np.meshgrid(np.arange(len(tup)), np.arange(len(tup)), indexing="ij")
            sel = ki >= kj
            ki, kj = ki[sel], kj[sel]
            n = np.where(ki == kj, counts[ki] * (counts[ki] + 1) // 2, counts[ki] * counts[kj])
            ti, tj = tup[ki], tup[kj]
            la_p, lb_p = ti[:, 0] & 7, (ti[:, 0] >> 3) & 7
            la_q, lb_q = tj[:, 0] & 7, (tj[:, 0] >> 3) & 7
            sp, sq = (la_p + lb_p) * 16 + la_p, (la_q + lb_q) * 16 + la_q
            swap = sp < sq
            tb = np.where(swap[:, None], tj, ti)
            tk = np.where(swap[:, None], ti, tj)
            # two different tuples of equal weight: the device puts the pair of higher
            # index in the bra, so either tuple may be the bra -- count both orders
            tie = (sp == sq) & (ki != kj)
            tb, tk = np.concatenate([tb, tj[tie]]), np.concatenate([tk, ti[tie]])
            n = np.concatenate([n, n[tie]])
            dense = ((tb[:, 0] >> 6) & 1) | ((tk[:, 0] >> 6) & 1)
            for phase, (xb, xk) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
                kb, kk = tb[:, xb], tk[:, xk]
                ok = (kb >= 0) & (kk >= 0)
                kb, kk = kb[ok], kk[ok]
                cls = (((kb & 7) * 8 + ((kb >> 3) & 7)) * 8 + (kk & 7)) * 8 + ((kk >> 3) & 7)
                key = queue_key(phase, cls, dense[ok], 0, (kb >> 8) * (kk >> 8))
                np.add.at(cap, key, n[ok])
        self._capacity = cap
        return cap

    def schwarz_buckets(self):
        """(bra == ket) quartet lists of the Schwarz pass, bucketed as class_buckets."""
        p = np.arange(self.npair, dtype=np.int32)
        return self.class_buckets(p, p)

    def unique_counts(self, bra, ket):
        """Number of canonical packed slots each (bra, ket) quartet writes."""
        nout = self.shell_info[:, 5].astype(np.int64)
        sa, sb = self.pair_info[:, 0], self.pair_info[:, 1]
        npf = np.where(sa == sb, nout[sa] * (nout[sa] + 1) // 2, nout[sa] * nout[sb])
        nb, nk = npf[bra], npf[ket]
        return np.where(bra == ket, nb * (nb + 1) // 2, nb * nk)

    def class_buckets(self, bra, ket):
        """Group global (bra, ket) quartets by class, dense-transform flag and
        primitive-count bin.

        Returns a list of (cls, nprimq_max, dense, quartets (n,2) int32) sorted by class.
        """
        la, lb = self.pair_la[bra], self.pair_lb[bra]
        lc, ld = self.pair_la[ket], self.pair_lb[ket]
        nprim = self.pair_info[:, 3].astype(np.int64)
        npq = nprim[bra] * nprim[ket]
        binid = np.ceil(np.log2(np.maximum(npq, 1))).astype(np.int64)
        dense = (self.pair_dense[bra] | self.pair_dense[ket]).astype(np.int64)
        key = ((((la * 8 + lb) * 8 + lc) * 8 + ld) * 2 + dense) * 64 + binid
        order = np.lexsort((nprim[ket], nprim[bra], key))
        key_s = key[order]
        bounds = np.flatnonzero(np.diff(key_s)) + 1
        starts = np.concatenate([[0], bounds])
        ends = np.concatenate([bounds, [key_s.size]])
        out = []
        for s, e in zip(starts.tolist(), ends.tolist()):
            idx = order[s:e]
            k = int(key_s[s])
            cls = ((k >> 16) & 7, (k >> 13) & 7, (k >> 10) & 7, (k >> 7) & 7)
            quart = np.stack([bra[idx], ket[idx]], axis=1).astype(np.int32)
            out.append((cls, int(npq[idx].max()), bool((k >> 6) & 1), quart))
        return out
