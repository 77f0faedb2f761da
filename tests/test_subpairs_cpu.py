"""Primitive pairs split by build centre (host tables, CPU).

The device evaluates a shell quartet as the sum of sub-quartets whose pair entries
hold the primitive pairs built on the first or on the second centre
(gbasis_amd.tables.second_centre_group).  These tests replay, in NumPy, exactly what
the kernels compute from the tables -- records, pair entries, class programs in the
entries' own orientation, phases -- and compare with the binary128 oracle on shells
with exponents up to 1e5, where the first-centre build alone loses digits."""
import numpy as np
import pytest

from gbasis_amd.contractions import GeneralizedContractionShell
from gbasis_amd.programs import PH_HRRK, PH_VRR, class_program, comps
from gbasis_amd.tables import (BatchTables, MoleculeTables, decode_head, queue_key,
                               second_centre_group)
from oracle.boys import boys_hyp1f1
from oracle.hp import hp_block_raw


def _f2(c):
    def df(n):
        r = 1
        while n > 1:
            r *= n
            n -= 2
        return r
    return 1 / np.sqrt(df(2 * c[0] - 1) * df(2 * c[1] - 1) * df(2 * c[2] - 1))


def replay_entry_quartet(b, eb, ek, prim_acc=None):
    """Cartesian block (entry orientation, canonical component order, factorial2 norms,
    no norm_cont) of the sub-quartet (eb|ek) of BatchTables b, as the kernels compute
    it: primitive quartets of the two entries' records, the class program of the
    entries' angular momenta, horizontal recursion with the entries' A-B and C-D.
    ``prim_acc(prog, rb, rk)``, if given, replaces the Obara-Saika primitive part: it
    returns the contracted [e0|f0] in the program's accumulator order."""
    la, lb = int(b.pair_la[eb]), int(b.pair_lb[eb])
    lc, ld = int(b.pair_la[ek]), int(b.pair_lb[ek])
    prog = class_program(la, lb, lc, ld)
    ib, nb = int(b.pair_info[eb, 2]), int(b.pair_info[eb, 3])
    ik, nk = int(b.pair_info[ek, 2]), int(b.pair_info[ek, 3])
    rb = np.repeat(b.prim_pairs[ib:ib + nb], nk, axis=0)
    rk = np.tile(b.prim_pairs[ik:ik + nk], (nb, 1))
    acc = np.zeros(prog.nacc)
    if prim_acc is not None:
        acc = prim_acc(prog, rb, rk)
    elif rb.shape[0]:
        zeta, eta = rb[:, 0], rk[:, 0]
        inv = 1 / (zeta + eta)
        PQ = rb[:, 2:5] - rk[:, 2:5]
        T = zeta * eta * inv * np.sum(PQ * PQ, axis=1)
        pref = rb[:, 8] * rk[:, 8] * np.sqrt(inv)
        PA, QC = rb[:, 5:8], rk[:, 5:8]
        roz, roe = eta * inv, zeta * inv
        WP, WQ = -roz[:, None] * PQ, roe[:, None] * PQ
        i2z, i2e, i2s = rb[:, 1], rk[:, 1], 0.5 * inv
        S = np.zeros((prog.prim_slots, rb.shape[0]))
        L = prog.L
        S[:L + 1] = boys_hyp1f1(np.arange(L + 1)[:, None], T[None, :]) * pref[None, :]
        for kind, items in prog.prim_phases:
            for dst, s0, s1, s2, s3, s4, code, _ in items:
                di, k1, k2 = code & 15, (code >> 4) & 255, (code >> 12) & 255
                if kind == PH_VRR:
                    v = PA[:, di] * S[s0] + WP[:, di] * S[s1]
                    if s2 >= 0:
                        v = v + k1 * i2z * (S[s2] - roz * S[s3])
                else:
                    v = QC[:, di] * S[s0] + WQ[:, di] * S[s1]
                    if s2 >= 0:
                        v = v + k1 * i2e * (S[s2] - roe * S[s3])
                    if s4 >= 0:
                        v = v + k2 * i2s * S[s4]
                S[dst] = v
        for it in prog.acc_phase[1]:
            acc[it[0]] += S[it[1]].sum()
    H = np.zeros(prog.h_slots)
    H[:prog.nacc] = acc
    AB, CD = b.pair_geom[eb, :3], b.pair_geom[ek, :3]
    for kind, items in prog.hrr_phases:
        g = CD if kind == PH_HRRK else AB
        for it in items:
            H[it[0]] = H[it[1]] + g[it[3]] * H[it[2]]
    outs = [(a, bb, c, d) for a in comps(la) for bb in comps(lb) for c in comps(lc)
            for d in comps(ld)]
    vals = np.array([H[s] * _f2(a) * _f2(bb) * _f2(c) * _f2(d)
                     for s, (a, bb, c, d) in zip(prog.out_slots, outs)])
    return vals.reshape([len(comps(x)) for x in (la, lb, lc, ld)])


def replay_quartet(b, bra, ket, prim_acc=None):
    """Sum of the sub-quartets of shell quartet (bra|ket), in the orientation of the
    pairs (a, b | c, d)."""
    sa, sb = b.pair_info[bra, :2]
    total = None
    for phase, xb, xk in b.sub_quartets([bra], [ket]):
        eb, ek = int(xb[0]), int(xk[0])
        blk = replay_entry_quartet(b, eb, ek, prim_acc)
        if b.pair_info[eb, 0] != sa:        # bra entry (b, a)
            blk = blk.transpose(1, 0, 2, 3)
        if b.pair_info[ek, 0] != b.pair_info[ket, 0]:
            blk = blk.transpose(0, 1, 3, 2)
        total = blk if total is None else total + blk
    return total


def _wide_basis(seed, ls, k=3):
    rng = np.random.default_rng(seed)
    out = []
    for i, l in enumerate(ls):
        exps = np.exp(rng.uniform(np.log(0.02), np.log(1e5), k))
        out.append(GeneralizedContractionShell(l, rng.uniform(0, 3.0, 3), rng.uniform(0.1, 1, k),
                                               exps, "cartesian"))
    return out


def test_second_centre_group_rule():
    """Pairs move only with l_second > 0, the larger exponent on the second centre and a
    large amplification r^l_second; coincident centres never move.  For exponents 0.1
    and 50 at 2 bohr, r = 40: enough for l = 3 (64000 > 4096), not for l = 2 (1600)."""
    ab = np.array([2.0, 2.0, 2.0, 0.0, 2.0, 2.0, 2.0])
    e1 = np.array([0.1, 50.0, 0.1, 0.1, 0.1, 1.0, 0.1])
    e2 = np.array([50.0, 0.1, 50.0, 50.0, 0.12, 1.0, 50.0])
    l2 = np.array([3, 3, 0, 3, 4, 4, 2])
    assert second_centre_group(l2, e1, e2, ab).tolist() == [True, False, False, False, False,
                                                            False, False]


def test_sub_pair_entries():
    """Entries of a split pair: the first-centre part is a prefix of the pair's records,
    the second-centre part is (b, a) with P - B, and the parts hold every primitive pair
    once."""
    basis = _wide_basis(3, [2, 1, 0, 2])
    m = MoleculeTables(basis)
    b = BatchTables([m])
    split = np.flatnonzero(b.pair_sub[:, 1] >= 0)
    assert split.size > 0 and b.npair_all == b.npair + 2 * split.size
    for p in split.tolist():
        ea, eb = b.pair_sub[p]
        assert ea >= b.npair and eb == ea + 1
        a, bb, off, n = b.pair_info[p]
        assert tuple(b.pair_info[ea, :3]) == (a, bb, off)
        assert tuple(b.pair_info[eb, :2]) == (bb, a)
        assert b.pair_info[ea, 3] + b.pair_info[eb, 3] == n
        assert (b.pair_la[eb], b.pair_lb[eb]) == (b.pair_lb[p], b.pair_la[p])
        np.testing.assert_array_equal(b.pair_geom[eb, :3], -b.pair_geom[p, :3])
        recs_b = b.prim_pairs[b.pair_info[eb, 2]:b.pair_info[eb, 2] + b.pair_info[eb, 3]]
        rest = b.prim_pairs[off + b.pair_info[ea, 3]:off + n]
        np.testing.assert_array_equal(recs_b[:, [0, 1, 2, 3, 4, 8]], rest[:, [0, 1, 2, 3, 4, 8]])
        Bc = m.shells[bb].coord
        np.testing.assert_allclose(recs_b[:, 5:8], recs_b[:, 2:5] - Bc, rtol=0, atol=1e-15)
    # unsplit pairs are their own first-centre entry
    uns = np.flatnonzero(b.pair_sub[:, 1] < 0)
    assert np.array_equal(b.pair_sub[uns, 0], uns)


def _replay_worst(ls):
    basis = _wide_basis(11 + sum(ls), ls)
    m = MoleculeTables(basis)
    b = BatchTables([m])
    bra, ket = m.quartets()
    worst = 0.0
    for p, q in list(zip(bra.tolist(), ket.tolist()))[::3]:
        got = replay_quartet(b, p, q)
        sh = [m.shells[int(x)] for x in (*b.pair_info[p, :2], *b.pair_info[q, :2])]
        ref = hp_block_raw([s.l for s in sh], [s.coord for s in sh], [s.exps for s in sh],
                           [s.coeffs for s in sh], "quad", 0)
        worst = max(worst, np.abs(got - ref).max() / max(1.0, np.abs(ref).max()))
    return worst, b


@pytest.mark.parametrize("ls", [[1, 1, 0, 0], [2, 1, 1, 0], [2, 1, 2, 1], [2, 2, 1, 1],
                                [3, 2, 1, 0], [3, 3, 2, 2]])
def test_replayed_split_quartets_vs_binary128(ls):
    """Quartets of shells with exponents 0.02 .. 1e5 as the device computes them from the
    split tables, against the binary128 oracle: within 1e-14 of the largest value."""
    worst, _ = _replay_worst(ls)
    assert worst <= 1e-14, worst


def test_unsplit_tables_miss_the_bar(monkeypatch):
    """The same replay with every primitive pair built on its first centre (the
    reference's orientation) is off by more than 1e-13 on the (ff|dd) case: the bar of
    the test above discriminates."""
    import gbasis_amd.tables as tables

    monkeypatch.setattr(tables, "SPLIT_AMPLIFICATION", np.inf)
    worst, b = _replay_worst([3, 3, 2, 2])
    assert not (b.pair_sub[:, 1] >= 0).any()
    assert worst > 1e-13, worst


def test_queue_key_layout():
    """Bucket key fields: phase, class, dense, precision, primitive bin."""
    k = int(queue_key(3, ((1 * 8 + 2) * 8 + 2) * 8 + 1, 1, 1, 40000))
    assert k & 15 == 15
    assert decode_head(k >> 4) == (3, (1, 2, 2, 1), True, 1)
    assert int(queue_key(0, 0, 0, 0, 1)) == 0


@pytest.mark.parametrize("which", ["ethylene", "batch", "high_l", "wide"])
def test_bucket_capacity_bounds_every_bucket(which):
    """The host capacity of every work-queue bucket (used when the queue is built without
    a counting pass, e.g. inside a HIP graph) is at least the number of sub-quartets the
    queue builder puts there, enumerated here on the host with its orientation rule:
    pairs of equal weight are ordered by pair index, so sub-quartets (B_p|A_q) and
    (A_q|B_p) of tied pairs land in different buckets."""
    import os

    from gbasis_amd import workloads
    from gbasis_amd.parsers import fchk_basis, read_fchk

    if which == "ethylene":
        basis = fchk_basis(read_fchk(os.path.join(os.path.dirname(__file__), "golden",
                                                  "C2H4_hf_ccpvdz.fchk")))
        mols = [MoleculeTables(basis)]
    elif which == "batch":
        mols = [MoleculeTables(workloads.batch_molecule(s)) for s in (1000, 1001, 1002)]
    elif which == "high_l":
        mols = [MoleculeTables(workloads.high_l_pure())]
    else:
        mols = [MoleculeTables(_wide_basis(5, [2, 1, 1, 0, 2, 1]))]
    b = BatchTables(mols)
    exact = np.zeros(b.bucket_capacity().size, np.int64)
    for mi, m in enumerate(mols):
        bra, ket = m.quartets()
        po = b.mol_pair_off[mi]
        for phase, xb, xk in b.sub_quartets(bra.astype(np.int64) + po, ket.astype(np.int64) + po):
            cls = ((b.pair_la[xb] * 8 + b.pair_lb[xb]) * 8 + b.pair_la[xk]) * 8 + b.pair_lb[xk]
            dense = (b.pair_dense[xb] | b.pair_dense[xk]).astype(np.int64)
            npq = b.pair_info[xb, 3].astype(np.int64) * b.pair_info[xk, 3]
            np.add.at(exact, queue_key(phase, cls, dense, 0, npq), 1)
    cap = b.bucket_capacity()
    assert np.all(cap >= exact), int((exact - cap).clip(0).sum())
    if which in ("high_l", "wide"):
        assert (b.pair_sub[:, 1] >= 0).any()
