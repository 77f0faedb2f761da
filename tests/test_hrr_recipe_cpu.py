"""Recipe of the LDS horizontal-recursion pass (gbx_eri_hrr_lds), replayed in NumPy the
way the kernel runs it (ket levels on every row e, then bra levels on every column cd),
against the horizontal-recursion phases of the class program that the cooperative pass
interprets: the same steps, so the same values."""
import numpy as np
import pytest

from gbasis_amd.programs import PH_HRRK, class_program, comps, hrr_recipe, ncart


def replay_recipe(cls, acc, AB, CD):
    la, lb, lc, ld = cls
    rec = hrr_recipe(*cls)
    NA, NC, nlk, nlb = (int(x) for x in rec[:4])
    sizes = [int(x) for x in rec[4:4 + nlk + nlb]]
    o = 4 + nlk + nlb
    cur = acc.reshape(NC, NA).T.copy()          # rows e, level 0 of the ket: c
    for s in sizes[:nlk]:
        ent = rec[o:o + s]
        o += s
        cur = cur[:, ent & 0xFFF] + CD[ent >> 24][None, :] * cur[:, (ent >> 12) & 0xFFF]
    assert cur.shape[1] == ncart(lc) * ncart(ld)
    cur = cur.T.copy()                          # columns cd, level 0 of the bra: e
    for s in sizes[nlk:]:
        ent = rec[o:o + s]
        o += s
        cur = cur[:, ent & 0xFFF] + AB[ent >> 24][None, :] * cur[:, (ent >> 12) & 0xFFF]
    assert o == rec.size and cur.shape[1] == ncart(la) * ncart(lb)
    return cur.T.reshape(ncart(la), ncart(lb), ncart(lc), ncart(ld))


def program_hrr(cls, acc, AB, CD):
    prog = class_program(*cls)
    H = np.zeros(prog.h_slots)
    H[:prog.nacc] = acc
    for kind, items in prog.hrr_phases:
        g = CD if kind == PH_HRRK else AB
        writes = [(it[0], H[it[1]] + g[it[3]] * H[it[2]]) for it in items]
        for dst, v in writes:
            H[dst] = v
    return H[prog.out_slots].reshape([len(comps(x)) for x in cls])


@pytest.mark.parametrize("cls", [(2, 1, 0, 0), (1, 1, 1, 1), (3, 1, 2, 0), (2, 2, 2, 2),
                                 (3, 2, 3, 1), (3, 3, 3, 3), (4, 0, 4, 0), (4, 2, 3, 3),
                                 (4, 4, 4, 3)])
def test_recipe_matches_program_hrr(cls):
    rng = np.random.default_rng(sum(cls))
    prog = class_program(*cls)
    acc = rng.standard_normal(prog.nacc)
    AB, CD = rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)
    got = replay_recipe(cls, acc, AB, CD)
    ref = program_hrr(cls, acc, AB, CD)
    assert np.array_equal(got, ref)


def test_recipe_header():
    rec = hrr_recipe(4, 4, 4, 4)
    assert list(rec[:4]) == [145, 145, 4, 4]
    assert rec[4 + 3] == 225 and rec[4 + 4 + 3] == 225   # last levels: the (c,d), (a,b) blocks
    ent = rec[4 + 8:]
    assert ent.min() >= 0 and (ent >> 24).max() <= 2
