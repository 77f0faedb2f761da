"""Class programs (the recursion schedule both kernel families execute), run on the
CPU in NumPy, against the oracle and an extended-precision evaluation."""
import numpy as np
import pytest

from gbasis_amd.programs import PH_HRRK, PH_VRR, class_program, comps
from oracle.boys import boys_hyp1f1
from oracle.os_hgp import os_hgp_quartet


def _f2(c):
    def df(n):
        r = 1
        while n > 1:
            r *= n
            n -= 2
        return r
    return 1 / np.sqrt(df(2 * c[0] - 1) * df(2 * c[1] - 1) * df(2 * c[2] - 1))


def run_program(cls, A, B, C, D, ex, co, compact=False):
    """Execute a class program exactly as the kernels do (FP64, NumPy).

    ``compact`` runs the slot-reusing form the cooperative kernel interprets, with
    every phase evaluated as one parallel step (all reads before any write).
    """
    la, lb, lc, ld = cls
    prog = class_program(*cls)
    prim_phases = prog.c_prim_phases if compact else prog.prim_phases
    acc_phase = prog.c_acc_phase if compact else prog.acc_phase
    hrr_phases = prog.c_hrr_phases if compact else prog.hrr_phases
    out_slots = prog.c_out_slots if compact else prog.out_slots
    nslot = prog.c_prim_slots if compact else prog.prim_slots
    hslot = prog.c_h_slots if compact else prog.h_slots
    L = prog.L
    acc = np.zeros(prog.nacc)

    def N(e, l):
        return (2 * e / np.pi) ** 0.75 * (4 * e) ** (l / 2)

    for i, a in enumerate(ex[0]):
        for j, b in enumerate(ex[1]):
            for k, c in enumerate(ex[2]):
                for l, d in enumerate(ex[3]):
                    z, e = a + b, c + d
                    inv = 1 / (z + e)
                    P, Q = (a * A + b * B) / z, (c * C + d * D) / e
                    PQ = P - Q
                    T = z * e * inv * PQ @ PQ
                    kab = co[0][i] * co[1][j] * N(a, la) * N(b, lb) * np.exp(-a * b / z * np.sum((A - B) ** 2))
                    kcd = co[2][k] * co[3][l] * N(c, lc) * N(d, ld) * np.exp(-c * d / e * np.sum((C - D) ** 2))
                    S = np.zeros(nslot)
                    S[:L + 1] = boys_hyp1f1(np.arange(L + 1), T) * 2 * np.pi ** 2.5 / (z * e) * np.sqrt(inv) * kab * kcd
                    PA, WP, roz, i2z = P - A, -e * inv * PQ, e * inv, 0.5 / z
                    QC, WQ, roe, i2e, i2s = Q - C, z * inv * PQ, z * inv, 0.5 / e, 0.5 * inv
                    for kind, items in prim_phases:
                        writes = []
                        for dst, s0, s1, s2, s3, s4, code, _ in items:
                            di, k1, k2 = code & 15, (code >> 4) & 255, (code >> 12) & 255
                            if kind == PH_VRR:
                                v = PA[di] * S[s0] + WP[di] * S[s1]
                                if s2 >= 0:
                                    v += k1 * i2z * (S[s2] - roz * S[s3])
                            else:
                                v = QC[di] * S[s0] + WQ[di] * S[s1]
                                if s2 >= 0:
                                    v += k1 * i2e * (S[s2] - roe * S[s3])
                                if s4 >= 0:
                                    v += k2 * i2s * S[s4]
                            writes.append((dst, v))
                        for dst, v in writes:
                            S[dst] = v
                    for it in acc_phase[1]:
                        acc[it[0]] += S[it[1]]
    H = np.zeros(hslot)
    H[:prog.nacc] = acc
    for kind, items in hrr_phases:
        g = (C - D) if kind == PH_HRRK else (A - B)
        writes = [(it[0], H[it[1]] + g[it[3]] * H[it[2]]) for it in items]
        for dst, v in writes:
            H[dst] = v
    outs = [(a, b, c, d) for a in comps(la) for b in comps(lb) for c in comps(lc) for d in comps(ld)]
    vals = np.array([H[s] * _f2(a) * _f2(b) * _f2(c) * _f2(d)
                     for s, (a, b, c, d) in zip(out_slots, outs)])
    return vals.reshape([len(comps(x)) for x in cls])


@pytest.mark.parametrize("cls", [(0, 0, 0, 0), (1, 0, 0, 0), (1, 1, 0, 0), (1, 0, 1, 0),
                                 (1, 1, 1, 1), (2, 0, 1, 0), (2, 1, 1, 1), (2, 2, 2, 1),
                                 (3, 1, 2, 0), (2, 2, 2, 2), (3, 0, 3, 0)])
@pytest.mark.parametrize("compact", [False, True])
def test_program_matches_oracle(cls, compact):
    rng = np.random.default_rng(sum(cls) * 7 + cls[0])
    A, B, C, D = rng.uniform(-1, 1, (4, 3))
    ex = [rng.uniform(0.2, 3, 2) for _ in range(4)]
    co = [rng.uniform(0.1, 1, 2) for _ in range(4)]
    la, lb, lc, ld = cls
    got = run_program(cls, A, B, C, D, ex, co, compact)
    ref = os_hgp_quartet(A, la, comps(la), ex[0], co[0], B, lb, comps(lb), ex[1], co[1],
                         C, lc, comps(lc), ex[2], co[2], D, ld, comps(ld), ex[3], co[3])
    ref = ref[..., 0, 0, 0, 0]
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_program_sizes_monotone():
    """Sanity of the schedule: intermediates grow with l, and every output is reachable."""
    small, big = class_program(1, 1, 1, 1), class_program(2, 2, 2, 2)
    assert big.prim_slots > small.prim_slots and big.h_slots > small.h_slots
    for p in (small, big):
        assert len(set(p.out_slots.tolist())) == p.nout
        assert len(set(p.c_out_slots.tolist())) == p.nout
        assert p.c_prim_slots < p.prim_slots and p.c_h_slots < p.h_slots


@pytest.mark.parametrize("cls", [(0, 0, 0, 0), (1, 0, 1, 0), (2, 2, 2, 1), (3, 3, 3, 3),
                                 (4, 4, 4, 4)])
def test_encoded_items_round_trip(cls):
    """The 16-byte items the cooperative kernel reads decode to the compact program."""
    from gbasis_amd.programs import PH_ACC, PH_KVRR

    p = class_program(*cls)
    header, items, phases = p.encode()
    w = items.view(np.uint32).reshape(-1, 4)
    assert header[7] == w.shape[0]
    progs = p.c_prim_phases + [p.c_acc_phase] + p.c_hrr_phases
    for (kind, its), (k2, i0, i1) in zip(progs, phases.reshape(-1, 3)):
        assert kind == k2 and i1 - i0 == len(its)
        for it, row in zip(its, w[i0:i1]):
            if kind in (PH_VRR, PH_KVRR):
                f = [row[0] & 0xFFFF, row[0] >> 16, row[1] & 0xFFFF, row[1] >> 16,
                     row[2] & 0xFFFF, row[2] >> 16]
                dec = [int(v) if v != 0xFFFF else -1 for v in f] + [int(row[3])]
                assert dec == list(it[:7])
            elif kind == PH_ACC:
                assert [int(row[0]), int(row[1])] == list(it[:2])
            else:
                assert [int(x) for x in row] == list(it[:4])


def _prim_acc(prog, boys, sc):
    """Accumulators of one primitive quartet from a compact program: boys = S[0..L],
    sc = (PA, WP, roz, i2z, QC, WQ, roe, i2e, i2s).  Returns {acc row: value}."""
    PA, WP, roz, i2z, QC, WQ, roe, i2e, i2s = sc
    S = np.zeros(prog.c_prim_slots)
    S[:prog.L + 1] = boys
    for kind, items in prog.c_prim_phases:
        writes = []
        for dst, s0, s1, s2, s3, s4, code, _ in items:
            di, k1, k2 = code & 15, (code >> 4) & 255, (code >> 12) & 255
            if kind == PH_VRR:
                v = PA[di] * S[s0] + WP[di] * S[s1]
                if s2 >= 0:
                    v += k1 * i2z * (S[s2] - roz * S[s3])
            else:
                v = QC[di] * S[s0] + WQ[di] * S[s1]
                if s2 >= 0:
                    v += k1 * i2e * (S[s2] - roe * S[s3])
                if s4 >= 0:
                    v += k2 * i2s * S[s4]
            writes.append((dst, v))
        for dst, v in writes:
            S[dst] = v
    return {it[0]: S[it[1]] for it in prog.c_acc_phase[1]}


@pytest.mark.parametrize("cls,parts", [((2, 1, 1, 0), 2), ((3, 2, 2, 1), 3), ((2, 2, 2, 2), 4)])
def test_subprograms_cover_the_accumulators(cls, parts):
    """Primitive-pass programs over accumulator subsets give every accumulator once, with
    the value of the full program."""
    from gbasis_amd.programs import class_subprograms

    rng = np.random.default_rng(sum(cls) + parts)
    full = class_program(*cls)
    boys = rng.uniform(0.1, 1.0, full.L + 1)
    sc = (rng.normal(size=3), rng.normal(size=3), 0.3, 0.7, rng.normal(size=3),
          rng.normal(size=3), 0.4, 0.6, 0.25)
    ref = _prim_acc(full, boys, sc)
    got = {}
    subs = class_subprograms(*cls, parts)
    for q in subs:
        part = _prim_acc(q, boys, sc)
        assert not set(part) & set(got)
        got.update(part)
        assert q.c_prim_slots <= full.prim_slots
    assert sorted(got) == list(range(full.nacc))
    for f, v in ref.items():
        assert abs(got[f] - v) <= 1e-12 * max(1.0, abs(v)), f


def test_encoded_program_disk_cache(tmp_path, monkeypatch):
    """The on-disk cache returns exactly the arrays the builder encodes."""
    from gbasis_amd import programs

    monkeypatch.setenv("GBX_PROGRAM_CACHE", str(tmp_path))
    programs._cache_dir.cache_clear()
    try:
        for cls, sub in (((2, 1, 2, 0), None), ((3, 2, 2, 1), (3, 1))):
            first = programs.encoded_program(cls, sub)
            again = programs.encoded_program(cls, sub)     # read back from disk
            p = class_program(*cls) if sub is None else \
                programs.class_subprograms(*cls, sub[0])[sub[1]]
            h, it, ph = p.encode()
            for a in (first, again):
                assert np.array_equal(a[0], h) and np.array_equal(a[1], it)
                assert np.array_equal(a[2], ph) and np.array_equal(a[3], p.c_out_slots)
        assert len(list(tmp_path.rglob("*.npz"))) == 2
    finally:
        monkeypatch.delenv("GBX_PROGRAM_CACHE")
        programs._cache_dir.cache_clear()
