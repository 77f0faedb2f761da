"""Host tables: shell expansion, pairs, canonical quartet enumeration and the
packed-slot placement rule the kernels apply (CPU)."""
import numpy as np
import pytest

from gbasis_amd.contractions import GeneralizedContractionShell
from gbasis_amd.tables import BatchTables, MoleculeTables
from tests.molecules import random_basis


def _tri(i, j):
    return i * (i + 1) // 2 + j if i >= j else j * (j + 1) // 2 + i


def _slot_writes(b, bra, ket):
    """Replay the kernels' placement rule; count writes per packed slot."""
    counts = np.zeros(b.n_slots, dtype=np.int64)
    info = b.shell_info
    for p, q in zip(bra.tolist(), ket.tolist()):
        pb, pk = b.pair_info[p], b.pair_info[q]
        sh = [pb[0], pb[1], pk[0], pk[1]]
        funcs = []
        for s in sh:
            off, kind, moff, nout = (int(x) for x in info[s, [0, 2, 3, 5]])
            if kind == 0:
                ncart = (int(info[s, 1]) + 1) * (int(info[s, 1]) + 2) // 2
                funcs.append([off + int(b.comp_map[moff + i]) for i in range(ncart)])
            else:
                funcs.append([off + i for i in range(nout)])
        for fa in funcs[0]:
            for fb in funcs[1]:
                if pb[0] == pb[1] and fa < fb:
                    continue
                for fc in funcs[2]:
                    for fd in funcs[3]:
                        if pk[0] == pk[1] and fc < fd:
                            continue
                        ij, kl = _tri(fa, fb), _tri(fc, fd)
                        if p == q and ij < kl:
                            continue
                        counts[int(b.pair_out_base[p]) + _tri(ij, kl)] += 1
    return counts


@pytest.mark.parametrize("ls,ct", [([0, 1, 2, 1], "cartesian"), ([2, 0, 3], "spherical"),
                                   ([1, 1, 0], "spherical")])
def test_every_slot_written_once(ls, ct):
    basis = random_basis(3, ls, kmax=2, coord_type=ct)
    m = MoleculeTables(basis)
    b = BatchTables([m])
    bra, ket = m.quartets()
    counts = _slot_writes(b, bra, ket)
    assert np.all(counts == 1)
    assert int(b.unique_counts(bra.astype(np.int64), ket.astype(np.int64)).sum()) == b.n_slots


def test_generalized_contraction_expansion():
    """M columns become M consecutive segmented shells (m-major function order)."""
    rng = np.random.default_rng(0)
    sh = GeneralizedContractionShell(1, np.zeros(3), rng.uniform(0.1, 1, (3, 2)),
                                     np.array([0.5, 1.5, 4.0]), "cartesian")
    m = MoleculeTables([sh])
    assert m.nshell == 2 and m.nfunc == 6
    assert [s.func_off for s in m.shells] == [0, 3]
    np.testing.assert_allclose(m.shells[1].comp_scale, sh.norm_cont[1])


def test_canonical_orientation_and_buckets():
    basis = random_basis(9, [0, 2, 1, 1], kmax=3)
    m = MoleculeTables(basis)
    b = BatchTables([m])
    bra, ket = m.quartets()
    assert np.all(m.pair_la >= m.pair_lb)
    lab_b = m.pair_la[bra] + m.pair_lb[bra]
    lab_k = m.pair_la[ket] + m.pair_lb[ket]
    assert np.all((lab_b > lab_k) | ((lab_b == lab_k) & (m.pair_la[bra] >= m.pair_la[ket])))
    total = 0
    for cls, npq, dense, quart in b.class_buckets(bra.astype(np.int64), ket.astype(np.int64)):
        assert not dense
        assert np.all(m.pair_la[quart[:, 0]] == cls[0]) and np.all(m.pair_lb[quart[:, 1]] == cls[3])
        total += quart.shape[0]
    assert total == bra.size


def test_norm_cont_closed_form_is_unit_norm():
    """norm_cont normalises each contracted Cartesian function (closed-form overlap)."""
    rng = np.random.default_rng(1)
    for l in range(4):
        exps = np.exp(rng.uniform(-2, 3, 4))
        sh = GeneralizedContractionShell(l, np.zeros(3), rng.uniform(0.1, 1, 4), exps, "cartesian")
        # numerical self-overlap of the contracted x^l function on a radial grid
        from scipy.integrate import quad
        N = sh.norm_prim_cart[0]  # component (l,0,0)
        c = sh.coeffs[:, 0] * N * sh.norm_cont[0, 0]

        def f(r, axis_pow):
            return np.sum(c * np.exp(-exps * r * r))

        # separable: integral of (sum_p c_p x^l e^{-a r^2})^2 over R^3
        tot = 0.0
        for p in range(4):
            for q in range(4):
                a = exps[p] + exps[q]
                ix = quad(lambda x: x ** (2 * l) * np.exp(-a * x * x), -np.inf, np.inf)[0]
                iy = np.sqrt(np.pi / a)
                tot += c[p] * c[q] * ix * iy * iy
        assert abs(tot - 1.0) < 1e-10


def test_bucket_capacity_deep_contractions():
    """Shells with K = 14 primitives (cc-pVTZ S/Cl, cc-pV5Z C s shells): 14^4 = 38416
    primitive quartets exceed the 4-bit primitive-count bin, which must stay in the last
    bin and never carry into the precision or output-transform bits of the key.  The
    capacities cover every unscreened candidate exactly once."""
    from gbasis_amd.tables import MAX_PRIM_BIN, prim_bin

    assert prim_bin(np.array([1, 2, 3, 16384, 16385, 38416, 2 ** 30])).tolist() == \
        [0, 1, 2, 14, 15, 15, 15]
    basis = random_basis(5, [0, 0, 1, 0], kmin=14, kmax=14, box=4.0)
    m = MoleculeTables(basis)
    b = BatchTables([m])
    cap = b.bucket_capacity()
    keys = np.flatnonzero(cap)
    assert int(cap[:int(cap.size) // 4].sum()) == b.n_cand   # phase 0: one per candidate
    from gbasis_amd.tables import decode_head

    heads = [decode_head(int(k) >> 4) for k in keys]
    assert np.all((keys & 15) == MAX_PRIM_BIN)
    assert not any(fp32 or dense for _, _, dense, fp32 in heads)
    got = {cls for phase, cls, _, _ in heads if phase == 0}
    assert got == {(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 0),
                   (1, 1, 1, 1)}
