"""Parity with the reference's own outputs (tests/golden/reference_vectors.npz, written
by tests/golden/make_reference_vectors.py from gbasis itself).

What the vectors show:
  * ``norm_cont`` of the generated shells equals the reference's;
  * the Schwarz bound matrices Q and hence the screened quartet lists at 1e-12
    (gbasis/integrals/_schwarz_screening.py:62-90,169-191) -- the lists are
    identical, entry for entry;
  * the reference's FP64 ERIs on the dense s,s,p,d workload (N = 44).  Against
    the extended-precision oracle, over the 4000 sampled entries, the
    reference's own FP64 error reaches 5.34e-8 Eh in (dd|dd) quartets, where
    its electron-transfer step loses digits, and 6.2e-9 elsewhere.  The device
    path is held to the reference within that error (its largest deviation
    from the reference is that same 5.34e-8 entry), and to the oracle at 1e-12
    by the other tests.
"""
import os

import numpy as np
import pytest

from gbasis_amd.workloads import batch_molecule, dense_small

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"))
THRESHOLD = 1e-12
REF_FP64_ERR = 6e-8      # the reference's measured FP64 error at the samples: 5.34e-8
# bars for Q: the reference's FP64 Q against the extended-precision oracle differs by
# at most 1.1e-8 (dense_small), 1.4e-7 (batch1000) and 4.0e-7 (batch1001, exponents
# up to 1e3), measured when the vectors were written
Q_TOL = {"dense_small": 5e-8, "batch1000": 5e-7, "batch1001": 1.2e-6}


def _screened(Q):
    """Kept (ij, kl) pairs, ij <= kl over i <= j shell pairs: Q_ij Q_kl >= threshold."""
    n = Q.shape[0]
    i, j = np.triu_indices(n)
    q = Q[i, j]
    P = np.outer(q, q)
    return np.triu(P >= THRESHOLD)


def _molecules():
    return {"dense_small": dense_small(1), "batch1000": batch_molecule(1000),
            "batch1001": batch_molecule(1001)}


def test_norm_cont_equals_reference():
    nc = np.concatenate([np.asarray(s.norm_cont).ravel() for s in dense_small(1)])
    np.testing.assert_allclose(nc, VEC["dense_small_norm_cont"], rtol=1e-14)


@pytest.mark.parametrize("name", ["dense_small", "batch1000"])
def test_oracle_schwarz_and_screened_list(name):
    from oracle.assembly import schwarz_bounds

    Q = schwarz_bounds(_molecules()[name])
    Qr = VEC[f"{name}_Q"]
    assert np.abs(Q - Qr).max() <= Q_TOL[name]
    kept, kept_ref = _screened(Q), _screened(Qr)
    assert kept.sum() > 0 and np.array_equal(kept, kept_ref)


def test_oracle_dense_samples():
    """The extended-precision oracle at sampled entries: within the reference's FP64
    error, which is largest in (dd|dd) (5.34e-8) and below 1e-8 elsewhere."""
    from oracle.assembly import basis_offsets, normalized_block

    basis = dense_small(1)
    ct = [s.coord_type for s in basis]
    off = basis_offsets(basis, ct)
    shell_of = np.searchsorted(off, np.arange(off[-1]), side="right") - 1
    idx = VEC["dense_small_idx"].astype(np.int64)
    val = VEC["dense_small_val"]
    blocks = {}
    worst = {"dddd": 0.0, "other": 0.0}
    for r in range(0, 4000, 50):                 # 80 entries
        e = idx[r]
        s = tuple(int(x) for x in shell_of[e])
        if s not in blocks:
            blocks[s] = normalized_block(basis, ct, *s, exact=True)
        v = blocks[s][tuple(e - off[list(s)])]
        kind = "dddd" if all(basis[x].angmom == 2 for x in s) else "other"
        worst[kind] = max(worst[kind], abs(v - val[r]))
    assert worst["other"] <= 1e-8, worst
    assert worst["dddd"] <= REF_FP64_ERR, worst


@pytest.mark.gpu
def test_device_vs_reference_outputs(gpu):
    """Device Q and screened lists identical to the reference's; the device dense ERIs
    at 4000 sampled entries and in aggregate within the reference's FP64 error."""
    from gbasis_amd.electron_repulsion import electron_repulsion_integral_improved
    from gbasis_amd.schwarz import compute_schwarz_bounds

    print("[gpu case] reference vectors: Schwarz matrices and dense s,s,p,d", flush=True)
    for name, basis in _molecules().items():
        Q = compute_schwarz_bounds(basis)
        Qr = VEC[f"{name}_Q"]
        assert np.abs(Q - Qr).max() <= Q_TOL[name], name
        assert np.array_equal(_screened(Q), _screened(Qr)), name
    eri = electron_repulsion_integral_improved(dense_small(1), notation="chemist")
    assert eri.shape[0] == int(VEC["dense_small_n"])
    idx = VEC["dense_small_idx"].astype(np.int64)
    got = eri[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]
    assert np.abs(got - VEC["dense_small_val"]).max() <= REF_FP64_ERR
    # aggregates over all 3.75e6 entries: the reference's per-entry FP64 error summed
    # (the oracle's FP64 restatement differs from them by 1.1e-9 of sum|x|)
    sumabs = float(VEC["dense_small_sumabs"])
    assert abs(eri.sum() - float(VEC["dense_small_sum"])) <= 1e-8 * sumabs
    assert abs(np.abs(eri).sum() / sumabs - 1) <= 1e-8
    assert abs((eri * eri).sum() / float(VEC["dense_small_sumsq"]) - 1) <= 1e-8
