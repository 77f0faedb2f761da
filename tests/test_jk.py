"""Coulomb / exchange digestion from packed ERIs.

CPU: argument validation of the C ABI and the packed-index helper used below.
GPU: ``gbx_jk_packed`` against the definitions J_mn = sum_ls (mn|ls) D_ls and
K_ml = sum_ns (mn|ls) D_ns evaluated with einsum on the dense array (random
8-fold-symmetric arrays, so every degeneracy case i = j, k = l, ij = kl occurs),
J/K of a real basis against the oracle's dense ERIs, and the ethylene
HF/cc-pVDZ energies against the values published in
notebooks/tutorial/JCP_paper.ipynb (cells 28 and 33).
"""
import os

import numpy as np
import pytest

from gbasis_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
FCHK = os.path.join(HERE, "golden", "C2H4_hf_ccpvdz.fchk")


def _packed_from_dense(eri):
    """Canonical packed slots (ij(ij+1)/2 + kl) of an 8-fold symmetric chemists' array."""
    n = eri.shape[0]
    i, j = np.tril_indices(n)
    p_i, p_j = i, j
    P = p_i.size
    a, b = np.tril_indices(P)
    return eri[p_i[a], p_j[a], p_i[b], p_j[b]]


def _tri(a, b):
    hi, lo = np.maximum(a, b), np.minimum(a, b)
    return hi * (hi + 1) // 2 + lo


def _random_symmetric_eri(n, seed):
    """Dense 8-fold symmetric array filled from random canonical slots."""
    rng = np.random.default_rng(seed)
    npair = n * (n + 1) // 2
    packed = rng.standard_normal(npair * (npair + 1) // 2)
    idx = np.arange(n)
    kl = _tri(idx[:, None], idx[None, :])                      # (n, n)
    out = np.empty((n, n, n, n))
    for m in range(n):
        ij = _tri(m, idx)                                      # (n,)
        out[m] = packed[_tri(ij[:, None, None], kl[None, :, :])]
    return out


def test_packed_helper_order():
    n = 4
    eri = _random_symmetric_eri(n, 0)
    packed = _packed_from_dense(eri)
    # slot of (ij|kl) with i >= j, k >= l, ij >= kl
    for (i, j, k, l) in [(3, 1, 2, 2), (1, 0, 0, 0), (3, 3, 3, 3), (2, 1, 2, 0)]:
        ij, kl = i * (i + 1) // 2 + j, k * (k + 1) // 2 + l
        assert packed[ij * (ij + 1) // 2 + kl] == eri[i, j, k, l]


def test_jk_argument_errors_without_device():
    lib = _lib.load()
    assert lib.gbx_jk_packed(None, 4, None, None, None, None) == -1
    assert b"null" in lib.gbx_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 5, 23, 90, 110])
def test_jk_random_packed_vs_einsum(gpu, n):
    """n = 90 needs more than 64 KiB of LDS-private accumulators (the raised LDS
    limit); n = 110 exceeds them (n <= 101 on gfx950) and takes the global-atomic
    kernel."""
    import torch
    from gbasis_amd.jk import coulomb_exchange_packed

    print(f"[gpu case] jk random n={n}", flush=True)
    eri = _random_symmetric_eri(n, n)
    rng = np.random.default_rng(100 + n)
    D = rng.standard_normal((n, n))
    D = D + D.T
    packed = torch.as_tensor(_packed_from_dense(eri), device=gpu)
    J, K = coulomb_exchange_packed(packed, n, D)
    J, K = J.cpu().numpy(), K.cpu().numpy()
    Jr = np.einsum("mnls,ls->mn", eri, D)
    Kr = np.einsum("mnls,ns->ml", eri, D)
    scale = np.abs(eri).max() * np.abs(D).sum()
    assert np.abs(J - Jr).max() <= 1e-13 * scale
    assert np.abs(K - Kr).max() <= 1e-13 * scale
    assert np.array_equal(J, J.T) and np.array_equal(K, K.T)


@pytest.mark.gpu
def test_jk_basis_vs_oracle(gpu):
    from gbasis_amd.jk import coulomb_exchange
    from oracle.assembly import dense_eri
    from tests.molecules import random_basis

    print("[gpu case] jk basis vs oracle", flush=True)
    basis = random_basis(3, [0, 1, 2, 1, 0])
    ref = dense_eri(basis, notation="chemist")
    n = ref.shape[0]
    rng = np.random.default_rng(5)
    C = rng.standard_normal((n, 3))
    D = 2.0 * C @ C.T
    J, K = coulomb_exchange(basis, D)
    assert np.abs(J - np.einsum("mnls,ls->mn", ref, D)).max() <= 1e-11
    assert np.abs(K - np.einsum("mnls,ns->ml", ref, D)).max() <= 1e-11


@pytest.mark.gpu
def test_jk_ethylene_energies(gpu):
    from gbasis_amd.jk import coulomb_exchange, coulomb_exchange_energies
    from gbasis_amd.parsers import fchk_basis, fchk_rhf_orbitals, read_fchk

    print("[gpu case] jk ethylene", flush=True)
    f = read_fchk(FCHK)
    basis = fchk_basis(f)
    c, occ = fchk_rhf_orbitals(f)
    D = (c * occ) @ c.T
    J, K = coulomb_exchange(basis, D)
    ej, ek = coulomb_exchange_energies(J, K, D)
    print(f"E_J {ej:.14f} E_K {ek:.14f}", flush=True)
    # gbasis: 70.53719756188356 / 11.756060443132313; libcint: 70.5371975684723 / 11.756060454118147
    assert abs(ej - 70.53719756188356) < 2e-8 and abs(ej - 70.5371975684723) < 2e-8
    assert abs(ek - 11.756060443132313) < 2e-8 and abs(ek - 11.756060454118147) < 2e-8
