"""Four-index transform of packed ERIs into a new basis (``gbx_mo_transform``).

CPU: argument validation and the workspace size of the C ABI.
GPU: the two-pass transform against einsum of the dense array with the transform,
(ab|cd) = sum_ijkl T_ai T_bj T_ck T_dl (ij|kl) (gbasis/base_four_symm.py:611-615), on
random 8-fold-symmetric arrays (square, narrowing and widening transforms, sizes that
cover several LDS chunks and row tiles, both notations) and on a real basis against
the oracle's dense ERIs.
"""
import numpy as np
import pytest

from gbasis_amd import _lib

from tests.test_jk import _packed_from_dense, _random_symmetric_eri


def test_mo_transform_argument_errors_without_device():
    import ctypes

    lib = _lib.load()
    assert lib.gbx_mo_transform(None, 4, None, None, 4, None, None, 0, None) == -1
    assert b"null" in lib.gbx_last_error()
    n = ctypes.c_int64()
    assert lib.gbx_mo_workspace(5, 3, ctypes.byref(n)) == 0
    assert n.value == 15 * 6
    assert lib.gbx_mo_workspace(0, 3, ctypes.byref(n)) == -1


def _reference(eri, T, notation):
    out = np.einsum("ai,bj,ck,dl,ijkl->abcd", T, T, T, T, eri, optimize=True)
    return out.transpose(0, 2, 1, 3) if notation == "physicist" else out


@pytest.mark.gpu
@pytest.mark.parametrize("n,m", [(1, 1), (6, 4), (20, 20), (37, 40)])
@pytest.mark.parametrize("notation", ["chemist", "physicist"])
def test_mo_transform_random_vs_einsum(gpu, n, m, notation):
    import torch

    from gbasis_amd.engine import get_engine

    print(f"[gpu case] mo transform random n={n} m={m} {notation}", flush=True)
    eri = _random_symmetric_eri(n, 7 * n + m)
    T = np.random.default_rng(n + 3 * m).standard_normal((m, n))
    packed = torch.as_tensor(_packed_from_dense(eri), device=gpu)
    got = get_engine("cuda").mo_transform(packed, n, T, notation).cpu().numpy()
    ref = _reference(eri, T, notation)
    scale = np.abs(eri).max() * np.abs(T).sum(axis=1).max() ** 4
    assert got.shape == (m,) * 4
    assert np.abs(got - ref).max() <= 1e-14 * scale


@pytest.mark.gpu
def test_mo_transform_basis_vs_oracle(gpu):
    from gbasis_amd.electron_repulsion import electron_repulsion_integral
    from oracle.assembly import dense_eri
    from tests.molecules import random_basis

    print("[gpu case] mo transform basis vs oracle", flush=True)
    basis = random_basis(4, [0, 1, 2, 1])
    ref = dense_eri(basis, notation="chemist")
    n = ref.shape[0]
    T = np.random.default_rng(11).standard_normal((n - 2, n)) / np.sqrt(n)
    for notation in ("chemist", "physicist"):
        got = electron_repulsion_integral(basis, transform=T, notation=notation)
        assert np.abs(got - _reference(ref, T, notation)).max() <= 1e-11
