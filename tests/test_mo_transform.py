"""Four-index transform of packed ERIs into a new basis (``gbx_mo_transform``).

CPU: argument validation and the workspace size of the C ABI.
GPU: the two-pass transform against einsum of the dense array with the transform,
(ab|cd) = sum_ijkl T_ai T_bj T_ck T_dl (ij|kl) (gbasis/base_four_symm.py:611-615), on
random 8-fold-symmetric arrays (square, narrowing and widening transforms, sizes that
cover several K steps and row and column tiles of the GEMMs, both notations), on a
real basis against the oracle's dense ERIs, and at a size whose first pass runs in two
chunks against FP64 tensordots of the unpacked array on the device.
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
    # the half-transformed array (15 x 6 pairs), then the larger chunk of GEMM staging:
    # all 15 AO-pair matrices of the first pass at 5 x 3 doubles each
    assert n.value == 15 * 6 + 15 * 5 * 3
    assert lib.gbx_mo_workspace(400, 400, ctypes.byref(n)) == 0
    assert n.value == (400 * 401 // 2) ** 2 + (2 ** 27 // 160000) * 160000
    assert lib.gbx_mo_workspace(0, 3, ctypes.byref(n)) == -1


def _reference(eri, T, notation):
    out = np.einsum("ai,bj,ck,dl,ijkl->abcd", T, T, T, T, eri, optimize=True)
    return out.transpose(0, 2, 1, 3) if notation == "physicist" else out


@pytest.mark.gpu
@pytest.mark.parametrize("n,m", [(1, 1), (6, 4), (20, 20), (37, 40), (66, 70)])
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


@pytest.mark.gpu
def test_mo_transform_chunked_vs_tensordot(gpu):
    """n = m = 130: the first pass's 8515 matrices of 130 x 130 outnumber one chunk of
    GEMM staging (2^27 doubles), so the chunk loop runs twice."""
    import torch

    from gbasis_amd.engine import get_engine

    print("[gpu case] mo transform chunked n=m=130", flush=True)
    n = m = 130
    eng = get_engine("cuda")
    npair = n * (n + 1) // 2
    g = torch.Generator(device=gpu).manual_seed(5)
    # any packed vector is the canonical part of an 8-fold-symmetric array
    packed = torch.rand(npair * (npair + 1) // 2, dtype=torch.float64, device=gpu,
                        generator=g) - 0.5
    T = np.random.default_rng(9).standard_normal((m, n)) / np.sqrt(n)
    got = eng.mo_transform(packed, n, T, "chemist")
    dense = eng.unpack_dense(packed, n, "chemist")
    Td = torch.as_tensor(T, device=gpu)
    ref = dense
    for _ in range(4):  # contract the leading AO index, rotating it to the back
        ref = torch.tensordot(ref, Td, dims=([0], [1]))
    del dense
    assert torch.isfinite(got).all()
    # |T| row sums ~ sqrt(n) * 0.8: a bound on the magnitude of each output element
    scale = 0.5 * float(np.abs(T).sum(axis=1).max()) ** 4
    assert float((got - ref).abs().max()) <= 1e-13 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("notation", ["chemist", "physicist"])
def test_mo_transform_wide_sampled(gpu, notation):
    """m = 200 > 192 new functions: two column tiles of the GEMMs.  The 200^4 output is
    checked at 3000 random positions against the quadruple sum over the 5^4 array."""
    import torch

    from gbasis_amd.engine import get_engine

    print(f"[gpu case] mo transform wide n=5 m=200 {notation}", flush=True)
    n, m = 5, 200
    eri = _random_symmetric_eri(n, 17)
    T = np.random.default_rng(21).standard_normal((m, n))
    packed = torch.as_tensor(_packed_from_dense(eri), device=gpu)
    got = get_engine("cuda").mo_transform(packed, n, T, notation)
    idx = np.random.default_rng(4).integers(0, m, (3000, 4))
    vals = got[tuple(torch.as_tensor(idx[:, j], device=gpu) for j in range(4))].cpu().numpy()
    del got
    if notation == "physicist":  # <ac|bd> at [a][c][b][d] = (ab|cd)
        idx = idx[:, [0, 2, 1, 3]]
    ref = np.einsum("pi,pj,pk,pl,ijkl->p", T[idx[:, 0]], T[idx[:, 1]], T[idx[:, 2]],
                    T[idx[:, 3]], eri, optimize=True)
    scale = np.abs(eri).max() * np.abs(T).sum(axis=1).max() ** 4
    assert np.abs(vals - ref).max() <= 1e-14 * scale
