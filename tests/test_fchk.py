"""Gaussian fchk input: the reader and shell builder on CPU, and the ethylene
HF/cc-pVDZ Coulomb and exchange energies on the GPU.

The fixture ``golden/C2H4_hf_ccpvdz.fchk`` is the data file the reference's
notebooks load (notebooks/tutorial/C2H4_hf_ccpvdz.fchk).  The published values
of notebooks/tutorial/JCP_paper.ipynb (cells 28 and 33) are the known answers:
J = sum_ij 2 (ii|jj), K = sum_ij (ij|ij) over the eight occupied MOs,
70.53719756188356 / 11.756060443132313 Eh from gbasis and
70.5371975684723 / 11.756060454118147 Eh from libcint.
"""
import os

import numpy as np
import pytest

from gbasis_amd.parsers import fchk_basis, fchk_rhf_orbitals, read_fchk
from gbasis_amd.programs import comps
from gbasis_amd.tables import expand_shells, f2norm_table, prim_norm

HERE = os.path.dirname(os.path.abspath(__file__))
FCHK = os.path.join(HERE, "golden", "C2H4_hf_ccpvdz.fchk")

J_GBASIS, K_GBASIS = 70.53719756188356, 11.756060443132313
J_LIBCINT, K_LIBCINT = 70.5371975684723, 11.756060454118147


def _overlap_1d(la, lb, a, b, A, B):
    """Obara-Saika overlap s[i, j] of x^i e^{-a(x-A)^2} and x^j e^{-b(x-B)^2}
    (Cartesian powers about A and B)."""
    p = a + b
    P = (a * A + b * B) / p
    s = np.zeros((la + 2, lb + 2))
    s[0, 0] = np.sqrt(np.pi / p) * np.exp(-a * b / p * (A - B) ** 2)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:
                s[i, j] = (P - A) * s[i - 1, j] + ((i - 1) * (s[i - 2, j] if i > 1 else 0.0)
                                                  + j * (s[i - 1, j - 1] if j > 0 else 0.0)) / (2 * p)
            else:
                s[i, j] = (P - B) * s[i, j - 1] + (j - 1) * (s[i, j - 2] if j > 1 else 0.0) / (2 * p)
    return s


def _out_matrix(es):
    """(nout, ncart) map from canonical Cartesian components to the shell's outputs."""
    ncart = (es.l + 1) * (es.l + 2) // 2
    if es.kind == 1:
        return es.out_mat
    t = np.zeros((es.nout, ncart))
    t[es.comp_map, np.arange(ncart)] = es.comp_scale
    return t


def _overlap_matrix(basis):
    shells, nfunc = expand_shells(basis, [s.coord_type for s in basis])
    f2 = f2norm_table()
    f2off = np.cumsum([0] + [(l + 1) * (l + 2) // 2 for l in range(9)])
    S = np.zeros((nfunc, nfunc))
    for si in shells:
        for sj in shells:
            ci, cj = np.array(comps(si.l)), np.array(comps(sj.l))
            blk = np.zeros((len(ci), len(cj)))
            ni, nj = prim_norm(si.exps, si.l), prim_norm(sj.exps, sj.l)
            for p, a in enumerate(si.exps):
                for q, b in enumerate(sj.exps):
                    w = si.coeffs[p] * sj.coeffs[q] * ni[p] * nj[q]
                    s1 = [_overlap_1d(si.l, sj.l, a, b, si.coord[d], sj.coord[d]) for d in range(3)]
                    blk += w * (s1[0][np.ix_(ci[:, 0], cj[:, 0])] * s1[1][np.ix_(ci[:, 1], cj[:, 1])]
                                * s1[2][np.ix_(ci[:, 2], cj[:, 2])])
            blk *= f2[f2off[si.l]:f2off[si.l + 1]][:, None] * f2[f2off[sj.l]:f2off[sj.l + 1]][None, :]
            S[si.func_off:si.func_off + si.nout, sj.func_off:sj.func_off + sj.nout] = \
                _out_matrix(si) @ blk @ _out_matrix(sj).T
    return S


def test_fchk_reader_ethylene():
    f = read_fchk(FCHK)
    assert f["Number of basis functions"] == 48
    assert f["Number of alpha electrons"] == 8
    assert f["Shell types"].tolist()[:6] == [0, 0, 0, 1, 1, -2]
    basis = fchk_basis(f)
    # JCP_paper.ipynb cell 11: 24 shells, 48 functions
    assert len(basis) == 24
    n = sum(2 * s.angmom + 1 if s.coord_type == "spherical" else (s.angmom + 1) * (s.angmom + 2) // 2
            for s in basis)
    assert n == 48
    assert [s.angmom for s in basis if s.coord_type == "spherical"] == [2, 2]


def test_fchk_orbitals_orthonormal():
    """C^T S C = 1 with S from the shells as built: pins the component orders, the
    Cartesian->pure transform and the contraction normalisation of the reader."""
    f = read_fchk(FCHK)
    basis = fchk_basis(f)
    c, occ = fchk_rhf_orbitals(f)
    S = _overlap_matrix(basis)
    assert np.allclose(S, S.T, atol=1e-12)
    err = np.abs(c.T @ S @ c - np.eye(c.shape[1])).max()
    assert err < 1e-6, err
    assert occ.sum() == 16


def test_overlap_helper_closed_form():
    """The test's own overlap against closed forms: two normalised s functions and
    the p-p self overlap."""
    a, b, R = 0.8, 1.3, 0.7
    s = _overlap_1d(1, 1, a, b, 0.0, R)
    ref = np.sqrt(np.pi / (a + b)) * np.exp(-a * b / (a + b) * R * R)
    assert abs(s[0, 0] - ref) < 1e-15
    # <x|x> about one centre: sqrt(pi/p) / (2p)
    s = _overlap_1d(1, 1, a, b, 0.0, 0.0)
    assert abs(s[1, 1] - np.sqrt(np.pi / (a + b)) / (2 * (a + b))) < 1e-15
    # <x^2|1> about one centre equals <x|x>
    s2 = _overlap_1d(2, 0, a, b, 0.0, 0.0)
    assert abs(s2[2, 0] - s[1, 1]) < 1e-15


@pytest.mark.gpu
def test_fchk_ethylene_coulomb_exchange(gpu):
    from gbasis_amd.electron_repulsion import electron_repulsion_integral

    print("[gpu case] fchk ethylene J/K", flush=True)
    f = read_fchk(FCHK)
    basis = fchk_basis(f)
    c, occ = fchk_rhf_orbitals(f)
    mo = electron_repulsion_integral(basis, transform=c.T, notation="chemist")
    no = int((occ > 0).sum())
    ii = np.arange(no)
    j_coul = 2.0 * mo[ii[:, None], ii[:, None], ii[None, :], ii[None, :]].sum()
    k_ex = mo[ii[:, None], ii[None, :], ii[:, None], ii[None, :]].sum()
    print(f"J {j_coul:.14f} K {k_ex:.14f}", flush=True)
    # the two published values differ by 5e-9 (J) and 1.1e-8 (K)
    assert abs(j_coul - J_GBASIS) < 2e-8 and abs(j_coul - J_LIBCINT) < 2e-8
    assert abs(k_ex - K_GBASIS) < 2e-8 and abs(k_ex - K_LIBCINT) < 2e-8


def test_fchk_unrestricted_pure_f_orthonormal():
    """Water (uwB97XD/def2-TZVPD, notebooks/tutorial/water.fchk): pure d and f shells;
    alpha and beta orbitals orthonormal under the shells as built, which pins the pure
    component order c0, c1, s1, c2, s2, c3, s3 and the f normalisation."""
    from gbasis_amd.parsers import fchk_orbitals

    f = read_fchk(os.path.join(HERE, "golden", "water.fchk"))
    basis = fchk_basis(f)
    assert sorted({(s.angmom, s.coord_type) for s in basis}) == [
        (0, "cartesian"), (1, "cartesian"), (2, "spherical"), (3, "spherical")]
    S = _overlap_matrix(basis)
    for spin in ("alpha", "beta"):
        c, occ = fchk_orbitals(f, spin)
        assert c.shape == (58, 58) and occ.sum() == 5
        assert np.abs(c.T @ S @ c - np.eye(58)).max() < 1e-6
