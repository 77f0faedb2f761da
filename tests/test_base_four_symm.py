"""The per-quartet plugin point (gbasis/base_four_symm.py): the host assembly around
a subclass's ``construct_array_contraction``, after the cases of the reference's
tests/test_base_four_symm.py, and the Be+C STO-6G HORTON fixture assembled from
the oracle's quartet blocks through the plugin."""
import os

import numpy as np
import pytest

from gbasis_amd.base_four_symm import BaseFourIndexSymmetric
from gbasis_amd.contractions import GeneralizedContractionShell
from gbasis_amd.spherical import generate_transformation

HERE = os.path.dirname(os.path.abspath(__file__))


def _plugin(fn):
    class Test(BaseFourIndexSymmetric):
        construct_array_contraction = fn
    return Test


def _exps_product(self, c1, c2, c3, c4):
    shape = []
    for c in (c1, c2, c3, c4):
        shape += [c.num_seg_cont, c.num_cart]
    return np.ones(shape) * c1.exps * c2.exps * c3.exps * c4.exps


def test_init_and_contractions():
    cont = GeneralizedContractionShell(1, np.array([1, 2, 3]), np.ones(1), np.ones(1), "spherical")
    Test = _plugin(_exps_product)
    assert Test([cont]).contractions[0] is cont
    with pytest.raises(TypeError):
        Test([cont], [cont])
    with pytest.raises(TypeError):
        Test(cont)
    with pytest.raises(TypeError):
        Test([1, 2])
    with pytest.raises(ValueError):
        Test([])
    with pytest.raises(TypeError):       # abstract
        BaseFourIndexSymmetric([cont])


def test_cartesian_blocks():
    """Two shells (p, d) with block value e1 e2 e3 e4: every one of the 16 shell
    blocks of the dense array holds the product of its shells' exponents."""
    c1 = GeneralizedContractionShell(1, np.array([1, 2, 3]), np.ones((1, 1)), np.ones(1),
                                     "spherical")
    c2 = GeneralizedContractionShell(2, np.array([2, 3, 4]), np.ones((1, 1)), 2 * np.ones(1),
                                     "spherical")
    got = _plugin(_exps_product)([c1, c2]).construct_array_cartesian()
    sl = (slice(0, 3), slice(3, 9))
    e = (1.0, 2.0)
    assert got.shape == (9, 9, 9, 9)
    for a in range(2):
        for b in range(2):
            for c in range(2):
                for d in range(2):
                    assert np.all(got[sl[a], sl[b], sl[c], sl[d]] == e[a] * e[b] * e[c] * e[d])


def test_spherical_and_mix_against_full_transform():
    """Pure output = the Cartesian array with each shell's transform applied on all
    four axes (a block-diagonal matrix built here)."""
    c1 = GeneralizedContractionShell(2, np.array([0.0, 0.5, 1.0]), np.ones((1, 2)),
                                     np.array([0.7]), "spherical")
    c2 = GeneralizedContractionShell(1, np.array([1.0, 0.0, 0.0]), np.ones((1, 1)),
                                     np.array([1.3]), "spherical")
    c1.norm_cont = np.array([[1.0, 2.0, 3.0, 4.0, 5.0, 6.0], [0.5, 1.0, 1.5, 2.0, 2.5, 3.0]])
    c2.norm_cont = np.array([[1.0, 0.5, 0.25]])
    Test = _plugin(_exps_product)
    t = Test([c1, c2])
    cart = t.construct_array_cartesian()
    blocks = []
    for c, sph in ((c1, True), (c2, False)):
        T = (generate_transformation(c.angmom, c.angmom_components_cart, c.angmom_components_sph,
                                     "left") if sph else np.eye(c.num_cart))
        for _ in range(c.num_seg_cont):
            blocks.append(T)
    n_out = sum(b.shape[0] for b in blocks)
    B = np.zeros((n_out, cart.shape[0]))
    r = q = 0
    for b in blocks:
        B[r:r + b.shape[0], q:q + b.shape[1]] = b
        r += b.shape[0]
        q += b.shape[1]
    ref = np.einsum("ai,bj,ck,dl,ijkl->abcd", B, B, B, B, cart)
    np.testing.assert_allclose(t.construct_array_mix(["spherical", "cartesian"]), ref, rtol=1e-13)
    with pytest.raises(TypeError):
        t.construct_array_mix("spherical")
    with pytest.raises(ValueError):
        t.construct_array_mix(["spherical"])
    with pytest.raises(ValueError):
        t.construct_array_mix(["spherical", "bad"])


def test_lincomb_keeps_the_last_permutation():
    """One shell and a non-symmetric block: the last of the eight writes wins, so the
    array is the block read ijkl->lkji (reference test_construct_array_lincomb)."""
    cont = GeneralizedContractionShell(1, np.array([1, 2, 3]), np.ones(1), np.ones(1),
                                       "spherical")
    cont.norm_cont = np.ones((1, 3))
    Test = _plugin(lambda self, c1, c2, c3, c4, a=2:
                   np.arange(81, dtype=float).reshape(1, 3, 1, 3, 1, 3, 1, 3) * a)
    t = Test([cont])
    orb = np.random.default_rng(3).random((3, 3))
    base = np.einsum("ijkl->lkji", np.arange(81.0).reshape(3, 3, 3, 3)) * 2
    np.testing.assert_allclose(t.construct_array_lincomb(orb, ["cartesian"]),
                               np.einsum("ijkl,ai,bj,ck,dl->abcd", base, orb, orb, orb, orb))
    S = generate_transformation(1, cont.angmom_components_cart, cont.angmom_components_sph, "left")
    sph = np.einsum("ijkl,ai,bj,ck,dl->abcd", base, S, S, S, S)
    np.testing.assert_allclose(t.construct_array_lincomb(orb, ["spherical"]),
                               np.einsum("ijkl,ai,bj,ck,dl->abcd", sph, orb, orb, orb, orb))
    np.testing.assert_allclose(t.construct_array_cartesian(a=5), base / 2 * 5)
    with pytest.raises(TypeError):
        t.construct_array_lincomb(orb, "spherical", bad_keyword=3)
    with pytest.raises(TypeError):
        t.construct_array_lincomb(orb, "bad", keyword=3)


def test_mix_reads_no_pure_convention_of_cartesian_shells():
    class SpecialShell(GeneralizedContractionShell):
        @property
        def angmom_components_sph(self):
            raise NotImplementedError

    c = SpecialShell(1, np.array([1, 2, 3]), np.ones((1, 2)), np.ones(1), "spherical")
    Test = _plugin(lambda self, c1, c2, c3, c4, a=2:
                   np.arange(6.0 ** 4).reshape(2, 3, 2, 3, 2, 3, 2, 3) * a)
    t = Test([c, c])
    np.testing.assert_allclose(t.construct_array_cartesian(a=3),
                               t.construct_array_mix(["cartesian"] * 2, a=3))


def test_trailing_axes_follow_the_block():
    c = GeneralizedContractionShell(0, np.zeros(3), np.ones((1, 1)), np.ones(1), "cartesian")
    Test = _plugin(lambda self, c1, c2, c3, c4: np.ones((1, 1, 1, 1, 1, 1, 1, 1, 2)) *
                   np.array([1.0, -1.0]))
    got = Test([c, c]).construct_array_cartesian()
    assert got.shape == (2, 2, 2, 2, 2)
    assert np.all(got[..., 0] == 1.0) and np.all(got[..., 1] == -1.0)


def test_horton_bec_through_the_plugin():
    """Be+C STO-6G (HORTON data_horton_bec_cart_elec_repulsion.npy, physicist) from the
    oracle's quartet blocks assembled by the plugin point, and the ERI class
    routing a subclass's own ``construct_array_contraction`` to this assembly."""
    from gbasis_amd.electron_repulsion import ElectronRepulsionIntegral
    from gbasis_amd.parsers import make_contractions, parse_nwchem
    from oracle.assembly import contraction_block

    d = parse_nwchem(os.path.join(HERE, "golden", "sto6g_be_c.nwchem"))
    basis = make_contractions(d, ["Be", "C"], np.array([[0, 0, 0], [1.0, 0, 0]]), "cartesian")
    ref = np.load(os.path.join(HERE, "golden", "horton_bec_cart_elec_repulsion.npy"))

    class OracleEri(ElectronRepulsionIntegral):
        def construct_array_contraction(self, c1, c2, c3, c4, **kwargs):
            return contraction_block(c1, c2, c3, c4)

    got = OracleEri(basis).construct_array_cartesian()
    np.testing.assert_allclose(np.transpose(got, (0, 2, 1, 3)), ref, rtol=1e-10, atol=1e-15)
    assert OracleEri(basis)._plugin() and not ElectronRepulsionIntegral(basis)._plugin()
