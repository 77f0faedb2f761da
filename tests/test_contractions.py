"""The shell data model (gbasis/contractions.py) that the ERI path reads: setters and
their errors, the default component orders, counts, and the primitive and
contraction normalisation, against the known answers of the reference's
tests/test_contractions.py and closed-form overlaps."""
import numpy as np
import pytest

from gbasis_amd.contractions import GeneralizedContractionShell


def _shell(l=1, coeffs=None, exps=None, coord=None):
    return GeneralizedContractionShell(
        l, np.zeros(3) if coord is None else coord,
        np.array([1.0]) if coeffs is None else coeffs,
        np.array([0.25]) if exps is None else exps, "spherical")


def test_setters():
    s = _shell(coord=np.array([1, 2, 3]))
    assert s.coord.dtype == float and np.all(s.coord == [1.0, 2.0, 3.0])
    for bad in ([0.0, 0.0, 0.0], np.zeros(2), np.array(["a", "b", "c"])):
        with pytest.raises(TypeError):
            _shell(coord=bad)
    with pytest.raises(TypeError):
        GeneralizedContractionShell(1.0, np.zeros(3), np.ones(1), np.ones(1), "cartesian")
    with pytest.raises(ValueError):
        GeneralizedContractionShell(-1, np.zeros(3), np.ones(1), np.ones(1), "cartesian")
    with pytest.raises(TypeError):
        _shell(exps=np.array([1]))
    with pytest.raises(TypeError):
        _shell(coeffs=np.array([1]))
    with pytest.raises(ValueError):
        _shell(coeffs=np.ones((1, 1, 1)))
    with pytest.raises(ValueError):
        _shell(coeffs=np.ones(2), exps=np.ones(3))
    with pytest.raises(ValueError):
        GeneralizedContractionShell(1, np.zeros(3), np.ones(1), np.ones(1), "bad")
    assert _shell(coeffs=np.ones(3), exps=np.ones(3)).coeffs.shape == (3, 1)
    assert GeneralizedContractionShell(0, np.zeros(3), np.ones(1), np.ones(1), "c").coord_type \
        == "cartesian"
    assert GeneralizedContractionShell(0, np.zeros(3), np.ones(1), np.ones(1), "p").coord_type \
        == "spherical"


def test_component_orders_and_counts():
    assert [tuple(c) for c in _shell(0).angmom_components_cart] == [(0, 0, 0)]
    assert [tuple(c) for c in _shell(1).angmom_components_cart] == [(1, 0, 0), (0, 1, 0),
                                                                     (0, 0, 1)]
    assert [tuple(c) for c in _shell(2).angmom_components_cart] == [
        (2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]
    assert [tuple(c) for c in _shell(3).angmom_components_cart] == [
        (3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (1, 0, 2), (0, 3, 0), (0, 2, 1),
        (0, 1, 2), (0, 0, 3)]
    assert len(_shell(10).angmom_components_cart) == 66
    assert _shell(0).angmom_components_sph == ("c0",)
    assert _shell(1).angmom_components_sph == ("c1", "s1", "c0")
    assert _shell(2).angmom_components_sph == ("s2", "s1", "c0", "c1", "c2")
    assert _shell(3).angmom_components_sph == ("s3", "s2", "s1", "c0", "c1", "c2", "c3")
    for l in range(12):
        s = _shell(l)
        assert s.num_cart == (l + 1) * (l + 2) // 2 and s.num_sph == 2 * l + 1
    assert _shell(1, coeffs=np.ones((2, 5)), exps=np.ones(2)).num_seg_cont == 5


def test_norm_prim_cart_known_answers():
    """Values of the reference's test_norm_prim_cart."""
    s = GeneralizedContractionShell(0, np.zeros(3), np.array([1.0]), np.array([0.25]), "spherical")
    assert np.isclose(s.norm_prim_cart, 0.2519794355383807303479140)
    s = GeneralizedContractionShell(3, np.zeros(3), np.array([1.0]), np.array([0.5]), "spherical")
    assert np.isclose(s.norm_prim_cart[7], 0.6920252830162908851679097)


def _gauss_moment(n, p):
    """integral of x^n exp(-p x^2) over the line (n even)."""
    num = 1.0
    for k in range(1, n, 2):
        num *= k
    return num / (2 * p) ** (n // 2) * np.sqrt(np.pi / p)


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_assign_norm_cont_normalises_every_component(l):
    """After assign_norm_cont, every Cartesian component of a contracted shell with
    generalised columns has unit self-overlap, from the one-centre closed form."""
    exps = np.array([3.1, 0.9, 0.27])
    coeffs = np.array([[0.3, 0.1], [0.5, -0.4], [0.4, 1.0]])
    s = GeneralizedContractionShell(l, np.zeros(3), coeffs, exps, "cartesian")
    s.assign_norm_cont()
    npc = s.norm_prim_cart
    for m in range(coeffs.shape[1]):
        for ci, comp in enumerate(s.angmom_components_cart):
            ovl = 0.0
            for p, a in enumerate(exps):
                for q, b in enumerate(exps):
                    w = coeffs[p, m] * coeffs[q, m] * npc[ci, p] * npc[ci, q]
                    ovl += w * np.prod([_gauss_moment(2 * n, a + b) for n in comp])
            assert abs(ovl * s.norm_cont[m, ci] ** 2 - 1.0) < 1e-12
    one = GeneralizedContractionShell(2, np.zeros(3), np.array([1.0]), np.array([0.25]),
                                      "spherical")
    one.assign_norm_cont()
    assert np.allclose(one.norm_cont, 1.0)
