"""Cartesian->pure matrices vs the HORTON cartpure tables held in the reference tests."""
import json
import os

import numpy as np
import pytest

from gbasis_amd.spherical import generate_transformation
from gbasis_amd.utils import default_cart_components, default_sph_components

HERE = os.path.dirname(os.path.abspath(__file__))


def test_horton_cartpure():
    data = json.load(open(os.path.join(HERE, "golden", "horton_cartpure.json")))
    assert len(data["records"]) == 8
    for r in data["records"]:
        got = generate_transformation(r["l"], np.array(r["cart"]), tuple(r["sph"]), "left")
        np.testing.assert_allclose(got, np.array(r["matrix"]), atol=1e-14)


@pytest.mark.parametrize("l", range(6))
def test_pure_functions_orthonormal(l):
    """Pure functions built from normalised Cartesians are orthonormal: T S T^T = 1."""
    cart = default_cart_components(l)
    T = generate_transformation(l, cart, default_sph_components(l), "left")
    # overlap of normalised Cartesian components on one centre (same exponent)
    def df(n):
        r = 1
        while n > 1:
            r *= n
            n -= 2
        return r
    S = np.zeros((len(cart), len(cart)))
    for i, a in enumerate(cart):
        for j, b in enumerate(cart):
            s = a + b
            if np.any(s % 2):
                continue
            num = np.prod([df(x - 1) for x in s])
            den = np.sqrt(np.prod([df(2 * x - 1) for x in a]) * np.prod([df(2 * x - 1) for x in b]))
            S[i, j] = num / den
    np.testing.assert_allclose(T @ S @ T.T, np.eye(2 * l + 1), atol=1e-13)


def test_bad_labels():
    with pytest.raises(ValueError):
        generate_transformation(1, default_cart_components(1), ("c0", "c1", "s0"), "left")
    with pytest.raises(ValueError):
        generate_transformation(1, default_cart_components(1), ("c0", "c1", "s1"), "middle")
