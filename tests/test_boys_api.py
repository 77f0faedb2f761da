"""The Boys-function interface of gbasis/integrals/boys_functions.py, evaluated on
the device: argument checks on the CPU; values, range-separated identities
(tests/test_boys_functions.py:296-436 of the reference) and mpmath on the GPU."""
import numpy as np
import pytest

from gbasis_amd.boys_functions import get_boys_function


def test_argument_errors():
    from gbasis_amd.boys_functions import boys_function_standard
    with pytest.raises(ValueError):
        boys_function_standard(np.array([0, -1]), np.ones(2))
    with pytest.raises(ValueError):
        boys_function_standard(np.array([0.5]), np.ones(1))
    with pytest.raises(ValueError):
        boys_function_standard(np.array([0]), -np.ones(1))
    assert boys_function_standard(np.zeros((0, 1), dtype=int), np.ones(3)).shape == (0, 3)
    from gbasis_amd.boys_functions import boys_function_all_orders
    with pytest.raises(ValueError):
        boys_function_all_orders(-1, np.ones(2))
    with pytest.raises(ValueError):
        boys_function_all_orders(3, -np.ones(2))
    assert boys_function_all_orders(4, np.zeros((0, 2))).shape == (5, 0, 2)


def test_factory_argument_errors():
    with pytest.raises(ValueError):
        get_boys_function("yukawa")
    with pytest.raises(ValueError):
        get_boys_function("erf")
    with pytest.raises(ValueError):
        get_boys_function("erfc_coulomb", omega=None)
    with pytest.raises(ValueError, match="omega parameter is required for erfc"):
        get_boys_function("ERFC")
    f = get_boys_function("erf", omega=0.5)
    with pytest.raises(ValueError, match="rho parameter is required for erf-"):
        f(np.arange(3), np.ones(3))
    for name in ("coulomb", "Standard", "1/r"):
        assert callable(get_boys_function(name))


@pytest.mark.gpu
def test_boys_functions_on_device(gpu):
    from gbasis_amd.boys_functions import (boys_function_erf, boys_function_erfc,
                                           boys_function_standard)
    from oracle.boys import boys_mpmath

    print("[gpu case] boys api", flush=True)
    orders = np.arange(0, 17)[:, None]
    T = np.array([0.0, 1e-8, 0.4, 3.0, 17.0, 39.9, 41.0, 300.0, 1e4])[None, :]
    F = boys_function_standard(orders, T)
    assert F.shape == (17, 9)
    for m in range(17):
        for k, t in enumerate(T[0]):
            ref = boys_mpmath(m, t)
            assert abs(F[m, k] - ref) <= 1e-14 * ref + 1e-300
    rho = np.array([0.3, 1.7, 9.0])[None, :]
    Tr = np.array([0.2, 5.0, 60.0])[None, :]
    for om in (0.1, 0.8, 4.0):
        e = boys_function_erf(orders, Tr, rho, om)
        c = boys_function_erfc(orders, Tr, rho, om)
        np.testing.assert_allclose(e + c, boys_function_standard(orders, Tr), rtol=1e-14,
                                   atol=1e-300)
        s = om ** 2 / (om ** 2 + rho)
        for m in (0, 3, 9):
            for k in range(3):
                ref = s[0, k] ** (m + 0.5) * boys_mpmath(m, s[0, k] * Tr[0, k])
                assert abs(e[m, k] - ref) <= 1e-14 * ref + 1e-300
    # omega -> infinity recovers 1/r; omega -> 0 vanishes
    np.testing.assert_allclose(boys_function_erf(orders, Tr, rho, 1e8),
                               boys_function_standard(orders, Tr), rtol=1e-12)
    assert np.all(np.abs(boys_function_erf(orders, Tr, rho, 1e-8)) < 1e-7)
    # the C ABI: a negative or NaN argument gives NaN for every order
    from gbasis_amd.engine import get_engine
    v = get_engine().boys(None, np.array([-1.0, np.nan, 2.0]), 3).cpu().numpy()
    assert np.all(np.isnan(v[:2])) and np.all(np.isfinite(v[2]))
    from gbasis_amd.boys_functions import boys_function_all_orders
    Ta = np.array([[0.0, 2.5], [45.0, 1e3]])
    A = boys_function_all_orders(20, Ta)
    assert A.shape == (21, 2, 2)
    np.testing.assert_array_equal(A, boys_function_standard(np.arange(21)[:, None, None],
                                                            Ta[None]))
    f = get_boys_function("erfc", omega=0.8)
    np.testing.assert_allclose(f(orders, Tr, rho=rho), boys_function_erfc(orders, Tr, rho, 0.8))


def test_boys_func_class_hook():
    """The reference's class attribute ``boys_func`` (electron_repulsion.py:84,303)
    selects the operator; other callables are rejected, not silently ignored."""
    from gbasis_amd.electron_repulsion import (ElectronRepulsionIntegral,
                                               ElectronRepulsionIntegralImproved)

    from gbasis_amd.point_charge import PointChargeIntegral

    assert ElectronRepulsionIntegral._operator() == ("coulomb", None)
    assert ElectronRepulsionIntegralImproved._operator() == ("coulomb", None)

    class Old(ElectronRepulsionIntegralImproved):    # as the reference's screening tests
        boys_func = PointChargeIntegral.boys_func

    assert Old._operator() == ("coulomb", None)

    class Erfc(ElectronRepulsionIntegralImproved):
        boys_func = get_boys_function("erfc_coulomb", omega=0.3)

    class Attr(ElectronRepulsionIntegralImproved):
        potential, omega = "erf", 0.7

    class Custom(ElectronRepulsionIntegral):
        boys_func = staticmethod(lambda orders, weighted_dist, rho=None: weighted_dist)

    assert Erfc._operator() == ("erfc", 0.3)
    assert Attr._operator() == ("erf", 0.7)
    with pytest.raises(NotImplementedError):
        Custom._operator()


def test_broadcasting_and_order_selection(monkeypatch):
    """Host logic of boys_function_standard / _all_orders with a stand-in device call
    that returns F[i, m] = 100 T_i + m: each argument is evaluated once, and the
    broadcast (orders, weighted_dist) pairs pick the right row and column."""
    import torch

    import gbasis_amd.boys_functions as bf

    calls = []

    class FakeEngine:
        boys_max_order = 24

        def boys(self, dev, T, mmax):
            calls.append(np.asarray(T).size)
            T = torch.as_tensor(np.asarray(T))
            return 100 * T[:, None] + torch.arange(mmax + 1, dtype=torch.float64)[None, :]

    monkeypatch.setattr(bf, "get_engine", lambda device="cuda": FakeEngine())
    orders = np.arange(4)[:, None, None]
    T = np.array([[0.5, 1.5, 2.5], [3.5, 4.5, 5.5]])[None]
    got = bf.boys_function_standard(orders, T)
    assert got.shape == (4, 2, 3) and calls == [6]
    np.testing.assert_array_equal(got, 100 * T + orders)
    np.testing.assert_array_equal(bf.boys_function_standard(2, 7.0), 702.0)
    allo = bf.boys_function_all_orders(3, T[0])
    np.testing.assert_array_equal(allo, 100 * T + orders)
    s = 0.4 ** 2 / (0.4 ** 2 + 0.8)
    np.testing.assert_allclose(bf.get_boys_function("erf", omega=0.4)(np.array([1]), np.array([2.0]),
                                                                       rho=0.8),
                               s ** 1.5 * (100 * s * 2.0 + 1))
