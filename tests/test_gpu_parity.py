"""GPU parity: HIP kernels against the oracle and the reference's fixtures."""
import os

import numpy as np
import pytest
import torch

from tests.molecules import random_basis

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu


def _print_case(name):
    print(f"[gpu case] {name}", flush=True)


def test_boys_device_vs_mpmath(gpu):
    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from oracle.boys import boys_mpmath

    _print_case("boys")
    basis = random_basis(0, [0])
    dev = DeviceTables(BatchTables([MoleculeTables(basis)]))
    T = np.array([0.0, 1e-9, 0.05, 0.3, 1.7, 9.99, 25.0, 39.96, 40.0, 40.2, 55.0, 300.0, 1e4, 1e5])
    F = get_engine().boys(dev, T, 16).cpu().numpy()
    for i, t in enumerate(T):
        for m in range(17):
            ref = boys_mpmath(m, t)
            assert abs(F[i, m] - ref) <= 1e-14 * abs(ref) + 1e-300, (t, m, F[i, m], ref)


def test_horton_bec_fixture(gpu):
    """Be+C STO-6G, physicist notation, vs the reference's HORTON fixture."""
    from gbasis_amd.electron_repulsion import electron_repulsion_integral_improved
    from gbasis_amd.parsers import make_contractions, parse_nwchem

    _print_case("horton_bec")
    d = parse_nwchem(os.path.join(HERE, "golden", "sto6g_be_c.nwchem"))
    basis = make_contractions(d, ["Be", "C"], np.array([[0, 0, 0], [1.0, 0, 0]]), "cartesian")
    ref = np.load(os.path.join(HERE, "golden", "horton_bec_cart_elec_repulsion.npy"))
    got = electron_repulsion_integral_improved(basis)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("ls,coord", [([0, 0, 1], "cartesian"), ([0, 1, 2], "cartesian"),
                                      ([1, 2, 2, 0], "cartesian"), ([2, 1, 3], "cartesian"),
                                      ([0, 1, 2, 3], "spherical"), ([4, 1], "spherical")])
def test_dense_vs_oracle(gpu, ls, coord):
    from gbasis_amd.electron_repulsion import electron_repulsion_integral
    from oracle.assembly import dense_eri

    _print_case(f"dense {ls} {coord}")
    basis = random_basis(len(ls) * 11 + ls[-1], ls, kmax=3, coord_type=coord)
    got = electron_repulsion_integral(basis, notation="chemist")
    exact = dense_eri(basis, notation="chemist", exact=True)
    ref64 = dense_eri(basis, notation="chemist")
    scale = max(1.0, np.abs(exact).max())
    # FP64 parity bar: 1e-12 Eh against the extended-precision evaluation
    assert np.abs(got - exact).max() <= 1e-12 * scale
    # against the reference's own FP64 algorithm: no further than that algorithm's own
    # rounding error (its electron transfer loses up to ~3e-10 relative on these inputs)
    ref_err = np.abs(ref64 - exact)
    assert np.all(np.abs(got - ref64) <= ref_err + 1e-12 * scale)


def test_deep_contractions_vs_oracle(gpu):
    """Shells of K = 14 primitives: their 38416 primitive quartets per (ss|ss) lie beyond
    the largest primitive-count bin of the work-queue key, which must keep them FP64 and
    in their own class (screened and mixed-precision runs included)."""
    from gbasis_amd.electron_repulsion import (electron_repulsion_integral_improved,
                                               electron_repulsion_integral_packed)
    from oracle.assembly import dense_eri

    _print_case("deep contractions K=14")
    basis = random_basis(14, [0, 0, 1], kmin=14, kmax=14, emin=0.05, emax=500.0)
    got = electron_repulsion_integral_improved(basis, notation="chemist",
                                               schwarz_threshold=1e-12)
    exact = dense_eri(basis, notation="chemist", exact=True)
    scale = max(1.0, np.abs(exact).max())
    assert np.abs(got - exact).max() <= 1e-12 * scale
    fp64, _ = electron_repulsion_integral_packed(basis, precision="fp64")
    mixed, _ = electron_repulsion_integral_packed(basis, precision="mixed")
    assert float((mixed - fp64).abs().max()) <= 1e-8


@pytest.mark.parametrize("potential,omega", [("erf", 0.4), ("erfc", 0.4), ("erf", 2.5)])
def test_range_separated_vs_oracle(gpu, potential, omega):
    """erf / erfc-attenuated operators (cooperative kernels, FP64) against the oracle's
    extended-precision evaluation, and erf + erfc = 1/r on the device."""
    from gbasis_amd.electron_repulsion import (electron_repulsion_integral,
                                               electron_repulsion_integral_improved)
    from oracle.assembly import dense_eri

    _print_case(f"range-separated {potential} omega={omega}")
    basis = random_basis(91, [0, 1, 2, 3], kmax=3, coord_type="spherical")
    got = electron_repulsion_integral_improved(basis, notation="chemist", potential=potential,
                                               omega=omega)
    exact = dense_eri(basis, notation="chemist", exact=True, potential=potential, omega=omega)
    scale = max(1.0, np.abs(exact).max())
    assert np.abs(got - exact).max() <= 1e-12 * scale
    other = "erfc" if potential == "erf" else "erf"
    got2 = electron_repulsion_integral_improved(basis, notation="chemist", potential=other,
                                                omega=omega)
    coul = electron_repulsion_integral(basis, notation="chemist")
    assert np.abs(got + got2 - coul).max() <= 1e-12 * scale
    # Schwarz screening keeps the Coulomb bounds, as the reference's screener; at 1e-2
    # some quartets of this basis are screened (smallest Q products 4.8e-3 .. 1.2e-2, none
    # within 2 % of the threshold)
    thr = 1e-2
    got_s = electron_repulsion_integral_improved(basis, notation="chemist", potential=potential,
                                                 omega=omega, schwarz_threshold=thr)
    ref_s = dense_eri(basis, notation="chemist", potential=potential, omega=omega,
                      schwarz_threshold=thr, exact=True)
    # a block screened by one bound and kept by the other would differ by its value
    assert np.abs(got_s - ref_s).max() <= 1e-12 * scale
    assert int((got_s == 0.0).sum()) > int((got == 0.0).sum())


def test_construct_array_contraction_vs_oracle(gpu):
    """The per-quartet class method: (M1, L1, ..., M4, L4) Cartesian blocks without
    norm_cont, in the shells' component order, for reordered and generalised cases."""
    from gbasis_amd.contractions import GeneralizedContractionShell
    from gbasis_amd.electron_repulsion import ElectronRepulsionIntegralImproved
    from oracle.assembly import contraction_block

    _print_case("construct_array_contraction")
    rng = np.random.default_rng(12)

    def shell(l, m):
        return GeneralizedContractionShell(l, rng.uniform(-1, 1, 3), rng.uniform(0.2, 1.0, (3, m)),
                                           np.exp(rng.uniform(np.log(0.3), np.log(6.0), 3)),
                                           "cartesian")

    for ls in ([1, 0, 2, 0], [0, 2, 1, 1], [2, 2, 0, 1]):
        conts = [shell(l, 1 + (i == 1)) for i, l in enumerate(ls)]
        got = ElectronRepulsionIntegralImproved.construct_array_contraction(*conts)
        ref = contraction_block(*conts, exact=True)
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), ls


@pytest.mark.gpu
def test_class_interface_vs_oracle(gpu):
    """ElectronRepulsionIntegral[Improved](basis).construct_array_{cartesian, spherical,
    mix, lincomb} on the device against the oracle's extended-precision dense arrays
    (chemists' notation, base_four_symm.py:160-616)."""
    from gbasis_amd.contractions import GeneralizedContractionShell
    from gbasis_amd.electron_repulsion import (ElectronRepulsionIntegral,
                                               ElectronRepulsionIntegralImproved)
    from oracle.assembly import dense_eri

    _print_case("class interface")
    rng = np.random.default_rng(77)
    basis = [GeneralizedContractionShell(l, rng.uniform(-1, 1, 3), rng.uniform(0.2, 1.0, (2, m)),
                                         np.exp(rng.uniform(np.log(0.3), np.log(4.0), 2)),
                                         "spherical")
             for l, m in ((0, 1), (1, 2), (2, 1))]

    def ref(types):
        return dense_eri(basis, notation="chemist", coord_types=types, exact=True)

    def close(a, b):
        assert a.shape == b.shape
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max())

    cart = ref(["cartesian"] * 3)
    mix_types = ["cartesian", "spherical", "spherical"]
    mix = ref(mix_types)
    for cls in (ElectronRepulsionIntegral, ElectronRepulsionIntegralImproved):
        obj = cls(basis)
        close(obj.construct_array_cartesian(), cart)
        close(obj.construct_array_spherical(), ref(["spherical"] * 3))
        close(obj.construct_array_mix(mix_types), mix)
        T = rng.standard_normal((4, mix.shape[0]))
        close(obj.construct_array_lincomb(T, mix_types),
              np.einsum("ai,bj,ck,dl,ijkl->abcd", T, T, T, T, mix))
    close(ElectronRepulsionIntegralImproved(basis).construct_array_cartesian(
        schwarz_threshold=1e-12), cart)


@pytest.mark.gpu
def test_engine_functions_vs_oracle(gpu):
    """compute_two_electron_integrals_os_hgp and the first engine's entry points on raw
    quartet data: (L_a, L_b, L_c, L_d, M_a, ..., M_d) against the oracle's quartet
    (extended precision), a permuted component order, and the all-s form."""
    from gbasis_amd.boys_functions import boys_function_standard, get_boys_function
    from gbasis_amd.two_elec_int import (_compute_two_elec_integrals,
                                         _compute_two_elec_integrals_angmom_zero,
                                         compute_two_electron_integrals_os_hgp)
    from gbasis_amd.utils import default_cart_components
    from oracle.boys import boys_longdouble
    from oracle.os_hgp import os_hgp_quartet

    _print_case("engine functions")
    rng = np.random.default_rng(31)

    def raw(l, m):
        return (rng.uniform(-1, 1, 3), l, default_cart_components(l),
                np.exp(rng.uniform(np.log(0.3), np.log(6.0), 3)), rng.uniform(0.2, 1.0, (3, m)))

    for ls in ([2, 1, 1, 0], [1, 1, 1, 1], [2, 0, 2, 2]):
        q = [raw(l, 1 + (i == 2)) for i, l in enumerate(ls)]
        flat = [x for a in q for x in a]
        ref = os_hgp_quartet(*flat, boys=boys_longdouble, dtype=np.longdouble,
                             ket="vrr").astype(np.float64)
        got = compute_two_electron_integrals_os_hgp(boys_function_standard, *flat)
        assert got.shape == ref.shape
        tol = 1e-12 * max(1.0, np.abs(ref).max())
        assert np.abs(got - ref).max() <= tol, ls
        old = _compute_two_elec_integrals(get_boys_function("coulomb"), *flat)
        assert np.abs(old - ref).max() <= tol, ls
        # a reversed component order on centre a permutes the first axis
        flat_p = list(flat)
        flat_p[2] = flat[2][::-1]
        got_p = compute_two_electron_integrals_os_hgp(boys_function_standard, *flat_p)
        assert np.abs(got_p - got[::-1]).max() <= tol, ls
    q = [raw(0, 2) for _ in range(4)]
    args = [x for a in q for x in (a[0], a[3], a[4])]
    flat = [x for a in q for x in a]
    ref = os_hgp_quartet(*flat, boys=boys_longdouble, dtype=np.longdouble,
                         ket="vrr").astype(np.float64)
    got = _compute_two_elec_integrals_angmom_zero(boys_function_standard, *args)
    assert got.shape == (1, 1, 1, 1, 2, 2, 2, 2)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_schwarz_bounds_and_screened_lists(gpu):
    """Q_ij on the device vs the oracle's screener; screened quartet sets identical."""
    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from gbasis_amd.workloads import batch_molecule
    from oracle.assembly import schwarz_bounds

    _print_case("schwarz")
    basis = batch_molecule(2024, n_atoms=5)
    eng = get_engine()
    mol = MoleculeTables(basis)
    dev = DeviceTables(BatchTables([mol]))
    q = eng.schwarz(dev).cpu().numpy()[:, 0]
    Qdev = mol.schwarz_orig(q)
    Qref = schwarz_bounds(basis)
    Qex = schwarz_bounds(basis, exact=True)
    # Q^2 is an integral: hold it to the 1e-12 Eh bar against the extended-precision value,
    # and to the reference's own FP64 error against the reference
    assert np.all(np.abs(Qdev ** 2 - Qex ** 2) <= 1e-12 * np.maximum(1.0, Qex ** 2))
    assert np.all(np.abs(Qdev ** 2 - Qref ** 2) <= np.abs(Qref ** 2 - Qex ** 2)
                  + 1e-12 * np.maximum(1.0, Qex ** 2))
    n = len(basis)
    iu = np.array([(i, j) for i in range(n) for j in range(i, n)])
    for thr in (1e-12, 1e-8, 1e-4):
        keep_dev = set()
        keep_ref = set()
        borderline = 0
        for p, (i, j) in enumerate(iu):
            for (k, l) in iu[p:]:
                if abs(Qex[i, j] * Qex[k, l] - thr) <= 1e-9 * thr:
                    borderline += 1  # within the FP64 rounding of either engine: not compared
                    continue
                if Qdev[i, j] * Qdev[k, l] >= thr:
                    keep_dev.add((i, j, k, l))
                if Qref[i, j] * Qref[k, l] >= thr:
                    keep_ref.add((i, j, k, l))
        assert borderline == 0
        assert keep_dev == keep_ref


def test_screened_dense_vs_oracle(gpu):
    from gbasis_amd.electron_repulsion import electron_repulsion_integral_improved
    from gbasis_amd.workloads import batch_molecule
    from oracle.assembly import dense_eri

    _print_case("screened dense")
    basis = batch_molecule(77, n_atoms=3)
    thr = 1e-6
    got = electron_repulsion_integral_improved(basis, notation="physicist", schwarz_threshold=thr)
    exact = dense_eri(basis, notation="physicist", schwarz_threshold=thr, exact=True)
    ref64 = dense_eri(basis, notation="physicist", schwarz_threshold=thr)
    scale = max(1.0, np.abs(exact).max())
    assert np.abs(got - exact).max() <= 1e-12 * scale
    # the reference algorithm's own FP64 error on these exponents (0.05..1e3) is large
    assert np.all(np.abs(got - ref64) <= np.abs(ref64 - exact) + 1e-12 * scale)
    # screened shell quartets are exact zero blocks
    from oracle.assembly import basis_offsets, schwarz_bounds

    Q = schwarz_bounds(basis, exact=True)
    off = basis_offsets(basis, [s.coord_type for s in basis])
    chem = got.transpose(0, 2, 1, 3)
    n = len(basis)
    screened = 0
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for l in range(n):
                    if Q[i, j] * Q[k, l] < thr:
                        screened += 1
                        blk = chem[off[i]:off[i + 1], off[j]:off[j + 1], off[k]:off[k + 1],
                                   off[l]:off[l + 1]]
                        assert np.all(blk == 0.0)
    assert screened > 0


def test_mixed_precision_error(gpu):
    """FP32-bulk quartets (Q_ij Q_kl < cutoff) stay within 1e-8 Eh of FP64."""
    from gbasis_amd.electron_repulsion import electron_repulsion_integral_packed
    from gbasis_amd.workloads import batch_molecule

    _print_case("mixed")
    basis = batch_molecule(31, n_atoms=8)
    p64, n = electron_repulsion_integral_packed(basis, schwarz_threshold=1e-12)
    pmx, _ = electron_repulsion_integral_packed(basis, schwarz_threshold=1e-12,
                                                precision="mixed", fp32_cutoff=1e-2)
    err = float((pmx - p64).abs().max())
    print(f"mixed max err {err:.3e}", flush=True)
    assert err <= 1e-8


def test_mixed_after_fp32_mask_changes(gpu):
    """Mixed precision stays within 1e-8 Eh of FP64 when the FP32 class mask changes
    between queue builds on the same tables (autotuners narrow it, a new one starts
    again from the default): each queue is sized for the mask it is filled with, so no
    bucket overflows and no quartet is dropped."""
    import torch

    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from gbasis_amd.workloads import dense_small

    _print_case("mixed after FP32 mask changes")
    eng = get_engine()
    mol = MoleculeTables(dense_small())
    dev = DeviceTables(BatchTables([mol]))
    bound = eng.pair_bounds(dev, eng.schwarz(dev))
    ref = eng.run(dev, eng.plan_device(dev, bound, 0.0, 0.0, stats=True))
    scale = max(1.0, float(ref.abs().max()))
    stats = eng.plan_device(dev, bound, 0.0, 1e-2, stats=True)
    fp32_classes = {(bk.cls, bool(bk.dense)) for bk in stats if bk.precision == 1}
    assert fp32_classes
    try:
        for allowed in ({k: False for k in fp32_classes}, {}):
            eng.set_fp32_allowed(allowed)
            out = torch.zeros_like(ref)
            eng.run(dev, eng.plan_device(dev, bound, 0.0, 1e-2, stats=False,
                                         device_counts=True), out)
            eng.check_queue(dev)
            err = float((out - ref).abs().max())
            print(f"mask {len(allowed)} classes off: mixed max err {err:.3e}", flush=True)
            assert err <= 1e-8 * scale
    finally:
        eng.set_fp32_allowed({})


def test_rys_vs_cooperative(gpu):
    """The Rys primitive pass in both accumulator layouts, with the default LDS batch (up
    to 32 primitive quartets) and with batches of at most 8 (every class with
    6 <= L <= 17 forced onto it, sub-quartet phases included), gives the cooperative
    kernel's packed array within 1e-12 of the largest value, pure and Cartesian shells
    up to g."""
    import torch

    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables

    _print_case("rys vs coop")
    eng = get_engine()
    for seed, ls, coord in ((41, [4, 3, 2, 1], "spherical"), (42, [4, 4, 0, 2], "cartesian")):
        mol = MoleculeTables(random_basis(seed, ls, kmax=3, coord_type=coord, emin=0.05,
                                          emax=2e3))
        dev = DeviceTables(BatchTables([mol]))
        plans = eng.plan_quartets(dev)
        eng.path = "coop"
        try:
            ref = eng.run(dev, plans)
        finally:
            eng.path = "auto"
        scale = max(1.0, float(ref.abs().max()))
        for variant in (("rys", eng.RYS_STRIDED), ("rys", eng.RYS_GROUPED),
                        ("rys", eng.RYS_STRIDED, 8), ("rys", eng.RYS_GROUPED, 8)):
            forced = 0
            for bk in plans:
                bk.variant = None
                if eng.rys_ok(bk.cls):
                    bk.variant = variant
                    forced += 1
            assert forced > 0
            got = eng.run(dev, plans)
            assert all(bk.path == "rys" for bk in plans if bk.variant)
            err = float((got - ref).abs().max())
            print(f"{ls} {coord} {variant}: {forced} buckets on the Rys pass, max diff "
                  f"{err:.2e}", flush=True)
            assert err <= 1e-12 * scale
        torch.cuda.synchronize()


def test_lds_hrr_vs_cooperative_hrr(gpu):
    """The horizontal-recursion pass in LDS (gbx_eri_hrr_lds, forced on every class that
    runs a separate recursion pass and fits) gives the cooperative pass's packed array
    within 1e-12 of the largest value: pure, Cartesian and mixed shells up to g, with
    the build-centre sub-quartet phases (stores and adds)."""
    import torch

    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables

    _print_case("lds hrr vs coop hrr")
    eng = get_engine()
    cases = ((43, [4, 3, 2, 1], "spherical"), (44, [4, 4, 3, 0], "cartesian"),
             (45, [3, 3, 2, 2], "cartesian"), (46, [4, 2, 3, 1], "mixed"))
    phases = set()
    for seed, ls, coord in cases:
        if coord == "mixed":
            basis = (random_basis(seed, ls[:2], kmax=3, coord_type="spherical", emin=0.05,
                                  emax=2e3)
                     + random_basis(seed + 100, ls[2:], kmax=3, coord_type="cartesian",
                                    emin=0.05, emax=2e3))
        else:
            basis = random_basis(seed, ls, kmax=3, coord_type=coord, emin=0.05, emax=2e3)
        dev = DeviceTables(BatchTables([MoleculeTables(basis)]))
        plans = eng.plan_quartets(dev)
        phases |= {bk.phase for bk in plans}
        try:
            eng.hrr_lds = "never"
            ref = eng.run(dev, plans)
            eng.hrr_lds = "always"
            got = eng.run(dev, plans)
        finally:
            eng.hrr_lds = "auto"
        scale = max(1.0, float(ref.abs().max()))
        err = float((got - ref).abs().max())
        print(f"{ls} {coord}: max diff {err:.2e} (scale {scale:.2e})", flush=True)
        assert err <= 1e-12 * scale
        torch.cuda.synchronize()
    assert phases == {0, 1, 2, 3}


def test_register_vs_cooperative(gpu):
    """Both kernel families give the same packed array (FP64)."""
    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from gbasis_amd.workloads import batch_molecule

    _print_case("reg vs coop")
    eng = get_engine()
    mol = MoleculeTables(batch_molecule(5, n_atoms=4))
    dev = DeviceTables(BatchTables([mol]))
    plans = eng.plan_quartets(dev)
    a = eng.run(dev, plans)
    eng.path = "coop"
    try:
        b = eng.run(dev, plans)
    finally:
        eng.path = "auto"
    assert float((a - b).abs().max()) <= 1e-13 * max(1.0, float(a.abs().max()))


def test_autotuned_variants_match_default(gpu):
    """Every variant the autotuner may pick gives the default launch's packed array:
    the tuned plan, and the split classes forced onto the cooperative kernel."""
    from gbasis_amd.autotune import Autotuner
    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from gbasis_amd.workloads import batch_molecule

    _print_case("autotuned variants")
    eng = get_engine()
    mol = MoleculeTables(batch_molecule(11, n_atoms=5))
    dev = DeviceTables(BatchTables([mol]))
    plans = eng.plan_device(dev, None, 0.0, 0.0, stats=True)
    ref = eng.run(dev, plans)
    scale = max(1.0, float(ref.abs().max()))
    tuner = Autotuner(eng, repeats=1)
    scratch = torch.zeros_like(ref)
    tuner.tune(dev, plans, scratch)
    got = eng.run(dev, plans)
    assert float((got - ref).abs().max()) <= 1e-13 * scale
    forced = 0
    for bk in plans:
        if eng.split_count(bk.cls) > 0 and not bk.dense:
            bk.variant = ("coop",)
            forced += 1
    assert forced > 0
    got = eng.run(dev, plans)
    assert float((got - ref).abs().max()) <= 1e-13 * scale


def test_device_queue_matches_host(gpu):
    """Device-built screened quartet lists equal the host enumeration, bucket by bucket."""
    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from gbasis_amd.workloads import batch_molecule

    _print_case("device queue")
    eng = get_engine()
    mols = [MoleculeTables(batch_molecule(s, n_atoms=n)) for s, n in ((1, 3), (2, 6), (3, 5))]
    b = BatchTables(mols)
    dev = DeviceTables(b)
    q = eng.schwarz(dev)
    bound = eng.pair_bounds(dev, q)
    qh = q.cpu().numpy()
    for thr, cut in ((0.0, 0.0), (1e-10, 0.0), (1e-10, 1e-2)):
        dplans = eng.plan_device(dev, bound if (thr or cut) else None, thr, cut)
        hplans = eng.plan_quartets(dev, qh, thr, cut)
        dd, hh = {}, {}
        for bk in dplans:
            dd.setdefault((bk.cls, bk.precision, bk.dense), set()).update(
                map(tuple, bk.host_quartets().tolist()))
        for bk in hplans:
            hh.setdefault((bk.cls, bk.precision, bk.dense), set()).update(
                map(tuple, bk.host_quartets().tolist()))
        assert dd == hh
        assert sum(bk.eris for bk in dplans) == sum(bk.eris for bk in hplans)


def test_device_counts_and_graph_replay(gpu):
    """Launches that read their counts on the device match host-count launches, and a
    HIP graph of the whole step (queue fill + every launch) reproduces the eager
    result on every replay, with eager work between the replays."""
    import torch

    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from gbasis_amd.workloads import batch_molecule

    _print_case("device counts + graph")
    eng = get_engine()
    mols = [MoleculeTables(batch_molecule(s, n_atoms=n)) for s, n in ((11, 4), (12, 6), (13, 3))]
    dev = DeviceTables(BatchTables(mols))
    bound = eng.pair_bounds(dev, eng.schwarz(dev))
    for cut, streams in ((0.0, 1), (1e-2, 1), (0.0, 4)):
        eng.launch_streams = streams
        ref = eng.run(dev, eng.plan_device(dev, bound, 1e-12, cut, stats=False))
        out = torch.zeros_like(ref)
        eng.run(dev, eng.plan_device(dev, bound, 1e-12, cut, stats=False, device_counts=True),
                out)
        tol = (1e-13 if cut == 0 else 1e-8) * max(1.0, float(ref.abs().max()))
        assert float((out - ref).abs().max()) <= tol
        graph, gout, _ = eng.capture(dev, bound, 1e-12, cut)
        eager = gout.clone()
        for rep in range(3):
            # the step rewrites every unscreened slot; screened slots keep their zeros
            gout.zero_()
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(gout, eager), rep
            ref.mul_(1.0)  # eager work between replays
            torch.cuda.synchronize()
    eng.launch_streams = 1


def test_per_molecule_fp32_cutoff(gpu):
    """FP64 and mixed molecules in one batch give what each gives alone."""
    import torch

    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from gbasis_amd.workloads import batch_molecule

    _print_case("per-molecule cutoff")
    eng = get_engine()
    bases = [batch_molecule(21, n_atoms=5), batch_molecule(22, n_atoms=4)]
    mols = [MoleculeTables(b) for b in bases]
    dev = DeviceTables(BatchTables(mols), mol_fp32_cutoff=[0.0, 1e-2])
    bound = eng.pair_bounds(dev, eng.schwarz(dev))
    both = eng.run(dev, eng.plan_device(dev, bound, 1e-12, 0.0, stats=False))
    offs = dev.batch.mol_slot_off
    for i, cut in enumerate((0.0, 1e-2)):
        d1 = DeviceTables(BatchTables([mols[i]]))
        b1 = eng.pair_bounds(d1, eng.schwarz(d1))
        alone = eng.run(d1, eng.plan_device(d1, b1, 1e-12, cut, stats=False))
        part = both[offs[i]:offs[i] + mols[i].n_slots]
        assert float((part - alone).abs().max()) <= 1e-13 * max(1.0, float(alone.abs().max()))
    assert torch.isfinite(both).all()


def test_serving_alone_equals_mix(gpu):
    """A request's packed ERIs are bit-identical alone and batched with others, and
    match the FP64 path (1e-12) or the mixed-precision bound (1e-8)."""
    import torch

    from gbasis_amd.electron_repulsion import electron_repulsion_integral_packed
    from gbasis_amd.serving import EriServer
    from gbasis_amd.workloads import batch_molecule

    _print_case("serving alone == mix")
    srv = EriServer(n_streams=2, hbm_budget=8 << 30, threshold=1e-12, fp32_cutoff=1e-2)
    specs = [(batch_molecule(900 + i, n_atoms=n), p) for i, (n, p) in
             enumerate(((3, "fp64"), (5, "mixed"), (2, "fp64"), (6, "mixed"), (4, "mixed"),
                        (4, "fp64")))]
    alone = []
    for basis, prec in specs:
        r = srv.submit(basis, prec)
        srv.drain()
        alone.append(r.result.clone())
        srv.release(r)
    assert srv.used == 0
    reqs = [srv.submit(b, p, dispatch=False) for b, p in specs]
    srv.drain()
    assert srv.batches_launched >= len(specs) + 1
    for a, r, (basis, prec) in zip(alone, reqs, specs):
        assert torch.equal(a, r.result), r.id
        ref, _ = electron_repulsion_integral_packed(basis, schwarz_threshold=1e-12)
        tol = (1e-12 if prec == "fp64" else 1e-8) * max(1.0, float(ref.abs().max()))
        assert float((r.result - ref).abs().max()) <= tol
        srv.release(r)
    assert srv.used == 0


def _sample_blocks(name):
    z = np.load(os.path.join(HERE, "golden", "config_samples.npz"))
    idx, shapes = z[f"{name}_idx"], z[f"{name}_shape"]
    quad, ref64 = z[f"{name}_quad"], z[f"{name}_fp64"]
    pos = 0
    for q, shp in zip(idx.tolist(), shapes.tolist()):
        n = int(np.prod(shp))
        yield q, shp, quad[pos:pos + n].reshape(shp), ref64[pos:pos + n].reshape(shp)
        pos += n


def _eval_sampled_quartets(basis, quartets, precision, fp32_cutoff=1e-2):
    """Packed array of the whole molecule in which only the given shell quartets
    (original shell indices) were evaluated: each is oriented as the work queue does,
    expanded to its sub-quartets (primitive pairs split by build centre) and launched
    per phase and class; with mixed precision a quartet runs the FP32 bulk iff the
    product of its pairs' precision bounds is below the cutoff and the engine allows
    FP32 for its class, as in the queue builder."""
    import torch

    from gbasis_amd._lib import GBX_FP32_BULK, GBX_FP64, GBX_MODE_PACKED, GBX_MODE_PACKED_ADD
    from gbasis_amd.tables import STORE_PHASES
    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables

    eng = get_engine()
    mol = MoleculeTables(basis)
    batch = BatchTables([mol])
    dev = DeviceTables(batch)
    pair_of = {}
    oi, oj = mol.pair_orig
    for pidx, (i, j) in enumerate(zip(oi.tolist(), oj.tolist())):
        pair_of[(max(i, j), min(i, j))] = pidx
    key = (batch.pair_la + batch.pair_lb) * 16 + batch.pair_la
    bra, ket = [], []
    for i, j, k, l in quartets:
        p, q = pair_of[(max(i, j), min(i, j))], pair_of[(max(k, l), min(k, l))]
        if key[p] < key[q]:
            p, q = q, p
        bra.append(p)
        ket.append(q)
    bra, ket = np.array(bra), np.array(ket)
    prec = np.full(bra.size, GBX_FP64)
    if precision == "mixed":
        pb = eng.schwarz(dev)[:, 1].cpu().numpy()
        mask = eng.fp32_ok.cpu().numpy()
        cls = ((batch.pair_la[bra] * 8 + batch.pair_lb[bra]) * 8 + batch.pair_la[ket]) * 8 \
            + batch.pair_lb[ket]
        dense = (batch.pair_dense[bra] | batch.pair_dense[ket]).astype(np.int64)
        ok = mask[cls * 2 + dense] == 1
        prec[(pb[bra] * pb[ket] < fp32_cutoff) & ok] = GBX_FP32_BULK
    out = torch.zeros(batch.n_slots, dtype=torch.float64, device="cuda")
    for pr in (GBX_FP64, GBX_FP32_BULK):
        sel = prec == pr
        if not sel.any():
            continue
        for phase, xb, xk in batch.sub_quartets(bra[sel], ket[sel]):
            mode = GBX_MODE_PACKED if phase in STORE_PHASES else GBX_MODE_PACKED_ADD
            for cls, nprimq, dense, quart in batch.class_buckets(xb, xk):
                eng.launch_class(dev, cls, torch.from_numpy(quart).cuda(), nprimq, int(pr),
                                 mode, out, dense)
            torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("name", ["dense_small", "large_spdf", "high_l_pure"])
def test_workload_samples_vs_oracle(gpu, name):
    """Sampled shell quartets of the dense s/s/p/d, large s/p/d/f and high-l pure
    s-g workloads (every class present) against the binary128 oracle
    (tests/golden/config_samples.npz, made by tests/golden/make_config_samples.py).

    FP64 bar, per quartet: max|gpu - exact| <= 1e-12 max(1, |block|), the north-star
    tolerance with no allowance for rounding floors.  The device is also never
    further from the reference algorithm's FP64 value than that value is from the
    exact one (plus the same bar: the reference's electron transfer and its
    first-centre horizontal recursion can be off by orders of magnitude here).
    Mixed precision: within 1e-8 of the FP64 result.  Only the sampled quartets are
    evaluated."""
    from gbasis_amd import workloads
    from oracle.assembly import basis_offsets, packed_index

    _print_case(f"workload samples {name}")
    basis = getattr(workloads, name)()
    samples = list(_sample_blocks(name))
    quartets = [s[0] for s in samples]
    got_all = {prec: _eval_sampled_quartets(basis, quartets, prec) for prec in ("fp64", "mixed")}
    off = basis_offsets(basis, [s.coord_type for s in basis])
    worst = {"fp64": 0.0, "mixed": 0.0}
    per_class = {}
    for q, shp, exact, ref64 in samples:
        a, b, c, d = np.meshgrid(*[off[x] + np.arange(m) for x, m in zip(q, shp)],
                                 indexing="ij")
        slots = np.vectorize(packed_index)(a, b, c, d)
        scale = max(1.0, float(np.abs(exact).max()))
        bar = 1e-12 * scale
        got64 = got_all["fp64"][slots]
        err = float(np.abs(got64 - exact).max())
        worst["fp64"] = max(worst["fp64"], err)
        ls = tuple(sorted((basis[x].angmom for x in q), reverse=True))
        per_class[ls] = max(per_class.get(ls, 0.0), err / scale)
        assert err <= bar, (q, err, scale)
        ref_err = np.abs(ref64 - exact)
        assert np.all(np.abs(got64 - ref64) <= ref_err * (1 + 1e-12) + bar), q
        gotmx = got_all["mixed"][slots]
        errmx = float(np.abs(gotmx - got64).max())
        worst["mixed"] = max(worst["mixed"], errmx)
        assert errmx <= 1e-8, (q, errmx)
    top = sorted(per_class.items(), key=lambda kv: -kv[1])[:6]
    print(f"{name}: {len(samples)} quartets within 1e-12; worst FP64 error {worst['fp64']:.2e} "
          f"(relative to max(1,|v|), worst classes {top}), mixed vs FP64 "
          f"{worst['mixed']:.2e}", flush=True)


def test_register_hrr_vs_cooperative_hrr(gpu):
    """The split classes' register horizontal-recursion kernels give the packed array
    of the cooperative horizontal-recursion pass (same recursion, FP64)."""
    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from gbasis_amd.workloads import batch_molecule

    _print_case("register HRR vs cooperative HRR")
    eng = get_engine()
    mols = [MoleculeTables(batch_molecule(s, n_atoms=6)) for s in (41, 42)]
    dev = DeviceTables(BatchTables(mols))
    plans = eng.plan_device(dev, None, 0.0, 0.0, stats=False)
    assert any(eng.has_hrr(bk.cls) for bk in plans)
    eng.hrr_reg = True
    try:
        a = eng.run(dev, plans)
    finally:
        eng.hrr_reg = False
    b = eng.run(dev, plans)
    assert float((a - b).abs().max()) <= 1e-13 * max(1.0, float(b.abs().max()))
