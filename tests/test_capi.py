"""The C-ABI library loads and exports every symbol include/gbasis_amd.h declares (CPU)."""
import ctypes
import os
import re

from gbasis_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header():
    text = open(os.path.join(ROOT, "include", "gbasis_amd.h")).read()
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _declared():
    return sorted(set(re.findall(r"\b(gbx_[a-z_]+)\s*\(", _header())))


def test_bound_argument_counts_match_header():
    """Every ctypes signature has as many arguments as the C prototype."""
    for name, args in re.findall(r"\b(gbx_[a-z_]+)\s*\(([^)]*)\)", _header()):
        n = 0 if args.strip() in ("", "void") else args.count(",") + 1
        assert len(_lib.SIGNATURES[name][1]) == n, name


def test_header_symbols_exported_and_bound():
    names = _declared()
    assert "gbx_eri_coop" in names and "gbx_eri_reg" in names and "gbx_boys" in names
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, f"{n} not bound in gbasis_amd/_lib.py"


def test_bucket_count_matches_header():
    """The host's work-queue key space is the one the device queue builder uses."""
    from gbasis_amd.tables import NUM_BUCKETS

    m = re.search(r"#define GBX_NUM_BUCKETS (\d+)", open(os.path.join(ROOT, "include",
                                                                    "gbasis_amd.h")).read())
    assert m and int(m.group(1)) == NUM_BUCKETS


def test_error_path_without_device():
    lib = _lib.load()
    assert lib.gbx_version() >= 1
    # argument validation happens before any device call
    rc = lib.gbx_eri_coop(None, None, 0, 0, 0, 0, None, 0, None, 0, 1, 0, None, None, 0, None)
    assert rc == -1
    assert b"null" in lib.gbx_last_error()
    assert lib.gbx_has_reg_kernel(0, 0, 0, 0) == 1
    assert lib.gbx_has_reg_kernel(4, 4, 4, 4) == 0
    fake = ctypes.c_void_p(0x1000)       # never dereferenced: validation fails first
    assert lib.gbx_unpack_dense(None, 4, 0, None, None) == -1
    assert lib.gbx_unpack_dense(fake, 4, 2, fake, None) == -1          # notation
    assert lib.gbx_unpack_dense(fake, 40000, 0, fake, None) == -1      # n^4 overflow
    assert b"32768" in lib.gbx_last_error()
    assert lib.gbx_unpack_dense(None, 0, 0, None, None) == 0           # empty
    assert lib.gbx_jk_packed(fake, 0, fake, fake, fake, None) == -1
    assert lib.gbx_jk_packed(fake, 70000, fake, fake, fake, None) == -1
    assert lib.gbx_jk_packed(None, 4, fake, fake, fake, None) == -1
    assert lib.gbx_boys(None, fake, 1, 0, fake, None) == -1


def test_empty_launch_accepts_null_workspace():
    """A launch of capacity 0 (a class and bin with no quartet in this precision) may pass
    null quartet and accumulator pointers: it succeeds without touching the device."""
    lib = _lib.load()
    tab = _lib.GbxTables()
    tab.boys_nord, tab.boys_nt, tab.boys_h, tab.potential = 33, 401, 0.1, _lib.GBX_POT_COULOMB
    fake = ctypes.c_void_p(0x1000)       # never dereferenced
    cls = (2, 2, 2, 2)
    assert lib.gbx_split_count(*cls) > 0
    assert lib.gbx_eri_split(ctypes.byref(tab), *cls, 0, None, 0, None, _lib.GBX_FP64, 1,
                             None, 0, None) == 0
    assert lib.gbx_eri_split(ctypes.byref(tab), *cls, 0, fake, 4, None, _lib.GBX_FP64, 1,
                             None, 4, None) == -1
    assert b"null" in lib.gbx_last_error()
    if lib.gbx_has_hrr_kernel(2, 2, 2, 1):
        assert lib.gbx_eri_hrr_reg(ctypes.byref(tab), 2, 2, 2, 1, None, 0, None, None, 0, 1,
                                   _lib.GBX_MODE_PACKED, fake, None) == 0


def test_range_separated_argument_errors():
    """Operator checks happen before any device work (CPU)."""
    import pytest

    from gbasis_amd.electron_repulsion import electron_repulsion_integral_packed
    from tests.molecules import random_basis

    basis = random_basis(0, [0, 1])
    with pytest.raises(ValueError):
        electron_repulsion_integral_packed(basis, potential="yukawa")
    with pytest.raises(ValueError):
        electron_repulsion_integral_packed(basis, potential="erf")
    with pytest.raises(ValueError):
        electron_repulsion_integral_packed(basis, potential="erfc", omega=0.3, precision="mixed")


def test_construct_array_contraction_type_checks():
    """As tests/test_electron_repulsion.py:20-35 of the reference: non-shell arguments
    raise TypeError before any device work."""
    import numpy as np
    import pytest

    from gbasis_amd.contractions import GeneralizedContractionShell
    from gbasis_amd.electron_repulsion import (ElectronRepulsionIntegral,
                                               ElectronRepulsionIntegralImproved)

    cont = GeneralizedContractionShell(0, np.zeros(3), np.array([1.0]), np.array([0.1]),
                                       "spherical")
    for cls in (ElectronRepulsionIntegral, ElectronRepulsionIntegralImproved):
        for k in range(4):
            args = [cont] * 4
            args[k] = None
            with pytest.raises(TypeError):
                cls.construct_array_contraction(*args)
