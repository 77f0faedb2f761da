"""ctypes binding of ``libgbasis_amd.so`` (C ABI declared in include/gbasis_amd.h).

The binding follows the reference's own native-call convention
(gbasis/integrals/libcint.py:275-304): explicit ``argtypes``, an ``int``
status return and caller-owned buffers.  Device buffers are torch tensors;
only their ``data_ptr()`` crosses the boundary.  A missing library is an
error: there is no CPU fallback.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgbasis_amd.so")

GBX_FP64 = 0
GBX_FP32_BULK = 1
GBX_MODE_PACKED = 0
GBX_MODE_SCHWARZ = 1
GBX_MODE_PACKED_ADD = 2
GBX_POT_COULOMB = 0
GBX_POT_ERF = 1
GBX_POT_ERFC = 2

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_vp = ctypes.c_void_p


class GbxTables(ctypes.Structure):
    _fields_ = [
        ("shell_info", c_vp), ("comp_map", c_vp), ("comp_scale", c_vp), ("out_mat", c_vp),
        ("pair_info", c_vp), ("pair_geom", c_vp), ("pair_out_base", c_vp), ("prim_pairs", c_vp),
        ("f2norm", c_vp), ("boys_table", c_vp), ("boys_nt", c_i32), ("boys_nord", c_i32),
        ("boys_h", ctypes.c_double), ("potential", c_i32), ("reserved", c_i32),
        ("omega", ctypes.c_double),
    ]


class GbxProgram(ctypes.Structure):
    _fields_ = [("items", c_vp), ("phases", c_vp), ("out_slots", c_vp), ("header", c_i32 * 16)]


# symbol -> (restype, argtypes); every symbol include/gbasis_amd.h declares
SIGNATURES = {
    "gbx_version": (ctypes.c_int, []),
    "gbx_last_error": (ctypes.c_char_p, []),
    "gbx_boys": (ctypes.c_int, [ctypes.POINTER(GbxTables), c_vp, c_i64, c_i32, c_vp, c_vp]),
    "gbx_coop_workspace": (ctypes.c_int, [ctypes.POINTER(GbxProgram), c_i32, c_i32, c_i32, c_i32,
                                          c_i32, c_i32, c_i32, ctypes.POINTER(c_i64),
                                          ctypes.POINTER(c_i64)]),
    "gbx_eri_coop_acc": (ctypes.c_int, [ctypes.POINTER(GbxTables), ctypes.POINTER(GbxProgram),
                                        c_i32, c_i32, c_i32, c_i32, c_vp, c_i64, c_vp, c_i32,
                                        c_i32, c_i32, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "gbx_eri_coop": (ctypes.c_int, [ctypes.POINTER(GbxTables), ctypes.POINTER(GbxProgram), c_i32,
                                    c_i32, c_i32, c_i32, c_vp, c_i64, c_vp, c_i32, c_i32, c_i32,
                                    c_vp, c_vp, c_i64, c_vp]),
    "gbx_eri_coop_noalloc": (ctypes.c_int, [ctypes.POINTER(GbxTables),
                                            ctypes.POINTER(GbxProgram), c_i32, c_i32, c_i32,
                                            c_i32, c_vp, c_i64, c_vp, c_i32, c_i32, c_i32,
                                            c_vp, c_vp]),
    "gbx_eri_rys_acc": (ctypes.c_int, [ctypes.POINTER(GbxTables), c_vp, c_i32, c_i32, c_vp,
                                       c_i32, c_i32, c_i32, c_i32, c_i32, c_vp, c_i64, c_vp,
                                       c_i32, c_vp, c_i64, c_i32, c_vp]),
    "gbx_rys_table_size": (c_i64, []),
    "gbx_eri_hrr_lds": (ctypes.c_int, [ctypes.POINTER(GbxTables), c_i32, c_i32, c_i32, c_i32,
                                       c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_i64, c_vp,
                                       c_i32, c_vp, c_i64, c_i32, c_i32, c_vp, c_vp]),
    "gbx_unpack_dense": (ctypes.c_int, [c_vp, c_i64, c_i32, c_vp, c_vp]),
    "gbx_jk_packed": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "gbx_mo_workspace": (ctypes.c_int, [c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "gbx_mo_transform": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_i32,
                                        c_vp]),
    "gbx_has_reg_kernel": (ctypes.c_int, [c_i32, c_i32, c_i32, c_i32]),
    "gbx_split_count": (ctypes.c_int, [c_i32, c_i32, c_i32, c_i32]),
    "gbx_eri_reg_segments": (ctypes.c_int, [ctypes.POINTER(GbxTables), c_i32, c_i32, c_i32,
                                            c_i32, c_vp, c_vp, c_vp, c_i64, c_i32, c_i32, c_vp,
                                            c_vp]),
    "gbx_eri_split": (ctypes.c_int, [ctypes.POINTER(GbxTables), c_i32, c_i32, c_i32, c_i32, c_i32,
                                     c_vp, c_i64, c_vp, c_i32, c_i32, c_vp, c_i64, c_vp]),
    "gbx_eri_hrr": (ctypes.c_int, [ctypes.POINTER(GbxTables), ctypes.POINTER(GbxProgram), c_i32,
                                   c_i32, c_i32, c_i32, c_vp, c_i64, c_vp, c_i32, c_vp, c_i64, c_i32,
                                   c_vp, c_vp, c_i64, c_vp]),
    "gbx_has_hrr_kernel": (ctypes.c_int, [c_i32, c_i32, c_i32, c_i32]),
    "gbx_eri_hrr_reg": (ctypes.c_int, [ctypes.POINTER(GbxTables), c_i32, c_i32, c_i32, c_i32,
                                       c_vp, c_i64, c_vp, c_vp, c_i64, c_i32, c_i32, c_vp,
                                       c_vp]),
    "gbx_quartets_count": (ctypes.c_int, [c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                          ctypes.c_double, ctypes.c_double, c_vp, c_vp, c_i64,
                                          c_vp, c_vp, c_vp, c_vp]),
    "gbx_quartets_fill": (ctypes.c_int, [c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                         ctypes.c_double, ctypes.c_double, c_vp, c_vp, c_i64,
                                         c_vp, c_vp, c_vp, c_vp]),
    "gbx_eri_reg": (ctypes.c_int, [ctypes.POINTER(GbxTables), c_i32, c_i32, c_i32, c_i32, c_vp,
                                   c_i64, c_vp, c_i32, c_i32, c_vp, c_vp]),
}

_lib = None


def load():
    """Load and bind the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"gbasis_amd: native library missing at {LIB_PATH}; run "
                           "`python -c 'import __graft_entry__ as g; g.build()'` first")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, what):
    if status != 0:
        msg = load().gbx_last_error()
        raise RuntimeError(f"{what} failed ({status}): {msg.decode() if msg else ''}")


def ptr(t):
    """Device (or host) address of a torch tensor / numpy array, None for None."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()
