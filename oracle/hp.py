"""ctypes binding of the high-precision ERI oracle (eri_hp.c).  TEST INFRASTRUCTURE ONLY.

``hp_block`` evaluates one contracted shell quartet in IEEE binary128 (``prec="quad"``)
or binary64 with the same operation order (``prec="f64"``), with the primitive pairs of
each shell pair built on the first centre (``mode=0``, the reference's orientation) or
each on the centre of the larger exponent where the horizontal recursion from the
first centre would amplify rounding by more than ``tau`` per unit of angular momentum
(``mode=1``; ``tau=0``: every pair whose second centre has the larger exponent) or by
more than ``tau`` in all (``mode=2``, r^l_second > tau, the device tables' rule).
The library is built on first use (oracle/build_hp.sh, gcc and libquadmath).
"""

import ctypes
import os
import subprocess
from functools import lru_cache

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_hp.so")


@lru_cache(maxsize=None)
def _lib():
    srcs = [os.path.join(_HERE, f) for f in ("eri_hp.c", "eri_hp_body.h")]
    if not os.path.exists(_LIB) or any(os.path.getmtime(s) > os.path.getmtime(_LIB)
                                       for s in srcs if os.path.exists(s)):
        subprocess.check_call([os.path.join(_HERE, "build_hp.sh")])
    lib = ctypes.CDLL(_LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    for sfx in ("q", "d"):
        f = getattr(lib, f"eri_block_{sfx}")
        f.restype = ctypes.c_int
        f.argtypes = [ip, dp, ip, dp, dp, ctypes.c_int, ctypes.c_double, dp, dp]
        g = getattr(lib, f"boys_values_{sfx}")
        g.restype = ctypes.c_int
        g.argtypes = [ctypes.c_double, ctypes.c_int, dp, dp]
    return lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def comps(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def hp_block_raw(ls, centres, exps, coefs, prec="quad", mode=1, tau=0.0, lo=False):
    """Cartesian block (n_c(l1), .., n_c(l4)) in canonical component order with the
    factorial2 norms, no norm_cont, one segmented contraction per shell.  ``exps`` and
    ``coefs`` are four 1-D arrays (raw contraction coefficients).  With ``lo`` also the
    remainder beyond double precision (binary128 only)."""
    lib = _lib()
    l = np.ascontiguousarray(ls, dtype=np.int32)
    c = np.ascontiguousarray(np.asarray(centres, dtype=np.float64).reshape(12))
    K = np.array([np.size(e) for e in exps], dtype=np.int32)
    e = np.ascontiguousarray(np.concatenate([np.ravel(x) for x in exps]), dtype=np.float64)
    w = np.ascontiguousarray(np.concatenate([np.ravel(x) for x in coefs]), dtype=np.float64)
    shape = tuple((x + 1) * (x + 2) // 2 for x in ls)
    hi = np.zeros(int(np.prod(shape)))
    rem = np.zeros_like(hi)
    f = lib.eri_block_q if prec == "quad" else lib.eri_block_d
    rc = f(_p(l, ctypes.c_int), _p(c, ctypes.c_double), _p(K, ctypes.c_int),
           _p(e, ctypes.c_double), _p(w, ctypes.c_double), int(mode), float(tau),
           _p(hi, ctypes.c_double), _p(rem, ctypes.c_double))
    if rc != 0:
        raise ValueError(f"eri_block_{prec}: unsupported input {ls}")
    if lo:
        return hi.reshape(shape), rem.reshape(shape)
    return hi.reshape(shape)


def hp_engine(s1, s2, s3, s4, prec="quad", mode=1, tau=0.0):
    """Contracted Cartesian block in chemists' notation of four shell objects, shape
    (n_1, n_2, n_3, n_4, M_1, M_2, M_3, M_4) with each shell's own component order,
    no norm_cont: the shape and conventions of oracle.os_hgp.os_hgp_quartet."""
    shells = (s1, s2, s3, s4)
    ls = [int(s.angmom) for s in shells]
    centres = np.array([np.asarray(s.coord, dtype=np.float64) for s in shells])
    cols = []
    for s in shells:
        co = np.asarray(s.coeffs, dtype=np.float64)
        cols.append(co[:, None] if co.ndim == 1 else co)
    perm = []
    for s, l in zip(shells, ls):
        pos = {c: i for i, c in enumerate(comps(l))}
        perm.append([pos[tuple(int(v) for v in c)] for c in np.asarray(
            s.angmom_components_cart).reshape(-1, 3).tolist()])
    M = [c.shape[1] for c in cols]
    out = np.zeros([len(p) for p in perm] + M)
    for m1 in range(M[0]):
        for m2 in range(M[1]):
            for m3 in range(M[2]):
                for m4 in range(M[3]):
                    blk = hp_block_raw(ls, centres, [s.exps for s in shells],
                                       [cols[0][:, m1], cols[1][:, m2], cols[2][:, m3],
                                        cols[3][:, m4]], prec, mode, tau)
                    out[..., m1, m2, m3, m4] = blk[np.ix_(*perm)]
    return out


def boys_hp(T, L, prec="quad"):
    """F_m(T), m = 0..L, from the oracle's own evaluation (hi, lo)."""
    lib = _lib()
    hi = np.zeros(L + 1)
    lo = np.zeros(L + 1)
    f = lib.boys_values_q if prec == "quad" else lib.boys_values_d
    if f(float(T), int(L), _p(hi, ctypes.c_double), _p(lo, ctypes.c_double)) != 0:
        raise ValueError("boys_values: bad input")
    return hi, lo
