"""Algorithmic FLOP and byte model of the ERI path and MI355X peaks.

The per-class FLOP counts follow the Obara-Saika / Head-Gordon-Pople operation
count of the reference algorithm (gbasis/integrals/_two_elec_int_improved.py),
evaluated in the canonical l-order la >= lb, lc >= ld, la >= lc:

  F_prim = 30 (geometry) + 20 + 4(L+1) (Boys)
         + 8 sum_{e=1}^{L} n_c(e) (L-e+1)                          (vertical recursion)
         + 8 sum_{f=1}^{Lcd} n_c(f) S(max(0, la-(Lcd-f)), Lab+Lcd-f) (transfer to C)
         + 2 S(la, Lab) S(lc, Lcd)                                 (accumulation)
  F_cont = 2 S(la,Lab) sum_{d=1}^{ld} n_c(d) S(lc, Lcd-d)           (ket HRR)
         + 2 n_c(lc) n_c(ld) sum_{b=1}^{lb} n_c(b) S(la, Lab-b)     (bra HRR)
         + n_c(la) n_c(lb) n_c(lc) n_c(ld)                          (norms)
         [+ 2 (ns_a nc_a nc_b nc_c nc_d + ns_a ns_b nc_b nc_c nc_d
              + ns_a ns_b ns_c nc_c nc_d + ns_a ns_b ns_c ns_d nc_d) for pure output]

with S(a, b) = sum_{l=a}^{b} n_c(l).  A shell quartet with K_a..K_d primitives
costs K_a K_b K_c K_d F_prim + F_cont.  Algorithmic bytes are 8 per packed slot
written plus the tables read once.
"""

from functools import lru_cache

import numpy as np

PEAK_FP64_VECTOR = 78.6e12   # FLOP/s, MI355X published
PEAK_FP32_VECTOR = 157.3e12  # FLOP/s
PEAK_HBM = 8.0e12            # B/s


def nc(l):
    return (l + 1) * (l + 2) // 2


def S(a, b):
    return sum(nc(l) for l in range(a, b + 1)) if b >= a else 0


def reference_order(la, lb, lc, ld):
    if la < lb:
        la, lb = lb, la
    if lc < ld:
        lc, ld = ld, lc
    if la < lc:
        la, lb, lc, ld = lc, ld, la, lb
    return la, lb, lc, ld


@lru_cache(maxsize=None)
def class_flops(la, lb, lc, ld, pure=False):
    """(F_prim, F_cont) of one class, counted in the reference's canonical order."""
    la, lb, lc, ld = reference_order(la, lb, lc, ld)
    L = la + lb + lc + ld
    Lab, Lcd = la + lb, lc + ld
    f_prim = 30 + 20 + 4 * (L + 1)
    f_prim += 8 * sum(nc(e) * (L - e + 1) for e in range(1, L + 1))
    f_prim += 8 * sum(nc(f) * S(max(0, la - (Lcd - f)), Lab + Lcd - f) for f in range(1, Lcd + 1))
    f_prim += 2 * S(la, Lab) * S(lc, Lcd)
    f_cont = 2 * S(la, Lab) * sum(nc(d) * S(lc, Lcd - d) for d in range(1, ld + 1))
    f_cont += 2 * nc(lc) * nc(ld) * sum(nc(b) * S(la, Lab - b) for b in range(1, lb + 1))
    f_cont += nc(la) * nc(lb) * nc(lc) * nc(ld)
    if pure:
        ns = [2 * l + 1 for l in (la, lb, lc, ld)]
        c = [nc(l) for l in (la, lb, lc, ld)]
        f_cont += 2 * (ns[0] * c[0] * c[1] * c[2] * c[3] + ns[0] * ns[1] * c[1] * c[2] * c[3]
                       + ns[0] * ns[1] * ns[2] * c[2] * c[3] + ns[0] * ns[1] * ns[2] * ns[3] * c[3])
    return f_prim, f_cont


def quartet_flops(cls, nprimq, pure=False, contracted=True):
    """Algorithmic FLOPs of quartets of one class; nprimq may be an array.
    ``contracted=False``: the primitive part only (a sub-quartet of a later phase,
    whose shell quartet's contracted part is counted once, in phase 0)."""
    fp, fc = class_flops(*cls, pure=pure)
    return np.asarray(nprimq, dtype=np.float64) * fp + (fc if contracted else 0)
