"""Per-class recursion programs for the cooperative ERI kernel.

For one angular-momentum class (la, lb, lc, ld), already in the kernel's
canonical orientation (la >= lb, lc >= ld, bra "heavier" than ket), this module
enumerates exactly the intermediates the Obara-Saika / Head-Gordon-Pople scheme
of gbasis/integrals/_two_elec_int_improved.py:591-732 needs:

  V(a, m)    vertical recursion on A          (:87-196)
  K(c, a, m) build of the ket centre C         (replaces the electron transfer :199-361)
  G(a,b,c,d) horizontal recursion, D then B   (:364-570)

and lays them out as flat "programs": phases of independent items, each item
naming its destination slot, its source slots and its integer coefficients.
The HIP cooperative kernel interprets a program with all threads of a
workgroup; items inside one phase are independent, phases are separated by a
workgroup barrier.  Direction choice per component is the reference's sweep
order (z if a_z > 0, else y if a_y > 0, else x).

The ket centre is built with the Obara-Saika two-centre vertical recursion

  [a|c+1_i]^(m) = (Q-C)_i [a|c]^(m) + (rho/eta)(P-Q)_i [a|c]^(m+1)
                  + c_i/(2 eta) ([a|c-1_i]^(m) - (rho/eta) [a|c-1_i]^(m+1))
                  + a_i/(2(zeta+eta)) [a-1_i|c]^(m+1)

instead of the reference's electron-transfer relation.  Both give the same
integrals; the transfer relation multiplies by zeta/eta, which for the
exponent ranges of the benchmark workloads (zeta/eta up to ~1e5) amplifies
FP64 rounding to 1e-10 relative and more (measured against an x87
extended-precision evaluation), while every coefficient above is bounded
(rho/eta <= 1).  The cost is 0.9-1.5x the transfer relation's for l <= 2
shells.

Item encoding (8 int32 per item; code = dir | k1 << 4 | k2 << 12):
  VRR : dst, s0=V(a-1i,m), s1=V(a-1i,m+1), s2=V(a-2i,m)|-1, s3=V(a-2i,m+1)|-1, -1, code(k1=a_i-1)
  KVRR: dst, s0=K(c-1i,a,m), s1=K(c-1i,a,m+1), s2=K(c-2i,a,m)|-1, s3=K(c-2i,a,m+1)|-1,
        s4=K(c-1i,a-1i,m+1)|-1, code(k1=c_i-1, k2=a_i)
  ACC : dst=acc index, s0=K/V slot of the final item, rest 0
  HRRK: dst, s0=G(a,0,c+1i,d-1i), s1=G(a,0,c,d-1i), dir, 0...   (ket: + CD_i * s1)
  HRRB: dst, s0=G(a+1i,b-1i,c,d), s1=G(a,b-1i,c,d), dir, 0...   (bra: + AB_i * s1)
"""

from functools import lru_cache
from itertools import groupby

import numpy as np

ITEM_INTS = 8
PH_VRR, PH_KVRR, PH_ACC, PH_HRRK, PH_HRRB = 0, 1, 2, 3, 4


def _code(direction_, k1=0, k2=0):
    return direction_ | (k1 << 4) | (k2 << 12)


@lru_cache(maxsize=None)
def _comps(l):
    return tuple((x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1))


def comps(l):
    """Canonical Cartesian component order (same as the reference default)."""
    return list(_comps(l))


@lru_cache(maxsize=None)
def comp_index(c):
    """Position of component c in comps(deg(c))."""
    return _comps(c[0] + c[1] + c[2]).index(c)


def ncart(l):
    return (l + 1) * (l + 2) // 2


@lru_cache(maxsize=None)
def direction(comp):
    for i in (2, 1, 0):
        if comp[i] > 0:
            return i
    return -1


@lru_cache(maxsize=None)
def sub(c, i, k=1):
    c = list(c)
    c[i] -= k
    return tuple(c)


@lru_cache(maxsize=None)
def add(c, i):
    c = list(c)
    c[i] += 1
    return tuple(c)


def deg(c):
    return c[0] + c[1] + c[2]


Z3 = (0, 0, 0)


def canonical_class(la, lb, lc, ld):
    """Canonical orientation used by the kernels.

    Bra and ket are each ordered l_first >= l_second; the pair with the larger
    (l1 + l2, l1) goes to the bra, so the ket build carries the smaller total
    angular momentum.  Returns (cls, bra_swapped, ket_swapped,
    braket_swapped) with swaps listed in the order they are applied.
    """
    bs = la < lb
    if bs:
        la, lb = lb, la
    ks = lc < ld
    if ks:
        lc, ld = ld, lc
    bk = (la + lb, la) < (lc + ld, lc)
    if bk:
        la, lb, lc, ld = lc, ld, la, lb
    return (la, lb, lc, ld), bs, ks, bk


class ClassProgram:
    """Recursion program of one class; see module docstring for the encoding."""

    def __init__(self, la, lb, lc, ld):
        self.cls = (la, lb, lc, ld)
        L = la + lb + lc + ld
        Lab, Lcd = la + lb, lc + ld
        self.L = L
        # ---- needed sets (top-down) ----
        final_e = [(c, a) for nc in range(lc, Lcd + 1) for c in comps(nc)
                   for na in range(la, Lab + 1) for a in comps(na)]
        need_k = set()
        stack = [(c, a, 0) for c, a in final_e]
        while stack:
            c, a, m = stack.pop()
            if (c, a, m) in need_k:
                continue
            need_k.add((c, a, m))
            if c == Z3:
                continue
            i = direction(c)
            c1 = sub(c, i)
            stack.append((c1, a, m))
            stack.append((c1, a, m + 1))
            if c[i] > 1:
                stack.append((sub(c, i, 2), a, m))
                stack.append((sub(c, i, 2), a, m + 1))
            if a[i] > 0:
                stack.append((c1, sub(a, i), m + 1))
        need_v = set()
        stack = [(a, m) for (c, a, m) in need_k if c == Z3]
        while stack:
            a, m = stack.pop()
            if (a, m) in need_v:
                continue
            need_v.add((a, m))
            if a == Z3:
                continue
            i = direction(a)
            a1 = sub(a, i)
            stack.append((a1, m))
            stack.append((a1, m + 1))
            if a[i] > 1:
                a2 = sub(a, i, 2)
                stack.append((a2, m))
                stack.append((a2, m + 1))
        # ---- slot assignment in the per-primitive-quartet store ----
        # the Boys values V(0, m), m = 0..L occupy slots 0..L
        slot = {}
        for m in range(L + 1):
            slot[("V", Z3, m)] = m
        v_items = sorted((x for x in need_v if x[0] != Z3),
                         key=lambda x: (deg(x[0]), x[1], comp_index(x[0])))
        for a, m in v_items:
            slot[("V", a, m)] = len(slot)
        k_items = sorted((x for x in need_k if x[0] != Z3),
                         key=lambda x: (deg(x[0]), comp_index(x[0]), deg(x[1]),
                                        comp_index(x[1]), x[2]))
        for c, a, m in k_items:
            slot[("K", c, a, m)] = len(slot)
        self.prim_slots = len(slot)

        def ks(c, a, m):
            return slot[("V", a, m)] if c == Z3 else slot[("K", c, a, m)]

        phases = []
        # VRR phases by degree of a (v_items are sorted by it)
        for _, group in groupby(v_items, key=lambda x: deg(x[0])):
            items = []
            for a, m in group:
                i = direction(a)
                a1 = sub(a, i)
                s2 = s3 = -1
                if a[i] > 1:
                    a2 = sub(a, i, 2)
                    s2, s3 = slot[("V", a2, m)], slot[("V", a2, m + 1)]
                items.append([slot[("V", a, m)], slot[("V", a1, m)], slot[("V", a1, m + 1)],
                              s2, s3, -1, _code(i, a[i] - 1), 0])
            if items:
                phases.append((PH_VRR, items))
        # ket-centre phases by degree of c (k_items are sorted by it)
        for _, group in groupby(k_items, key=lambda x: deg(x[0])):
            items = []
            for c, a, m in group:
                i = direction(c)
                c1 = sub(c, i)
                s2 = s3 = s4 = -1
                if c[i] > 1:
                    s2, s3 = ks(sub(c, i, 2), a, m), ks(sub(c, i, 2), a, m + 1)
                if a[i] > 0:
                    s4 = ks(c1, sub(a, i), m + 1)
                items.append([slot[("K", c, a, m)], ks(c1, a, m), ks(c1, a, m + 1), s2, s3, s4,
                              _code(i, c[i] - 1, a[i]), 0])
            if items:
                phases.append((PH_KVRR, items))
        # accumulation of the final [a0|c0] into double accumulators
        self.nacc = len(final_e)
        acc_index = {}
        items = []
        for f, (c, a) in enumerate(final_e):
            acc_index[(a, Z3, c, Z3)] = f
            items.append([f, ks(c, a, 0), 0, 0, 0, 0, 0, 0])
        self.acc_phase = (PH_ACC, items)
        self.prim_phases = phases

        # ---- HRR program on the contracted values (store H; first nacc = acc) ----
        hslot = dict(acc_index)
        hphases = []
        outs = [(a, b, c, d) for a in comps(la) for b in comps(lb) for c in comps(lc)
                for d in comps(ld)]
        need_g = set()
        stack = list(outs)
        while stack:
            g = stack.pop()
            if g in need_g:
                continue
            need_g.add(g)
            a, b, c, d = g
            if b != Z3:
                i = direction(b)
                stack.append((add(a, i), sub(b, i), c, d))
                stack.append((a, sub(b, i), c, d))
            elif d != Z3:
                i = direction(d)
                stack.append((a, b, add(c, i), sub(d, i)))
                stack.append((a, b, c, sub(d, i)))
        ket_items = sorted((g for g in need_g if g[1] == Z3 and g[3] != Z3),
                           key=lambda g: (deg(g[3]), g))
        bra_items = sorted((g for g in need_g if g[1] != Z3), key=lambda g: (deg(g[1]), g))
        for g in ket_items + bra_items:
            if g not in hslot:
                hslot[g] = len(hslot)
        for _, group in groupby(ket_items, key=lambda g: deg(g[3])):
            items = []
            for g in group:
                a, b, c, d = g
                i = direction(d)
                items.append([hslot[g], hslot[(a, b, add(c, i), sub(d, i))],
                              hslot[(a, b, c, sub(d, i))], i, 0, 0, 0, 0])
            if items:
                hphases.append((PH_HRRK, items))
        for _, group in groupby(bra_items, key=lambda g: deg(g[1])):
            items = []
            for g in group:
                a, b, c, d = g
                i = direction(b)
                items.append([hslot[g], hslot[(add(a, i), sub(b, i), c, d)],
                              hslot[(a, sub(b, i), c, d)], i, 0, 0, 0, 0])
            if items:
                hphases.append((PH_HRRB, items))
        self.hrr_phases = hphases
        self.h_slots = len(hslot)
        self.out_slots = np.array([hslot[g] for g in outs], dtype=np.int32)
        self.nout = len(outs)

        self._compact()

    def _compact(self):
        """Slot reuse for the cooperative kernel's storage.

        The program above is in single-assignment form (one slot per value), which
        the register-kernel generator needs.  The cooperative kernel keeps the values
        in LDS, so here every value gets a storage slot only from the phase that
        defines it to the last phase that reads it; a slot read for the last time in
        phase k is reused from phase k+1 on (items of one phase run concurrently).
        """
        self._compact_prim()
        self._compact_hrr()

    def _compact_prim(self):
        """Slot reuse of the per-primitive-quartet store (fields 1..5 of VRR/KVRR
        items are sources)."""
        nph = len(self.prim_phases)
        last = {}
        for k, (_, items) in enumerate(self.prim_phases):
            for it in items:
                for src in it[1:6]:
                    if src >= 0:
                        last[src] = k
        for it in self.acc_phase[1]:
            last[it[1]] = nph
        remap, nslots = _allocate([list(range(self.L + 1))] +
                                  [[it[0] for it in items] for _, items in self.prim_phases],
                                  last)
        # items of a phase are independent: group them by which optional sources they
        # read, so that the lanes of a wavefront mostly take the same path
        self.c_prim_phases = [(kind, sorted([[remap[it[0]]] + [remap[x] if x >= 0 else -1
                                                               for x in it[1:6]] + it[6:]
                                             for it in items],
                                            key=lambda r: (r[3] >= 0, r[5] >= 0)))
                              for kind, items in self.prim_phases]
        self.c_acc_phase = (PH_ACC, [[it[0], remap[it[1]]] + it[2:] for it in self.acc_phase[1]])
        self.c_prim_slots = nslots

    def _compact_hrr(self):
        """Slot reuse of the contracted store of the horizontal recursion (fields 1..2
        are sources)."""
        last = {}
        for k, (_, items) in enumerate(self.hrr_phases):
            for it in items:
                last[it[1]] = k
                last[it[2]] = k
        end = len(self.hrr_phases)
        for o in self.out_slots.tolist():
            last[o] = end
        remap, nslots = _allocate([list(range(self.nacc))] +
                                  [[it[0] for it in items] for _, items in self.hrr_phases], last)
        self.c_hrr_phases = [(kind, [[remap[it[0]], remap[it[1]], remap[it[2]]] + it[3:]
                                     for it in items]) for kind, items in self.hrr_phases]
        self.c_h_slots = nslots
        self.c_out_slots = np.array([remap[o] for o in self.out_slots.tolist()], dtype=np.int32)

    def encode(self):
        """Flatten (slot-reusing form) to (header int32[16], items int32[n*4],
        phases int32[3*nph]).

        One item is four int32 words.  Recursion items (VRR / ket build) pack their
        16-bit storage slots: dst | s0 << 16, s1 | s2 << 16, s3 | s4 << 16, code
        (0xFFFF marks an absent source).  Accumulation items are (acc index, slot,
        0, 0); horizontal-recursion items (dst, src1, src2, direction) with 32-bit
        slots.
        """
        if self.c_prim_slots >= 0xFFFF:
            raise ValueError(f"class {self.cls}: {self.c_prim_slots} primitive slots exceed "
                             "the 16-bit item encoding")

        def pack(lo, hi):
            return (lo & 0xFFFF) | ((hi & 0xFFFF) << 16)

        phases = self.c_prim_phases + [self.c_acc_phase] + self.c_hrr_phases
        rows = []
        table = []
        for kind, its in phases:
            table.append([kind, len(rows), len(rows) + len(its)])
            for it in its:
                if kind in (PH_VRR, PH_KVRR):
                    dst, s0, s1, s2, s3, s4, code = it[:7]
                    rows.append([pack(dst, s0), pack(s1, s2), pack(s3, s4), code])
                elif kind == PH_ACC:
                    rows.append([it[0], it[1], 0, 0])
                else:
                    rows.append([it[0], it[1], it[2], it[3]])
        n_prim_ph = len(self.c_prim_phases)
        header = np.zeros(16, dtype=np.int32)
        header[:8] = [self.c_prim_slots, self.nacc, self.c_h_slots, self.nout, n_prim_ph,
                      len(phases), self.L, len(rows)]
        flat = (np.array(rows, dtype=np.uint32).view(np.int32).reshape(-1) if rows
                else np.zeros(0, np.int32))
        return header, flat, np.array(table, dtype=np.int32).reshape(-1)


def _allocate(defs_by_phase, last_use):
    """Linear-scan slot allocation; defs_by_phase[0] are the inputs (phase -1).

    A value whose last read is in phase k frees its slot for phases > k.  Returns
    (old value -> slot, number of slots).
    """
    import heapq

    free = []
    nslots = 0
    remap = {}
    releases = {}
    for k, defs in enumerate(defs_by_phase):
        phase = k - 1
        for v in releases.pop(phase - 1, []):
            heapq.heappush(free, remap[v])
        for v in defs:
            if free:
                slot = heapq.heappop(free)
            else:
                slot = nslots
                nslots += 1
            remap[v] = slot
            releases.setdefault(last_use.get(v, phase), []).append(v)
    return remap, nslots


@lru_cache(maxsize=None)
def class_program(la, lb, lc, ld):
    return ClassProgram(la, lb, lc, ld)


def _finals_bra_order(p):
    """Accumulator indices ordered by bra component, then ket degree: neighbours share
    most of their ket-build values."""
    la, lb, lc, ld = p.cls
    final_e = [(c, a) for nc in range(lc, lc + ld + 1) for c in comps(nc)
               for na in range(la, la + lb + 1) for a in comps(na)]
    return sorted(range(len(final_e)),
                  key=lambda f: (deg(final_e[f][1]), comp_index(final_e[f][1]),
                                 deg(final_e[f][0])))


def _subprogram(p, finals):
    """Copy of program p whose primitive part produces only the accumulators in
    ``finals`` (global accumulator indices kept in the accumulation items), with its
    own slot reuse.  The horizontal-recursion part is p's."""
    import copy

    q = copy.copy(p)
    deps = {}
    for _, items in p.prim_phases:
        for it in items:
            deps[it[0]] = [x for x in it[1:6] if x >= 0]
    keep = set(finals)
    acc_items = [it for it in p.acc_phase[1] if it[0] in keep]
    need = set()
    stack = [it[1] for it in acc_items]
    while stack:
        x = stack.pop()
        if x in need:
            continue
        need.add(x)
        stack.extend(deps.get(x, []))
    phases = [(k, [it for it in items if it[0] in need]) for k, items in p.prim_phases]
    q.prim_phases = [(k, its) for k, its in phases if its]
    q.acc_phase = (PH_ACC, acc_items)
    q.nacc = len(acc_items)
    q._compact_prim()
    # a primitive-pass program: no horizontal recursion of its own
    q.hrr_phases = q.c_hrr_phases = []
    q.h_slots = q.c_h_slots = 0
    return q


def _hrr_side(lo, span, l_first, l_second):
    """Levels of one side of the horizontal recursion, [x, y+1_i] = [x+1_i, y] +
    G_i [x, y] (i = direction(y + 1_i)): level 0 lists the components x of degree
    lo..lo+span in comps order; level k the (x, y) with deg(y) = k that the final block
    needs, the last one being [(x, y) for x in comps(l_first) for y in comps(l_second)].
    Returns the list of levels, each a list of (s1, s2, dir) into the level below."""
    level0 = [c for n in range(lo, lo + span + 1) for c in comps(n)]
    final = [(x, y) for x in comps(l_first) for y in comps(l_second)]
    need = set()
    stack = list(final)
    while stack:
        x, y = stack.pop()
        if (x, y) in need or y == Z3:
            continue
        need.add((x, y))
        i = direction(y)
        stack.append((add(x, i), sub(y, i)))
        stack.append((x, sub(y, i)))
    index = {(x, Z3): j for j, x in enumerate(level0)}
    levels = []
    for k in range(1, l_second + 1):
        if k == l_second:
            entries = final
        else:
            entries = sorted((g for g in need if deg(g[1]) == k),
                             key=lambda g: (deg(g[0]), comp_index(g[0]), comp_index(g[1])))
        rows = []
        for x, y in entries:
            i = direction(y)
            rows.append((index[(add(x, i), sub(y, i))], index[(x, sub(y, i))], i))
        index = {g: j for j, g in enumerate(entries)}
        levels.append(rows)
    return levels


@lru_cache(maxsize=None)
def hrr_recipe(la, lb, lc, ld):
    """Recipe of the LDS horizontal-recursion pass (gbx_eri_hrr_lds) of a class: int32
    [NA, NC, l_d, l_b, ket level sizes, bra level sizes, entries s1 | s2 << 12 | dir << 24
    of the ket levels, then of the bra levels].  The steps are those of the class
    program's HRR phases (same sources, same direction), grouped by level."""
    ket = _hrr_side(lc, ld, lc, ld)
    bra = _hrr_side(la, lb, la, lb)
    na = sum(ncart(n) for n in range(la, la + lb + 1))
    nc = sum(ncart(n) for n in range(lc, lc + ld + 1))
    head = [na, nc, ld, lb] + [len(v) for v in ket] + [len(v) for v in bra]
    body = [s1 | (s2 << 12) | (i << 24) for v in ket + bra for s1, s2, i in v]
    assert all(len(v) < 4096 for v in ket + bra)
    return np.array(head + body, dtype=np.int32)


@lru_cache(maxsize=None)
def class_subprograms(la, lb, lc, ld, parts):
    """The primitive part of a class split into ``parts`` programs over contiguous runs
    of the bra-ordered accumulators (for classes whose primitive store does not fit
    one workgroup's LDS: each part recomputes the values its accumulators need)."""
    p = class_program(la, lb, lc, ld)
    order = _finals_bra_order(p)
    n = len(order)
    return tuple(_subprogram(p, order[k * n // parts:(k + 1) * n // parts]) for k in range(parts))


# ---------------------------------------------------------------------------
# Encoded programs on disk.  Building a class program is pure Python and takes
# seconds for g shells ((gg|gg): about 7 s); the encoded arrays are cached per
# class in a directory (GBX_PROGRAM_CACHE, default gbasis_amd/_program_cache),
# keyed by a hash of this file so that a changed builder never reads stale
# programs.  Set GBX_PROGRAM_CACHE to an empty string to disable the cache.

def _builder_tag():
    import hashlib
    import os

    with open(os.path.abspath(__file__), "rb") as fh:
        return hashlib.sha1(fh.read()).hexdigest()[:12]


@lru_cache(maxsize=None)
def _cache_dir():
    import os

    d = os.environ.get("GBX_PROGRAM_CACHE")
    if d == "":
        return None
    if d is None:
        d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_program_cache")
    return os.path.join(d, _builder_tag())


def encoded_program(cls, sub=None):
    """(header, items, phases, out_slots) of a class program (``sub = (parts, k)``: of
    the k-th of its primitive-pass programs), from the disk cache when present."""
    import os
    import tempfile

    d = _cache_dir()
    name = "_".join(map(str, cls)) + ("" if sub is None else f"_p{sub[0]}_{sub[1]}") + ".npz"
    path = None if d is None else os.path.join(d, name)
    if path is not None and os.path.exists(path):
        try:
            with np.load(path, allow_pickle=False) as z:
                return z["header"], z["items"], z["phases"], z["out_slots"]
        except (OSError, ValueError, KeyError):
            pass  # unreadable entry: rebuild it below
    prog = class_program(*cls) if sub is None else class_subprograms(*cls, sub[0])[sub[1]]
    header, items, phases = prog.encode()
    out_slots = np.asarray(prog.c_out_slots, dtype=np.int32)
    if path is not None:
        tmp = None
        try:
            os.makedirs(d, exist_ok=True)
            fd, tmp = tempfile.mkstemp(dir=d, suffix=".npz")
            with os.fdopen(fd, "wb") as fh:
                np.savez(fh, header=header, items=items, phases=phases, out_slots=out_slots)
            os.replace(tmp, path)
        except OSError:
            # a read-only or full location only costs the rebuild next time
            if tmp is not None and os.path.exists(tmp):
                try:
                    os.remove(tmp)
                except OSError:
                    pass
    return header, items, phases, out_slots
