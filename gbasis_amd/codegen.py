"""Generator of the per-class register kernels (gbasis_amd/csrc/gbx_reg_gen_<k>.hip, with
the dispatch in gbasis_amd/csrc/gbx_reg_gen.hip).

For each angular-momentum class in REG_CLASSES (kernel orientation, see
programs.canonical_class) this emits one HIP kernel in which every lane
evaluates one whole shell quartet: a loop over the bra and ket primitive pairs
with the Boys function, the Obara-Saika recursions and the accumulation into
FP64 registers written out as straight-line code, then the Head-Gordon-Pople
horizontal recursion (always FP64, on the FP64 contracted values), the
factorial2 norms and the placement into the packed array.  The straight-line code is produced from the same class program the
cooperative kernel interprets (programs.ClassProgram), so both paths evaluate
identical recursions; only the storage differs (registers instead of LDS).

Run as a script (``build()`` does) to regenerate the source.
"""

import os
import sys
from functools import lru_cache

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from gbasis_amd.programs import (PH_HRRK, PH_KVRR, PH_VRR, add, class_program, comps,
                                     direction, sub)
else:
    from .programs import PH_HRRK, PH_KVRR, PH_VRR, add, class_program, comps, direction, sub

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "gbx_reg_gen.hip")


REG_BUDGET = 220     # FP64 values a lane may hold (accumulators + live intermediates)
SPLIT_BUDGET = int(os.environ.get("GBX_SPLIT_BUDGET", "200"))   # same, per split part

# Inner (ket primitive) loop unroll of the fused kernels, per class.  Each lane's
# primitive loop is a serial chain (record load -> setup -> Boys table load ->
# recursion); unrolling puts two independent chains in flight per wave where the
# extra registers do not cost occupancy.  Classes not listed: 1.
INNER_UNROLL = {}
_U = int(os.environ.get("GBX_INNER_UNROLL_ALL", "0"))

# Fused kernels: Boys table columns of the class staged in LDS once per workgroup
# (table reads become LDS reads instead of divergent vector-memory loads), and
# the primitive loops ordered with the bra pair innermost (the lanes of a wave
# mostly share the bra pair, so its record loads hit few distinct addresses).
# Classes run as split kernels even though a fused kernel fits the budget (the
# fused kernel would spill a little); extra classes can be named in
# GBX_FORCE_SPLIT as e.g. "2111,2120".
FORCE_SPLIT = set()
for _c in filter(None, os.environ.get("GBX_FORCE_SPLIT", "").split(",")):
    FORCE_SPLIT.add(tuple(int(x) for x in _c))

# Split classes whose register HRR kernel would spill (two ket recursion levels keep
# too many values live per column): they keep the cooperative HRR pass.
HRR_REG_EXCLUDE = {(2, 2, 2, 2)}


def _hrr_reg(cls):
    """Split class with a register HRR kernel (classes with an f shell use the
    cooperative HRR pass)."""
    return cls not in HRR_REG_EXCLUDE and max(cls) <= 2

BOYS_LDS = os.environ.get("GBX_BOYS_LDS", "1") == "1"
LOOP_SWAP = os.environ.get("GBX_LOOP_SWAP", "1") == "1"


def _deps(p):
    """value -> list of source values of its recursion item (single-assignment program)."""
    deps = {}
    for _, items in p.prim_phases:
        for it in items:
            deps[it[0]] = [x for x in it[1:6] if x >= 0]
    return deps


def schedule(p, finals):
    """Register-pressure-aware order of the items that produce ``finals``.

    Greedy list scheduling over the recursion DAG: among the ready items pick the
    one that frees the most operands (their last reader) and creates no long-lived
    value; every final is accumulated as soon as it is produced.  Returns
    (ordered item destinations, peak number of live values).
    """
    deps = _deps(p)
    need = set()
    stack = list(finals)
    while stack:
        x = stack.pop()
        if x in need:
            continue
        need.add(x)
        stack.extend(deps.get(x, []))
    users = {}
    for d in need:
        if d <= p.L:
            continue
        for src in set(deps[d]):
            users.setdefault(src, set()).add(d)
    remaining = {v: len(u) for v, u in users.items()}
    pending = {d: sum(1 for src in set(deps[d]) if src > p.L) for d in need if d > p.L}
    ready = {d for d, c in pending.items() if c == 0}
    live = {m for m in range(p.L + 1) if m in need}
    finals = set(finals)
    order = []
    peak = len(live)
    while ready:
        best, best_key = None, None
        for d in ready:
            freed = sum(1 for src in set(deps[d]) if remaining.get(src, 0) == 1)
            created = 0 if (d in finals and d not in users) else 1
            key = (freed - created, -d)
            if best_key is None or key > best_key:
                best, best_key = d, key
        ready.discard(best)
        order.append(best)
        live.add(best)
        peak = max(peak, len(live))
        for src in set(deps[best]):
            remaining[src] -= 1
            if remaining[src] == 0:
                live.discard(src)
        if remaining.get(best, 0) == 0:
            live.discard(best)
        for u in users.get(best, ()):
            pending[u] -= 1
            if pending[u] == 0:
                ready.add(u)
    return order, peak


def _finals_by_bra(p):
    """Accumulator indices ordered by bra component first (splits share the ket build)."""
    la, lb, lc, ld = p.cls
    final_e = [(c, a) for nc in range(lc, lc + ld + 1) for c in comps(nc)
               for na in range(la, la + lb + 1) for a in comps(na)]
    idx = sorted(range(len(final_e)),
                 key=lambda f: (sum(final_e[f][1]), comps(sum(final_e[f][1])).index(final_e[f][1]),
                                sum(final_e[f][0])))
    return idx


@lru_cache(maxsize=None)
def class_plan(cls):
    """('fused', [(acc indices, order)]) or ('split', [(acc indices, order), ...])."""
    p = class_program(*cls)
    src_of = [it[1] for it in p.acc_phase[1]]          # acc index -> final value
    all_f = list(range(p.nacc))
    order, peak = schedule(p, [src_of[f] for f in all_f])
    if p.nacc + peak <= REG_BUDGET and cls not in FORCE_SPLIT:
        return "fused", [(all_f, order)]
    byb = _finals_by_bra(p)
    for S in range(2, 33):
        chunks = [byb[k * len(byb) // S:(k + 1) * len(byb) // S] for k in range(S)]
        parts = []
        ok = True
        for ch in chunks:
            o, pk = schedule(p, [src_of[f] for f in ch])
            if len(ch) + pk > SPLIT_BUDGET:
                ok = False
                break
            parts.append((ch, o))
        if ok:
            return "split", parts
    raise RuntimeError(f"class {cls}: no split fits the register budget")


# Classes with an f shell get split register kernels (FP64 only, cooperative HRR pass)
# when their accumulators split into at most F_MAX_PARTS register-sized parts; the
# others (e.g. (ff|ff), (ff|fd), (fd|fd)) stay on the cooperative kernel.  Classes with
# more accumulators than F_MAX_NACC are not tried.
F_SHELLS = os.environ.get("GBX_F_REGISTER", "1") == "1"
F_MAX_NACC = 2400
F_MAX_PARTS = 32


def _orientations(out):
    """The orientations of the sub-quartets whose bra or ket entry is built on its second
    centre (entries (lb, la) with lb > 0, and likewise on the ket)."""
    extra = set()
    for la, lb, lc, ld in out:
        for bra in ((la, lb), (lb, la)) if lb > 0 else ((la, lb),):
            for ket in ((lc, ld), (ld, lc)) if ld > 0 else ((lc, ld),):
                extra.add(bra + ket)
    return sorted(extra - set(out))


def _canonical(lmax):
    out = []
    for la in range(lmax + 1):
        for lb in range(la + 1):
            for lc in range(lmax + 1):
                for ld in range(lc + 1):
                    if (la + lb, la) < (lc + ld, lc):
                        continue
                    out.append((la, lb, lc, ld))
    return out


def _plannable(cls):
    """True if the class has a fused or split register plan."""
    if class_program(*cls).nacc > F_MAX_NACC:
        return False
    try:
        kind, parts = class_plan(cls)
    except RuntimeError:
        return False
    return kind == "fused" or len(parts) <= F_MAX_PARTS


def _reg_classes():
    """Classes that get generated register kernels: with l <= 2 per shell, the kernel
    orientations of the shell quartets (la >= lb, lc >= ld, bra heavier) and of their
    sub-quartets; then the classes with an f shell (and their sub-quartet orientations)
    that have a register plan."""
    low = _canonical(2)
    out = low + _orientations(low)
    if F_SHELLS:
        f = [c for c in _canonical(3) if max(c) == 3]
        f = f + _orientations(f)
        out += [c for c in f if _plannable(c)]
    return out


def fp64_only(cls):
    """Register kernels of classes with an f shell are instantiated in FP64 only (the
    mixed form keeps such classes in FP64)."""
    return max(cls) >= 3


def _f2(comp):
    def df(n):
        r = 1
        while n > 1:
            r *= n
            n -= 2
        return r
    return 1.0 / (df(2 * comp[0] - 1) * df(2 * comp[1] - 1) * df(2 * comp[2] - 1)) ** 0.5


def _item_table(p):
    table = {}
    for kind, items in p.prim_phases:
        for it in items:
            table[it[0]] = (kind, it)
    return table


def _emit_primitive_loop(w, p, order, acc_of_value, acc_ids, strided=False, lds=False,
                         swap=False):
    """Loop over bra x ket primitive pairs; straight-line recursion in ``order``;
    every final value is added to its FP64 accumulator as soon as it exists.

    ``strided``: the lane handles primitive quartets r, r + R, r + 2R, ... of the
    flattened bra x ket range (R lanes share one shell quartet)."""
    L = p.L
    table = _item_table(p)
    for f in acc_ids:
        w(f"  double acc{f} = 0.0;")
    if strided:
        w("  const int nprimq = pb.w * pk.w;")
        w("  for (int pq = r; pq < nprimq; pq += a.lanes_per_quartet) {")
        w("    const int ib = pq / pk.w, ik = pq - (pq / pk.w) * pk.w;")
        w("    const double* rb = t.prim_pairs + static_cast<int64_t>(pb.z + ib) * kPrimRec;")
        w("    {")
    elif swap:
        w("  for (int ik = 0; ik < pk.w; ++ik) {")
        w("    const double* rk = t.prim_pairs + static_cast<int64_t>(pk.z + ik) * kPrimRec;")
        u = _U or INNER_UNROLL.get(p.cls, 1)
        if u > 1:
            w(f"#pragma unroll {u}")
        w("    for (int ib = 0; ib < pb.w; ++ib) {")
        w("      const double* rb = t.prim_pairs + static_cast<int64_t>(pb.z + ib) * kPrimRec;")
    else:
        w("  for (int ib = 0; ib < pb.w; ++ib) {")
        w("    const double* rb = t.prim_pairs + static_cast<int64_t>(pb.z + ib) * kPrimRec;")
        u = _U or INNER_UNROLL.get(p.cls, 1)
        if u > 1:
            w(f"#pragma unroll {u}")
        w("    for (int ik = 0; ik < pk.w; ++ik) {")
    if not swap or strided:
        w("      const double* rk = t.prim_pairs + static_cast<int64_t>(pk.z + ik) * kPrimRec;")
    w("      PrimQuartet<TB> s;")
    w("      const double T = prim_quartet_setup(rb, rk, s);")
    w(f"      double F[{L + 1}];")
    if lds:
        w(f"      boys_eval_lds<{L}>(T, t, boys_lds, F);")
    else:
        w(f"      boys_eval_fixed<{L}>(T, t, F);")
    used = set()
    for d in order:
        used.update(x for x in table[d][1][1:6] if x >= 0)
    for m in range(L + 1):
        if m in used or m in acc_of_value:
            w(f"      const TB x{m} = static_cast<TB>(F[{m}] * s.pref);")
    k_vrr, k_kv1, k_kv2 = set(), set(), set()
    for d in order:
        kind, it = table[d]
        code = it[6]
        k1, k2 = (code >> 4) & 255, (code >> 12) & 255
        if kind == PH_VRR and it[3] >= 0:
            k_vrr.add(k1)
        if kind == PH_KVRR:
            if it[3] >= 0:
                k_kv1.add(k1)
            if it[5] >= 0:
                k_kv2.add(k2)
    for k in sorted(k_vrr):
        w(f"      const TB z{k} = static_cast<TB>({k}) * s.i2z;")
    for k in sorted(k_kv1):
        w(f"      const TB e{k} = static_cast<TB>({k}) * s.i2e;")
    for k in sorted(k_kv2):
        w(f"      const TB r{k} = static_cast<TB>({k}) * s.i2s;")
    # finals that are Boys values themselves (classes with L = 0)
    for m in range(L + 1):
        for f in acc_of_value.get(m, ()):
            w(f"      acc{f} += static_cast<double>(x{m});")
    for d in order:
        kind, it = table[d]
        dst, s0, s1, s2, s3, s4, code, _ = it
        i = code & 15
        k1, k2 = (code >> 4) & 255, (code >> 12) & 255
        if kind == PH_VRR:
            e = f"s.PA[{i}] * x{s0} + s.WP[{i}] * x{s1}"
            if s2 >= 0:
                e += f" + z{k1} * (x{s2} - s.roz * x{s3})"
        else:
            e = f"s.QC[{i}] * x{s0} + s.WQ[{i}] * x{s1}"
            if s2 >= 0:
                e += f" + e{k1} * (x{s2} - s.roe * x{s3})"
            if s4 >= 0:
                e += f" + r{k2} * x{s4}"
        w(f"      const TB x{dst} = {e};")
        for f in acc_of_value.get(dst, ()):
            w(f"      acc{f} += static_cast<double>(x{dst});")
    w("    }")
    w("  }")


def _head(w, name, strided=False, lds_L=None):
    w("template <typename TB>")
    w(f"__global__ __launch_bounds__(kRegThreads) void {name}(gbx_tables t, RegArgs a) {{")
    if lds_L is not None:
        # every thread of the workgroup helps stage the table before any lane leaves
        w("  __shared__ __align__(16) double boys_lds[kBoysRows * kBoysLdsStride];")
        w(f"  boys_stage_lds<{lds_L}>(t, boys_lds);")
        w("  __syncthreads();")
    if strided:
        # lane -> (r, q): consecutive lanes take consecutive quartets of one stride slot r
        w("  const int64_t gid = static_cast<int64_t>(blockIdx.x) * kRegThreads + threadIdx.x;")
        w("  if (gid >= a.nquart * a.lanes_per_quartet) return;")
        w("  const int r = static_cast<int>(gid / a.nquart);")
        w("  const int64_t q = gid - static_cast<int64_t>(r) * a.nquart;")
        w("  if (a.count && q >= *a.count) return;")
    else:
        w("  const int64_t q = lane_quartet(a, static_cast<int64_t>(blockIdx.x) * kRegThreads + "
          "threadIdx.x);")
        w("  if (q < 0) return;")
    w("  const int bra = a.quart[2 * q], ket = a.quart[2 * q + 1];")
    w("  const int4 pb = reinterpret_cast<const int4*>(t.pair_info)[bra];")
    w("  const int4 pk = reinterpret_cast<const int4*>(t.pair_info)[ket];")


def kernel_source(cls):
    """Fused kernel: primitive loop, FP64 horizontal recursion, norms, placement."""
    la, lb, lc, ld = cls
    p = class_program(*cls)
    kind, parts = class_plan(cls)
    assert kind == "fused"
    acc_ids, order = parts[0]
    acc_of_value = {}
    for f, it in enumerate(p.acc_phase[1]):
        acc_of_value.setdefault(it[1], []).append(f)
    name = f"eri_reg_{la}{lb}{lc}{ld}"
    o = []
    w = o.append
    w(f"// class ({la}{lb}|{lc}{ld}): {len(order)} recursion values, {p.nacc} accumulators, "
      f"{p.nout} outputs")
    _head(w, name, lds_L=p.L if BOYS_LDS else None)
    _emit_primitive_loop(w, p, order, acc_of_value, acc_ids, lds=BOYS_LDS, swap=LOOP_SWAP)
    # largest contracted [e0|f0] (input of the horizontal recursion), for the
    # precision bound of GBX_MODE_SCHWARZ
    w("  double accmax = 0.0;")
    w("  if (a.mode == GBX_MODE_SCHWARZ) {")
    for f in range(p.nacc):
        w(f"    accmax = fmax(accmax, fabs(acc{f}));")
    w("  }")
    # horizontal recursion on the FP64 contracted values
    w("  const double AB[3] = {t.pair_geom[4 * bra], t.pair_geom[4 * bra + 1], "
      "t.pair_geom[4 * bra + 2]};")
    w("  const double CD[3] = {t.pair_geom[4 * ket], t.pair_geom[4 * ket + 1], "
      "t.pair_geom[4 * ket + 2]};")
    for f in range(p.nacc):
        w(f"  const double h{f} = acc{f};")
    for kind_, items in p.hrr_phases:
        g = "CD" if kind_ == PH_HRRK else "AB"
        for it in items:
            w(f"  const double h{it[0]} = h{it[1]} + {g}[{it[3]}] * h{it[2]};")
    outs = [(a, b, c, d) for a in comps(la) for b in comps(lb) for c in comps(lc)
            for d in comps(ld)]
    nca, ncb, ncc, ncd = (len(comps(x)) for x in cls)
    w(f"  double out[{p.nout}];")
    for o_idx, ((a, b, c, d), hs) in enumerate(zip(outs, p.out_slots.tolist())):
        f2 = _f2(a) * _f2(b) * _f2(c) * _f2(d)
        if abs(f2 - 1.0) < 1e-300:
            w(f"  out[{o_idx}] = h{hs};")
        else:
            w(f"  out[{o_idx}] = h{hs} * {f2!r};")
    w(f"  store_block<{nca}, {ncb}, {ncc}, {ncd}>(t, a, bra, ket, pb, pk, out, accmax);")
    w("}")
    return "\n".join(o)


def split_kernel_source(cls, k):
    """Split kernel k: primitive loop for one subset of the accumulators.  R lanes
    (a.lanes_per_quartet) share a shell quartet, each summing a strided subset of
    its primitive quartets; lane r writes its partial sums to
    a.acc[(f * R + r) * a.acc_stride + q] for the horizontal-recursion pass."""
    la, lb, lc, ld = cls
    p = class_program(*cls)
    _, parts = class_plan(cls)
    acc_ids, order = parts[k]
    src_of = [it[1] for it in p.acc_phase[1]]
    acc_of_value = {}
    for f in acc_ids:
        acc_of_value.setdefault(src_of[f], []).append(f)
    name = f"eri_split_{la}{lb}{lc}{ld}_{k}"
    o = []
    w = o.append
    w(f"// class ({la}{lb}|{lc}{ld}) split {k}/{len(parts)}: {len(order)} recursion values, "
      f"{len(acc_ids)} accumulators")
    _head(w, name, strided=True)
    _emit_primitive_loop(w, p, order, acc_of_value, acc_ids, strided=True)
    w("  const int64_t R = a.lanes_per_quartet;")
    for f in acc_ids:
        w(f"  a.acc[({f} * R + r) * a.acc_stride + q] = acc{f};")
    w("}")
    return "\n".join(o)


def hrr_kernel_source(cls):
    """Horizontal-recursion kernel of a split class (FP64, packed output, shells with a
    scaled-permutation output map): one lane per shell quartet.  The lane sums the R
    partial accumulators of the split kernels and, one ket component pair (c, d) at a
    time, runs the HRR on D then B for that column in registers and stores the
    column's (a, b) block, so the live set stays one column wide."""
    la, lb, lc, ld = cls
    p = class_program(*cls)
    acc_index = {}
    # accumulator numbering of the class program: final_e order (c, a)
    f = 0
    for nc in range(lc, lc + ld + 1):
        for c in comps(nc):
            for na in range(la, la + lb + 1):
                for a in comps(na):
                    acc_index[(a, (0, 0, 0), c, (0, 0, 0))] = f
                    f += 1
    assert f == p.nacc
    nca, ncb, ncc, ncd = (len(comps(x)) for x in cls)
    name = f"eri_hrr_{la}{lb}{lc}{ld}"
    o = []
    w = o.append
    w(f"// class ({la}{lb}|{lc}{ld}): horizontal recursion per ket column, {p.nacc} accumulators")
    w(f"__global__ __launch_bounds__(kRegThreads) void {name}(gbx_tables t, RegArgs a) {{")
    w("  const int64_t q = lane_quartet(a, static_cast<int64_t>(blockIdx.x) * kRegThreads + "
      "threadIdx.x);")
    w("  if (q < 0) return;")
    w("  const int bra = a.quart[2 * q], ket = a.quart[2 * q + 1];")
    w("  const int4 pb = reinterpret_cast<const int4*>(t.pair_info)[bra];")
    w("  const int4 pk = reinterpret_cast<const int4*>(t.pair_info)[ket];")
    w("  const double AB[3] = {t.pair_geom[4 * bra], t.pair_geom[4 * bra + 1], "
      "t.pair_geom[4 * bra + 2]};")
    w("  const double CD[3] = {t.pair_geom[4 * ket], t.pair_geom[4 * ket + 1], "
      "t.pair_geom[4 * ket + 2]};")
    w(f"  ColumnStore<{nca}, {ncb}, {ncc}, {ncd}> st(t, a, bra, ket, pb, pk);")
    Z = (0, 0, 0)
    for ic, c in enumerate(comps(lc)):
        for id_, d in enumerate(comps(ld)):
            # depth-first post-order from each output of the column: every value is
            # computed right before its first use, accumulators are loaded there, and
            # each output is stored as soon as it exists
            names = {}
            w("  {")
            w(f"    const ColumnSlot cl = st.column({ic}, {id_});")

            def emit(g):
                if g in names:
                    return names[g]
                aa, bb, cc, dd = g
                if bb != Z:
                    i = direction(bb)
                    x, y = emit((add(aa, i), sub(bb, i), cc, dd)), emit((aa, sub(bb, i), cc, dd))
                    expr = f"{x} + AB[{i}] * {y}"
                elif dd != Z:
                    i = direction(dd)
                    x, y = emit((aa, bb, add(cc, i), sub(dd, i))), emit((aa, bb, cc, sub(dd, i)))
                    expr = f"{x} + CD[{i}] * {y}"
                else:
                    expr = f"sum_parts(a, {acc_index[g]}, q)"
                nm = f"v{len(names)}"
                names[g] = nm
                w(f"    const double {nm} = {expr};")
                return nm

            for ia, aa in enumerate(comps(la)):
                for ib, bb in enumerate(comps(lb)):
                    v = emit((aa, bb, c, d))
                    f2 = _f2(aa) * _f2(bb) * _f2(c) * _f2(d)
                    val = v if abs(f2 - 1.0) < 1e-300 else f"{v} * {f2!r}"
                    w(f"    st.put(cl, {ia}, {ib}, {val});")
            w("  }")
    w("}")
    return "\n".join(o)


N_PARTS = 8   # translation units of the generated kernels (compiled in parallel)
NOT_HERE = 1  # return value of a part's launcher for a class it does not hold


def _key(c):
    return ((c[0] * 8 + c[1]) * 8 + c[2]) * 8 + c[3]


def _launch_lines(cls, name):
    """Launch of a register kernel in the precision asked for (FP64-only classes refuse
    FP32)."""
    if fp64_only(cls):
        return ["      if (precision != GBX_FP64)",
                '        return fail(GBX_ERR_UNSUPPORTED, "register kernels of classes with '
                'an f shell are FP64 only");',
                f"      hipLaunchKernelGGL(({name}<double>), dim3(grid), dim3(kRegThreads), 0, "
                "s, *t, a);",
                f'      return check_launch("{name}");']
    return [f"      if (precision == GBX_FP64) hipLaunchKernelGGL(({name}<double>), "
            "dim3(grid), dim3(kRegThreads), 0, s, *t, a);",
            f"      else hipLaunchKernelGGL(({name}<float>), dim3(grid), "
            "dim3(kRegThreads), 0, s, *t, a);",
            f'      return check_launch("{name}");']


def _part_sources(classes):
    """Kernel sources of every class, grouped into N_PARTS translation units of similar
    size, each with its own launchers (which return NOT_HERE for other classes)."""
    units = []
    for cls in classes:
        kind, pl = class_plan(cls)
        src = []
        if kind == "fused":
            src.append(kernel_source(cls))
        else:
            for k in range(len(pl)):
                src.append(split_kernel_source(cls, k))
            if _hrr_reg(cls):
                src.append(hrr_kernel_source(cls))
        units.append((cls, kind, len(pl), "\n".join(src)))
    bins = [[] for _ in range(N_PARTS)]
    size = [0] * N_PARTS
    for u in sorted(units, key=lambda u: -len(u[3])):
        k = size.index(min(size))
        bins[k].append(u)
        size[k] += len(u[3])
    out = []
    for k, members in enumerate(bins):
        members.sort(key=lambda u: _key(u[0]))
        o = ["// GENERATED by gbasis_amd/codegen.py -- do not edit.",
             f"// Per-class register kernels, part {k} of {N_PARTS}: one lane = one shell quartet.",
             '#include "gbx_reg.h"', "", "namespace gbx {", ""]
        o += [u[3] + "\n" for u in members]
        o.append(f"int launch_reg_part{k}(const gbx_tables* t, int key, const RegArgs& a, "
                 "int precision, hipStream_t s) {")
        o.append("  const unsigned grid = static_cast<unsigned>((a.nquart + kRegThreads - 1) / "
                 "kRegThreads);")
        o.append("  (void)grid;")
        o.append("  switch (key) {")
        for cls, kind, _, _ in members:
            if kind != "fused":
                continue
            name = f"eri_reg_{''.join(map(str, cls))}"
            o.append(f"    case {_key(cls)}:")
            o += _launch_lines(cls, name)
        o.append("    default:")
        o.append(f"      return {NOT_HERE};")
        o.append("  }")
        o.append("}")
        o.append("")
        o.append(f"int launch_split_part{k}(const gbx_tables* t, int key, int split, "
                 "const RegArgs& a, int precision, hipStream_t s) {")
        o.append("  const unsigned grid = static_cast<unsigned>((a.nquart * a.lanes_per_quartet + "
                 "kRegThreads - 1) / kRegThreads);")
        o.append("  (void)grid;")
        o.append("  switch (key * 64 + split) {")
        for cls, kind, n, _ in members:
            if kind == "fused":
                continue
            for j in range(n):
                name = f"eri_split_{''.join(map(str, cls))}_{j}"
                o.append(f"    case {_key(cls) * 64 + j}:")
                o += _launch_lines(cls, name)
        o.append("    default:")
        o.append(f"      return {NOT_HERE};")
        o.append("  }")
        o.append("}")
        o.append("")
        o.append(f"int launch_hrr_part{k}(const gbx_tables* t, int key, const RegArgs& a, "
                 "hipStream_t s) {")
        o.append("  const unsigned grid = static_cast<unsigned>((a.nquart + kRegThreads - 1) / "
                 "kRegThreads);")
        o.append("  (void)grid;")
        o.append("  switch (key) {")
        for cls, kind, _, _ in members:
            if kind == "fused" or not _hrr_reg(cls):
                continue
            name = f"eri_hrr_{''.join(map(str, cls))}"
            o.append(f"    case {_key(cls)}:")
            o.append(f"      hipLaunchKernelGGL({name}, dim3(grid), dim3(kRegThreads), 0, s, "
                     "*t, a);")
            o.append(f'      return check_launch("{name}");')
        o.append("    default:")
        o.append(f"      return {NOT_HERE};")
        o.append("  }")
        o.append("}")
        o.append("")
        o.append("}  // namespace gbx")
        out.append("\n".join(o) + "\n")
    return out


def _write(path, src):
    if os.path.exists(path) and open(path).read() == src:
        return
    with open(path, "w") as fh:
        fh.write(src)


def _stamp():
    """Hash of everything the generated sources depend on: this generator, the class
    programs and the settings read from the environment."""
    import hashlib

    h = hashlib.sha1()
    here = os.path.dirname(os.path.abspath(__file__))
    for name in ("codegen.py", "programs.py"):
        with open(os.path.join(here, name), "rb") as fh:
            h.update(fh.read())
    h.update(repr((REG_BUDGET, SPLIT_BUDGET, sorted(FORCE_SPLIT), BOYS_LDS, LOOP_SWAP, _U,
                   F_SHELLS)).encode())
    return h.hexdigest()[:16]


def up_to_date(path=OUT):
    """True if ``path`` and every part were generated from the current inputs."""
    base = path[:-len(".hip")]
    if not all(os.path.exists(p) for p in [path] + [f"{base}_{k}.hip" for k in range(N_PARTS)]):
        return False
    with open(path) as fh:
        fh.readline()
        return fh.readline().strip() == f"// inputs {_stamp()}"


def generate(path=OUT):
    """Write the kernel parts gbx_reg_gen_<k>.hip and the dispatch file ``path``."""
    classes = _reg_classes()
    base = path[:-len(".hip")]
    for k, src in enumerate(_part_sources(classes)):
        _write(f"{base}_{k}.hip", src)
    fused = [c for c in classes if class_plan(c)[0] == "fused"]
    split = [(c, len(class_plan(c)[1])) for c in classes if class_plan(c)[0] != "fused"]
    hrr = [c for c, _ in split if _hrr_reg(c)]
    m = ["// GENERATED by gbasis_amd/codegen.py -- do not edit.",
         f"// inputs {_stamp()}",
         f"// Dispatch over the {N_PARTS} parts of the generated register kernels.",
         '#include "gbx_reg.h"', "", "namespace gbx {", ""]
    for k in range(N_PARTS):
        m.append(f"int launch_reg_part{k}(const gbx_tables* t, int key, const RegArgs& a, "
                 "int precision, hipStream_t s);")
        m.append(f"int launch_split_part{k}(const gbx_tables* t, int key, int split, "
                 "const RegArgs& a, int precision, hipStream_t s);")
        m.append(f"int launch_hrr_part{k}(const gbx_tables* t, int key, const RegArgs& a, "
                 "hipStream_t s);")
    m.append("")
    m.append("int launch_reg_generated(const gbx_tables* t, int la, int lb, int lc, int ld, "
             "const RegArgs& a, int precision, hipStream_t s) {")
    m.append("  const int key = ((la * 8 + lb) * 8 + lc) * 8 + ld;")
    m.append("  int rc;")
    for k in range(N_PARTS):
        m.append(f"  if ((rc = launch_reg_part{k}(t, key, a, precision, s)) != {NOT_HERE}) "
                 "return rc;")
    m.append('  return fail(GBX_ERR_UNSUPPORTED, "no fused register kernel for this class");')
    m.append("}")
    m.append("")
    m.append("int launch_split_generated(const gbx_tables* t, int la, int lb, int lc, int ld, "
             "int k, const RegArgs& a, int precision, hipStream_t s) {")
    m.append("  const int key = ((la * 8 + lb) * 8 + lc) * 8 + ld;")
    m.append("  int rc;")
    for k in range(N_PARTS):
        m.append(f"  if ((rc = launch_split_part{k}(t, key, k, a, precision, s)) != {NOT_HERE}) "
                 "return rc;")
    m.append('  return fail(GBX_ERR_UNSUPPORTED, "no split register kernel for this class");')
    m.append("}")
    m.append("")
    m.append("int launch_hrr_generated(const gbx_tables* t, int la, int lb, int lc, int ld, "
             "const RegArgs& a, hipStream_t s) {")
    m.append("  const int key = ((la * 8 + lb) * 8 + lc) * 8 + ld;")
    m.append("  int rc;")
    for k in range(N_PARTS):
        m.append(f"  if ((rc = launch_hrr_part{k}(t, key, a, s)) != {NOT_HERE}) return rc;")
    m.append('  return fail(GBX_ERR_UNSUPPORTED, "no register HRR kernel for this class");')
    m.append("}")
    m.append("")
    for fname, members in (("has_hrr_generated", hrr), ("has_reg_generated", fused)):
        m.append(f"bool {fname}(int la, int lb, int lc, int ld) {{")
        m.append("  switch (((la * 8 + lb) * 8 + lc) * 8 + ld) {")
        for cls in members:
            m.append(f"    case {_key(cls)}:")
        m.append("      return true;")
        m.append("    default:")
        m.append("      return false;")
        m.append("  }")
        m.append("}")
        m.append("")
    m.append("int split_count_generated(int la, int lb, int lc, int ld) {")
    m.append("  switch (((la * 8 + lb) * 8 + lc) * 8 + ld) {")
    for cls, n in split:
        m.append(f"    case {_key(cls)}:")
        m.append(f"      return {n};")
    m.append("    default:")
    m.append("      return 0;")
    m.append("  }")
    m.append("}")
    m.append("")
    m.append("}  // namespace gbx")
    _write(path, "\n".join(m) + "\n")
    return classes


if __name__ == "__main__":
    if "--force" not in sys.argv and up_to_date():
        print(f"register kernels up to date ({OUT})")
    else:
        cl = generate()
        print(f"generated register kernels for {len(cl)} classes -> {OUT}")
