"""Generator of the per-class register kernels (gbasis_amd/csrc/gbx_reg_gen.hip).

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

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from gbasis_amd.programs import (PH_HRRK, PH_KVRR, PH_VRR, class_program, comps)
else:
    from .programs import PH_HRRK, PH_KVRR, PH_VRR, class_program, comps

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "gbx_reg_gen.hip")


def _reg_classes():
    """Kernel-orientation classes with l <= 2 per shell whose working set fits registers."""
    out = []
    for la in range(3):
        for lb in range(la + 1):
            for lc in range(3):
                for ld in range(lc + 1):
                    if (la + lb, la) < (lc + ld, lc):
                        continue
                    p = class_program(la, lb, lc, ld)
                    if p.nacc <= REG_MAX_ACC and p.prim_slots <= REG_MAX_SLOTS:
                        out.append((la, lb, lc, ld))
    return out


REG_MAX_ACC = 96
REG_MAX_SLOTS = 260


def _f2(comp):
    def df(n):
        r = 1
        while n > 1:
            r *= n
            n -= 2
        return r
    return 1.0 / (df(2 * comp[0] - 1) * df(2 * comp[1] - 1) * df(2 * comp[2] - 1)) ** 0.5


def kernel_source(cls):
    la, lb, lc, ld = cls
    p = class_program(*cls)
    L = p.L
    name = f"eri_reg_{la}{lb}{lc}{ld}"
    o = []
    w = o.append
    w(f"// class ({la}{lb}|{lc}{ld}): {p.prim_slots} primitive intermediates, "
      f"{p.nacc} accumulators, {p.nout} outputs")
    w("template <typename TB>")
    w(f"__global__ __launch_bounds__(kRegThreads) void {name}(gbx_tables t, RegArgs a) {{")
    w("  const int64_t q = static_cast<int64_t>(blockIdx.x) * kRegThreads + threadIdx.x;")
    w("  if (q >= a.nquart) return;")
    w("  const int bra = a.quart[2 * q], ket = a.quart[2 * q + 1];")
    w("  const int4 pb = reinterpret_cast<const int4*>(t.pair_info)[bra];")
    w("  const int4 pk = reinterpret_cast<const int4*>(t.pair_info)[ket];")
    for f in range(p.nacc):
        w(f"  double acc{f} = 0.0;")
    w("  for (int ib = 0; ib < pb.w; ++ib) {")
    w("    const double* rb = t.prim_pairs + static_cast<int64_t>(pb.z + ib) * kPrimRec;")
    w("    for (int ik = 0; ik < pk.w; ++ik) {")
    w("      const double* rk = t.prim_pairs + static_cast<int64_t>(pk.z + ik) * kPrimRec;")
    w("      PrimQuartet<TB> s;")
    w("      const double T = prim_quartet_setup(rb, rk, s);")
    w(f"      double F[{L + 1}];")
    w(f"      boys_eval_fixed<{L}>(T, t, F);")
    for m in range(L + 1):
        w(f"      const TB x{m} = static_cast<TB>(F[{m}] * s.pref);")
    # coefficient multiples used by the items
    k_vrr = set()
    k_kv1 = set()
    k_kv2 = set()
    for kind, items in p.prim_phases:
        for it in items:
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
    for kind, items in p.prim_phases:
        for it in items:
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
    for it in p.acc_phase[1]:
        w(f"      acc{it[0]} += static_cast<double>(x{it[1]});")
    w("    }")
    w("  }")
    # HRR
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
    for kind, items in p.hrr_phases:
        g = "CD" if kind == PH_HRRK else "AB"
        for it in items:
            w(f"  const double h{it[0]} = h{it[1]} + {g}[{it[3]}] * h{it[2]};")
    # outputs with factorial2 norms
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


def generate(path=OUT):
    classes = _reg_classes()
    parts = [
        "// GENERATED by gbasis_amd/codegen.py -- do not edit.",
        "// Per-class register kernels: one lane = one shell quartet.",
        '#include "gbx_reg.h"',
        "",
        "namespace gbx {",
        "",
    ]
    for cls in classes:
        parts.append(kernel_source(cls))
        parts.append("")
    parts.append("int launch_reg_generated(const gbx_tables* t, int la, int lb, int lc, int ld, "
                 "const RegArgs& a, int precision, hipStream_t s) {")
    parts.append("  const int key = ((la * 8 + lb) * 8 + lc) * 8 + ld;")
    parts.append("  const unsigned grid = static_cast<unsigned>((a.nquart + kRegThreads - 1) / "
                 "kRegThreads);")
    parts.append("  switch (key) {")
    for cls in classes:
        la, lb, lc, ld = cls
        key = ((la * 8 + lb) * 8 + lc) * 8 + ld
        name = f"eri_reg_{la}{lb}{lc}{ld}"
        parts.append(f"    case {key}:")
        parts.append(f"      if (precision == GBX_FP64) hipLaunchKernelGGL(({name}<double>), "
                     "dim3(grid), dim3(kRegThreads), 0, s, *t, a);")
        parts.append(f"      else hipLaunchKernelGGL(({name}<float>), dim3(grid), "
                     "dim3(kRegThreads), 0, s, *t, a);")
        parts.append("      return check_launch(\"" + name + "\");")
    parts.append("    default:")
    parts.append("      return fail(GBX_ERR_UNSUPPORTED, \"gbx_eri_reg: no register kernel for "
                 "this class\");")
    parts.append("  }")
    parts.append("}")
    parts.append("")
    parts.append("bool has_reg_generated(int la, int lb, int lc, int ld) {")
    parts.append("  const int key = ((la * 8 + lb) * 8 + lc) * 8 + ld;")
    parts.append("  switch (key) {")
    for cls in classes:
        la, lb, lc, ld = cls
        parts.append(f"    case {((la * 8 + lb) * 8 + lc) * 8 + ld}:")
    parts.append("      return true;")
    parts.append("    default:")
    parts.append("      return false;")
    parts.append("  }")
    parts.append("}")
    parts.append("")
    parts.append("}  // namespace gbx")
    src = "\n".join(parts) + "\n"
    if os.path.exists(path) and open(path).read() == src:
        return classes
    with open(path, "w") as fh:
        fh.write(src)
    return classes


if __name__ == "__main__":
    cl = generate()
    print(f"generated {len(cl)} register kernels -> {OUT}")
