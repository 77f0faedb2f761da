// Cooperative (workgroup-per-shell-quartet) ERI kernel for gfx950, Boys device
// function and the C ABI declared in include/gbasis_amd.h.
//
// Algorithm: Obara-Saika vertical recursion on A, Obara-Saika build of C,
// contraction over primitive quartets into FP64 accumulators, Head-Gordon-Pople
// horizontal recursion on D then B (FP64 in both precision modes), factorial2 component norms, per-shell
// output transform (norm_cont, component order, Cartesian->pure) and placement
// into the canonical 8-fold packed array -- the stage order of
// gbasis/integrals/_two_elec_int_improved.py:591-732 and
// gbasis/base_four_symm.py:160-389.
//
// The recursion is driven by a per-class "program" (lists of independent items
// per phase, built on the host, with storage slots reused once a value is dead).
// One workgroup (several wavefronts) owns one shell
// quartet at a time and processes PF primitive quartets at once: every intermediate is stored
// slot-major with the PF copies adjacent, so lanes working on the same item for
// consecutive copies touch consecutive LDS words, and a phase boundary is a
// workgroup barrier.  Several such workgroups share a CU; each runs 4 to 16
// wavefronts, so that the few of them that fit in a CU's LDS still keep its four
// SIMDs busy.  Classes whose working set does not
// fit in the LDS one workgroup may allocate (160 KiB on gfx950) run the same code
// with 256-lane workgroups on a per-workgroup slice of a global scratch buffer,
// with few enough workgroups that the slices in use stay in the Infinity Cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

constexpr int kCoopThreads = 64;
constexpr int kCoopGlobalThreads = 256;  // lanes of a global-scratch workgroup
constexpr int kPhVrr = 0, kPhKvrr = 1, kPhAcc = 2, kPhHrrK = 3, kPhHrrB = 4;
constexpr int kScalars = 17;  // per-primitive-quartet scalars held in LDS
constexpr int kNoSlot = 0xFFFF;  // absent source of a recursion item

struct CoopArgs {
  const int32_t* quart;
  int64_t nquart;
  const int32_t* count;  // device quartet count (<= nquart); null: nquart
  const int32_t* items;
  const int32_t* phases;
  const int32_t* out_slots;
  int32_t prim_slots, nacc, h_slots, nout, n_prim_ph, n_ph, L;
  int32_t la, lb, lc, ld;
  int32_t pf, mode;
  double* out;
  char* scratch;
  int64_t stride;
  int64_t sc_off, acc_off, un_off, h_off, f0_off, f1_off;
  const double* acc_in;  // non-null: partial contracted [e0|f0] given,
                         //   acc_in[(f * acc_parts + r) * acc_stride + qi], summed over r
  int64_t acc_stride;
  int32_t acc_parts;
  double* acc_out;       // non-null: primitive pass only; work item (qi, r) sums primitive
                         //   quartets [r * chunk, (r + 1) * chunk) of quartet qi and writes
                         //   acc_out[(f * parts + r) * acc_stride + qi]
  int32_t parts;
};

struct CoopLayout {
  int64_t sc_off, acc_off, un_off, h_off, f0_off, f1_off, total;
  bool reg_acc;  // primitive pass keeps the accumulators in registers (no acc region)
};

inline int64_t align16(int64_t x) { return (x + 15) & ~int64_t(15); }

// Register accumulators of the primitive pass: kRegAccThreads lanes with
// kRegAccPerLane FP64 accumulators each.
constexpr int kRegAccThreads = 512;
constexpr int kRegAccPerLane = 24;

inline int64_t coop_lds_limit();

inline CoopLayout coop_layout_acc(const int32_t* hdr, int tb_bytes, int pf, bool prim_only,
                                  bool reg_acc) {
  const int64_t prim_slots = hdr[0], nacc = hdr[1], h_slots = hdr[2], nout = hdr[3];
  CoopLayout l{};
  l.reg_acc = reg_acc;
  l.sc_off = 0;
  l.acc_off = align16(l.sc_off + kScalars * pf * tb_bytes);
  l.un_off = reg_acc ? l.acc_off : align16(l.acc_off + nacc * 8);
  const int64_t prim_bytes = prim_slots * pf * tb_bytes;
  l.h_off = l.un_off;
  l.f0_off = align16(l.h_off + h_slots * 8);
  l.f1_off = align16(l.f0_off + nout * 8);
  const int64_t epi_end = align16(l.f1_off + nout * 8);
  l.total = prim_only ? align16(l.un_off + prim_bytes)
                      : std::max(align16(l.un_off + prim_bytes), epi_end);
  return l;
}

// Workspace of one workgroup.  The primitive pass (prim_only) moves its
// accumulators to registers when they would push the workspace out of LDS and
// the primitive store alone fits.
inline CoopLayout coop_layout(const int32_t* hdr, int tb_bytes, int pf, bool prim_only = false) {
  const CoopLayout l = coop_layout_acc(hdr, tb_bytes, pf, prim_only, false);
  if (!prim_only || l.total <= coop_lds_limit() ||
      hdr[1] > int64_t(kRegAccThreads) * kRegAccPerLane)
    return l;
  const CoopLayout r = coop_layout_acc(hdr, tb_bytes, pf, prim_only, true);
  return r.total <= coop_lds_limit() ? r : l;
}

__device__ inline uint64_t tri(uint64_t i, uint64_t j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

template <typename TB, bool kLDS, int NT, bool kRegAcc = false>
__global__ __launch_bounds__(NT) void eri_coop_kernel(gbx_tables t, CoopArgs a) {
  static_assert(!kRegAcc || NT == kRegAccThreads, "register accumulators need 512 lanes");
  double accr[kRegAcc ? kRegAccPerLane : 1];
  extern __shared__ __align__(16) char smem[];
  char* base = kLDS ? smem : (a.scratch + static_cast<int64_t>(blockIdx.x) * a.stride);
  TB* sc = reinterpret_cast<TB*>(base + a.sc_off);
  double* acc = reinterpret_cast<double*>(base + a.acc_off);
  TB* S = reinterpret_cast<TB*>(base + a.un_off);
  double* H = reinterpret_cast<double*>(base + a.h_off);
  double* F0 = reinterpret_cast<double*>(base + a.f0_off);
  double* F1 = reinterpret_cast<double*>(base + a.f1_off);
  __shared__ double s_red[NT / 64];
  __shared__ double s_red2[NT / 64];

  const int tid = threadIdx.x;
  const int PF = a.pf;  // a power of two (checked on the host)
  const int pf_shift = 31 - __builtin_clz(static_cast<unsigned>(PF));
  const int4* items4 = reinterpret_cast<const int4*>(a.items);
  const int4* pinfo = reinterpret_cast<const int4*>(t.pair_info);

  const int64_t nquart = a.count ? min(a.nquart, static_cast<int64_t>(*a.count)) : a.nquart;
  const int parts = a.acc_out ? a.parts : 1;
  for (int64_t wi = xcd_chunked(blockIdx.x, gridDim.x); wi < nquart * parts; wi += gridDim.x) {
    const int64_t qi = wi % nquart;
    const int part = static_cast<int>(wi / nquart);
    const int bra = a.quart[2 * qi], ket = a.quart[2 * qi + 1];
    const int4 pb = pinfo[bra];
    const int4 pk = pinfo[ket];
    const int nk = pk.w;
    const int nprimq = pb.w * nk;
    const int chunk = (nprimq + parts - 1) / parts;
    const int pbeg = min(nprimq, part * chunk);
    const int pend = min(nprimq, pbeg + chunk);

    __syncthreads();  // previous quartet's epilogue is done with the workspace
    if (a.acc_in) {
      for (int f = tid; f < a.nacc; f += NT) {
        double sum = 0.0;
        for (int r = 0; r < a.acc_parts; ++r)
          sum += a.acc_in[(static_cast<int64_t>(f) * a.acc_parts + r) * a.acc_stride + qi];
        acc[f] = sum;
      }
    } else if constexpr (kRegAcc) {
#pragma unroll
      for (int u = 0; u < kRegAccPerLane; ++u) accr[u] = 0.0;
    } else {
      for (int f = tid; f < a.nacc; f += NT) acc[f] = 0.0;
    }

    for (int p0 = pbeg; p0 < (a.acc_in ? pbeg : pend); p0 += PF) {
      const int ncopy = min(PF, pend - p0);
      __syncthreads();
      if (tid < ncopy) {
        const int p = p0 + tid;
        const int ib = p / nk, ik = p - ib * nk;
        const double* rb = t.prim_pairs + static_cast<int64_t>(pb.z + ib) * kPrimRec;
        const double* rk = t.prim_pairs + static_cast<int64_t>(pk.z + ik) * kPrimRec;
        PrimQuartet<double> s;
        const double Tval = prim_quartet_setup(rb, rk, s);
        sc[0 * PF + tid] = static_cast<TB>(s.PA[0]);
        sc[1 * PF + tid] = static_cast<TB>(s.PA[1]);
        sc[2 * PF + tid] = static_cast<TB>(s.PA[2]);
        sc[3 * PF + tid] = static_cast<TB>(s.WP[0]);
        sc[4 * PF + tid] = static_cast<TB>(s.WP[1]);
        sc[5 * PF + tid] = static_cast<TB>(s.WP[2]);
        sc[6 * PF + tid] = static_cast<TB>(s.roz);
        sc[7 * PF + tid] = static_cast<TB>(s.i2z);
        sc[8 * PF + tid] = static_cast<TB>(s.QC[0]);
        sc[9 * PF + tid] = static_cast<TB>(s.QC[1]);
        sc[10 * PF + tid] = static_cast<TB>(s.QC[2]);
        sc[11 * PF + tid] = static_cast<TB>(s.WQ[0]);
        sc[12 * PF + tid] = static_cast<TB>(s.WQ[1]);
        sc[13 * PF + tid] = static_cast<TB>(s.WQ[2]);
        sc[14 * PF + tid] = static_cast<TB>(s.roe);
        sc[15 * PF + tid] = static_cast<TB>(s.i2e);
        sc[16 * PF + tid] = static_cast<TB>(s.i2s);
        const double pref = s.pref;
        if (t.potential == GBX_POT_COULOMB || a.mode == GBX_MODE_SCHWARZ) {
          boys_eval(Tval, a.L, t,
                    [&](int m, double v) { S[m * PF + tid] = static_cast<TB>(v * pref); });
        } else {
          // range-separated operators (FP64 only, checked on the host):
          // G_m = att^(m+1/2) F_m(att T), att = omega^2 / (omega^2 + rho)
          const double w2 = t.omega * t.omega;
          const double rho = rb[0] * rk[0] / (rb[0] + rk[0]);
          const double att = w2 / (w2 + rho);
          const double att_sqrt = sqrt(att);
          const bool complement = t.potential == GBX_POT_ERFC;
          if (complement)
            boys_eval(Tval, a.L, t,
                      [&](int m, double v) { S[m * PF + tid] = static_cast<TB>(v * pref); });
          boys_eval(att * Tval, a.L, t, [&](int m, double v) {
            double pw = att_sqrt;
            for (int i = 0; i < m; ++i) pw *= att;
            const double g = v * pw * pref;
            S[m * PF + tid] =
                complement ? static_cast<TB>(S[m * PF + tid] - g) : static_cast<TB>(g);
          });
        }
      }
      __syncthreads();
      // NT is a multiple of PF, so a thread works on the same primitive copy c in every
      // phase: the recursion scalars of the current phase kind live in registers
      // (A-centre VRR: P-A, W-P, rho/zeta, 1/(2 zeta); C-centre build: Q-C, W-Q,
      // rho/eta, 1/(2 eta), 1/(2 (zeta + eta)))
      const int c = tid & (PF - 1);
      const bool active = c < ncopy;
      TB r_p[3] = {0, 0, 0}, r_w[3] = {0, 0, 0};
      TB r_ro = 0, r_i2 = 0, r_i2s = 0;
      int loaded = -1;
      for (int ph = 0; ph < a.n_prim_ph; ++ph) {
        const int kind = a.phases[3 * ph], i0 = a.phases[3 * ph + 1], i1 = a.phases[3 * ph + 2];
        const int nwork = (i1 - i0) << pf_shift;
        if (active) {
          if (kind != loaded) {
            const int o = kind == kPhVrr ? 0 : 8;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              r_p[i] = sc[(o + i) * PF + c];
              r_w[i] = sc[(o + 3 + i) * PF + c];
            }
            r_ro = sc[(o + 6) * PF + c];
            r_i2 = sc[(o + 7) * PF + c];
            r_i2s = sc[16 * PF + c];
            loaded = kind;
          }
          for (int x = tid; x < nwork; x += NT) {
            const int it = i0 + (x >> pf_shift);
            // one 16-byte item: 16-bit slots dst | s0, s1 | s2, s3 | s4 (0xFFFF: none), code
            const int4 d = items4[it];
            const uint32_t w0 = static_cast<uint32_t>(d.x), w1 = static_cast<uint32_t>(d.y),
                           w2 = static_cast<uint32_t>(d.z);
            const int dst = w0 & 0xFFFF, s0 = w0 >> 16, s1 = w1 & 0xFFFF, s2 = w1 >> 16;
            const int s3 = w2 & 0xFFFF, s4 = w2 >> 16;
            const int code = d.w;
            const int dir = code & 15;
            const TB k1 = static_cast<TB>((code >> 4) & 255);
            const TB pd = dir == 0 ? r_p[0] : (dir == 1 ? r_p[1] : r_p[2]);
            const TB wd = dir == 0 ? r_w[0] : (dir == 1 ? r_w[1] : r_w[2]);
            // A centre: [a+1_i]^(m) = PA_i [a]^(m) + WP_i [a]^(m+1)
            //   + a_i/(2z) ([a-1_i]^(m) - rho/z [a-1_i]^(m+1))
            // C centre: [a|c+1_i]^(m) = QC_i [a|c]^(m) + WQ_i [a|c]^(m+1)
            //   + c_i/(2e) ([a|c-1_i]^(m) - rho/e [a|c-1_i]^(m+1)) + a_i/(2(z+e)) [a-1_i|c]^(m+1)
            // slot * PF as a shift (PF is a power of two)
            TB v = pd * S[(s0 << pf_shift) + c] + wd * S[(s1 << pf_shift) + c];
            if (s2 != kNoSlot)
              v += k1 * r_i2 * (S[(s2 << pf_shift) + c] - r_ro * S[(s3 << pf_shift) + c]);
            if (s4 != kNoSlot) v += static_cast<TB>(code >> 12) * r_i2s * S[(s4 << pf_shift) + c];
            S[(dst << pf_shift) + c] = v;
          }
        }
        __syncthreads();
      }
      {
        const int i0 = a.phases[3 * a.n_prim_ph + 1];
        if constexpr (kRegAcc) {
#pragma unroll
          for (int u = 0; u < kRegAccPerLane; ++u) {
            const int f = tid + u * NT;
            if (f < a.nacc) {
              const int src = a.items[4 * (i0 + f) + 1];
              double sum = 0.0;
              for (int c = 0; c < ncopy; ++c) sum += static_cast<double>(S[src * PF + c]);
              accr[u] += sum;
            }
          }
        } else {
          for (int f = tid; f < a.nacc; f += NT) {
            const int src = a.items[4 * (i0 + f) + 1];
            double sum = 0.0;
            for (int c = 0; c < ncopy; ++c) sum += static_cast<double>(S[src * PF + c]);
            acc[f] += sum;
          }
        }
      }
    }
    if (a.acc_out) {
      // primitive pass only: each thread wrote acc[f] for its own f; the accumulation
      // item of f names its row of the class's accumulators (a program may produce a
      // subset of them)
      const int32_t* acc_items = a.items + 4 * a.phases[3 * a.n_prim_ph + 1];
      if constexpr (kRegAcc) {
#pragma unroll
        for (int u = 0; u < kRegAccPerLane; ++u) {
          const int f = tid + u * NT;
          if (f < a.nacc)
            a.acc_out[(static_cast<int64_t>(acc_items[4 * f]) * parts + part) * a.acc_stride +
                      qi] = accr[u];
        }
      } else {
        for (int f = tid; f < a.nacc; f += NT)
          a.acc_out[(static_cast<int64_t>(acc_items[4 * f]) * parts + part) * a.acc_stride +
                    qi] = acc[f];
      }
      continue;
    }
    // register accumulators serve the primitive pass only (acc_out is then set)
    if constexpr (!kRegAcc) {
    __syncthreads();

    // ---- horizontal recursion on the contracted values ----
    const double ABx = t.pair_geom[4 * bra], ABy = t.pair_geom[4 * bra + 1],
                 ABz = t.pair_geom[4 * bra + 2];
    const double CDx = t.pair_geom[4 * ket], CDy = t.pair_geom[4 * ket + 1],
                 CDz = t.pair_geom[4 * ket + 2];
    for (int f = tid; f < a.nacc; f += NT) H[f] = acc[f];
    __syncthreads();
    for (int ph = a.n_prim_ph + 1; ph < a.n_ph; ++ph) {
      const int kind = a.phases[3 * ph], i0 = a.phases[3 * ph + 1], i1 = a.phases[3 * ph + 2];
      const double gx = kind == kPhHrrK ? CDx : ABx;
      const double gy = kind == kPhHrrK ? CDy : ABy;
      const double gz = kind == kPhHrrK ? CDz : ABz;
      for (int it = i0 + tid; it < i1; it += NT) {
        const int4 d0 = items4[it];  // dst, src1, src2, direction
        const int dir = d0.w;
        const double g = dir == 0 ? gx : (dir == 1 ? gy : gz);
        H[d0.x] = H[d0.y] + g * H[d0.z];
      }
      __syncthreads();
    }

    // ---- epilogue: factorial2 norms, output transforms, placement ----
    const int nl[4] = {a.la, a.lb, a.lc, a.ld};
    int nc[4];
    for (int x = 0; x < 4; ++x) nc[x] = (nl[x] + 1) * (nl[x] + 2) / 2;
    const double* f2a = t.f2norm + f2_offset(a.la);
    const double* f2b = t.f2norm + f2_offset(a.lb);
    const double* f2c = t.f2norm + f2_offset(a.lc);
    const double* f2d = t.f2norm + f2_offset(a.ld);
    for (int o = tid; o < a.nout; o += NT) {
      int r = o;
      const int id = r % nc[3]; r /= nc[3];
      const int ic = r % nc[2]; r /= nc[2];
      const int ib = r % nc[1];
      const int ia = r / nc[1];
      F0[o] = H[a.out_slots[o]] * f2a[ia] * f2b[ib] * f2c[ic] * f2d[id];
    }
    __syncthreads();

    if (a.mode == GBX_MODE_SCHWARZ) {
      // out[2 bra]: sqrt(max |(ij|ij)|); out[2 bra + 1]: sqrt(max |contracted [e0|f0]|)
      double m = 0.0, m2 = 0.0;
      for (int o = tid; o < a.nout; o += NT) m = fmax(m, fabs(F0[o]));
      for (int f = tid; f < a.nacc; f += NT) m2 = fmax(m2, fabs(acc[f]));
      for (int off = 32; off > 0; off >>= 1) {
        m = fmax(m, __shfl_xor(m, off, 64));
        m2 = fmax(m2, __shfl_xor(m2, off, 64));
      }
      if ((tid & 63) == 0) {
        s_red[tid >> 6] = m;
        s_red2[tid >> 6] = m2;
      }
      __syncthreads();
      if (tid == 0) {
        double mm = s_red[0], mm2 = s_red2[0];
        for (int w = 1; w < NT / 64; ++w) {
          mm = fmax(mm, s_red[w]);
          mm2 = fmax(mm2, s_red2[w]);
        }
        a.out[2 * bra] = sqrt(mm);
        a.out[2 * bra + 1] = sqrt(mm2);
      }
      continue;
    }

    const int sh[4] = {pb.x, pb.y, pk.x, pk.y};
    int4 sinfo[4][2];
    for (int x = 0; x < 4; ++x) {
      const int4* si = reinterpret_cast<const int4*>(t.shell_info) + 2 * sh[x];
      sinfo[x][0] = si[0];
      sinfo[x][1] = si[1];
    }
    // dims of the current block; dense shells are transformed in place (ping-pong)
    int dim[4] = {nc[0], nc[1], nc[2], nc[3]};
    double* cur = F0;
    double* nxt = F1;
    for (int x = 0; x < 4; ++x) {
      if (sinfo[x][0].z != 1) continue;  // kind 0: applied at placement
      const int nout_x = sinfo[x][1].y;
      const double* M = t.out_mat + sinfo[x][1].x;
      int ndim[4] = {dim[0], dim[1], dim[2], dim[3]};
      ndim[x] = nout_x;
      int inner = 1;
      for (int y = x + 1; y < 4; ++y) inner *= dim[y];
      const int total = ndim[0] * ndim[1] * ndim[2] * ndim[3];
      for (int e = tid; e < total; e += NT) {
        const int lo = e % inner;
        const int jx = (e / inner) % nout_x;
        const int hi = e / (inner * nout_x);
        const double* src = cur + static_cast<int64_t>(hi) * dim[x] * inner + lo;
        const double* mrow = M + jx * dim[x];
        double s = 0.0;
        for (int k = 0; k < dim[x]; ++k) s += mrow[k] * src[k * inner];
        nxt[e] = s;
      }
      __syncthreads();
      dim[x] = nout_x;
      double* tmp = cur;
      cur = nxt;
      nxt = tmp;
    }

    const uint64_t obase = t.pair_out_base[bra];
    // bra and ket entries on the same two shells (a shell pair's (ij|ij) quartet or one
    // of its sub-quartets): only the ij >= kl half is written
    const bool bra_diag = pb.x == pb.y, ket_diag = pk.x == pk.y;
    const bool same_pair = (pb.x == pk.x && pb.y == pk.y) || (pb.x == pk.y && pb.y == pk.x);
    const bool add = a.mode == GBX_MODE_PACKED_ADD;
    const int total = dim[0] * dim[1] * dim[2] * dim[3];
    for (int e = tid; e < total; e += NT) {
      int idx[4];
      int r = e;
      idx[3] = r % dim[3]; r /= dim[3];
      idx[2] = r % dim[2]; r /= dim[2];
      idx[1] = r % dim[1];
      idx[0] = r / dim[1];
      double v = cur[e];
      uint32_t fn[4];
      for (int x = 0; x < 4; ++x) {
        int pos = idx[x];
        if (sinfo[x][0].z == 0) {
          pos = t.comp_map[sinfo[x][0].w + idx[x]];
          v *= t.comp_scale[sinfo[x][0].w + idx[x]];
        }
        fn[x] = static_cast<uint32_t>(sinfo[x][0].x + pos);
      }
      if (bra_diag && fn[0] < fn[1]) continue;
      if (ket_diag && fn[2] < fn[3]) continue;
      const uint64_t ij = tri(fn[0], fn[1]), kl = tri(fn[2], fn[3]);
      if (same_pair && ij < kl) continue;
      if (add)
        a.out[obase + tri(ij, kl)] += v;
      else
        a.out[obase + tri(ij, kl)] = v;
    }
    }  // !kRegAcc
  }
}

// Largest dynamic LDS allocation of one workgroup on the current device (160 KiB on
// gfx950), less the kernel's static LDS; cached per device.
inline int64_t coop_lds_limit() {
  static int64_t cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return kCoopLdsLimit;
  if (cached[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
      v = static_cast<int>(kCoopLdsLimit);
    cached[dev] = std::max<int64_t>(kCoopLdsLimit, static_cast<int64_t>(v) - 256);
  }
  return cached[dev];
}

// Compute units of the current device; cached per device.
inline int64_t coop_num_cu() {
  static int64_t cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
      v = 256;
    cached[dev] = v;
  }
  return cached[dev];
}

// Workgroups per CU of the global-scratch kernel (environment variable
// GBX_COOP_GLOBAL_WGS, 1..16, default 2).  Its workspaces are 0.2-2 MB; a few
// 256-lane workgroups per CU keep the scratch in flight within the 256 MB
// Infinity Cache instead of streaming it through HBM.
inline int coop_global_wgs_per_cu() {
  static int cached = 0;
  if (cached == 0) {
    const char* e = std::getenv("GBX_COOP_GLOBAL_WGS");
    const int v = e ? std::atoi(e) : 2;
    cached = (v >= 1 && v <= 16) ? v : 2;
  }
  return cached;
}

// Workgroups with LDS workspaces above this size run several wavefronts; smaller ones
// run one (environment variable GBX_COOP_WIDE_KB, default 0: every LDS workgroup is
// wide, which measured 14 % faster on the s-g workload than one-wavefront
// workgroups up to 40 KiB).
inline int64_t coop_wide_bytes() {
  static int64_t cached = -1;
  if (cached < 0) {
    const char* e = std::getenv("GBX_COOP_WIDE_KB");
    const int v = e ? std::atoi(e) : 0;
    cached = int64_t(v >= 0 && v <= 160 ? v : 0) * 1024;
  }
  return cached;
}

// Wavefronts per CU that a wide workgroup size aims for (environment variable
// GBX_COOP_WAVES, 4..32, default 16: 10.2 s on the s-g workload against 11.1 s
// with 8).
inline int coop_target_waves() {
  static int cached = 0;
  if (cached == 0) {
    const char* e = std::getenv("GBX_COOP_WAVES");
    const int v = e ? std::atoi(e) : 16;
    cached = (v >= 4 && v <= 32) ? v : 16;
  }
  return cached;
}

// Threads per workgroup for an LDS workspace of `bytes`: one wavefront up to
// coop_wide_bytes(), else enough wavefronts that the workgroups fitting on a CU
// together run about coop_target_waves() of them (a power of two, 256..1024).
inline int coop_threads(int64_t bytes) {
  if (bytes <= coop_wide_bytes()) return 64;
  const int64_t per_cu = std::max<int64_t>(1, coop_lds_limit() / bytes);
  const int64_t waves = (coop_target_waves() + per_cu - 1) / per_cu;
  int nt = 256;
  while (nt < 1024 && nt < 64 * waves) nt *= 2;
  return nt;
}

template <typename TB, bool kLDS, int NT, bool kRegAcc = false>
int launch_coop_nt(const CoopArgs& a, const gbx_tables* t, int64_t grid, size_t shmem,
                   hipStream_t stream) {
  if (shmem > 65536) {
    // the attribute is set once per device (it is per-device state)
    static bool raised[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(GBX_ERR_HIP, "eri_coop_kernel: no device");
    if (dev < 0 || dev >= 64 || !raised[dev]) {
      if (hipFuncSetAttribute(reinterpret_cast<const void*>(&eri_coop_kernel<TB, kLDS, NT, kRegAcc>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(coop_lds_limit())) != hipSuccess)
        return fail(GBX_ERR_HIP, "eri_coop_kernel: could not raise the LDS limit");
      if (dev >= 0 && dev < 64) raised[dev] = true;
    }
  }
  hipLaunchKernelGGL((eri_coop_kernel<TB, kLDS, NT, kRegAcc>), dim3(static_cast<unsigned>(grid)),
                     dim3(NT), shmem, stream, *t, a);
  return check_launch("eri_coop_kernel");
}

template <typename TB, bool kLDS>
int launch_coop(const gbx_tables* t, const gbx_program* prog, int la, int lb, int lc, int ld,
                const int32_t* d_quartets, int64_t nquart, const int32_t* d_count, int pf,
                int mode, double* d_out,
                void* d_scratch, int64_t scratch_bytes, hipStream_t stream,
                const double* acc_in = nullptr, int64_t acc_stride = 0, int acc_parts = 1,
                double* acc_out = nullptr, int parts = 1) {
  const CoopLayout l = coop_layout(prog->header, sizeof(TB), pf, acc_out != nullptr);
  CoopArgs a{};
  a.quart = d_quartets;
  a.nquart = nquart;
  a.count = d_count;
  a.items = prog->items;
  a.phases = prog->phases;
  a.out_slots = prog->out_slots;
  a.prim_slots = prog->header[0];
  a.nacc = prog->header[1];
  a.h_slots = prog->header[2];
  a.nout = prog->header[3];
  a.n_prim_ph = prog->header[4];
  a.n_ph = prog->header[5];
  a.L = prog->header[6];
  a.la = la;
  a.lb = lb;
  a.lc = lc;
  a.ld = ld;
  a.pf = pf;
  a.mode = mode;
  a.out = d_out;
  a.sc_off = l.sc_off;
  a.acc_off = l.acc_off;
  a.un_off = l.un_off;
  a.h_off = l.h_off;
  a.f0_off = l.f0_off;
  a.f1_off = l.f1_off;
  a.acc_in = acc_in;
  a.acc_stride = acc_stride;
  a.acc_parts = acc_parts;
  a.acc_out = acc_out;
  a.parts = parts;
  int64_t grid = nquart * (acc_out ? parts : 1);
  size_t shmem = 0;
  if constexpr (kLDS) {
    if (l.total > coop_lds_limit())
      return fail(GBX_ERR_ARG, "gbx_eri_coop: workspace exceeds the LDS of one workgroup");
    shmem = static_cast<size_t>(l.total);
    grid = std::min<int64_t>(grid, int64_t(1) << 20);
    if (l.reg_acc) {
      if (!acc_out)
        return fail(GBX_ERR_ARG, "gbx_eri_coop: register accumulators need the primitive pass");
      return launch_coop_nt<TB, kLDS, kRegAccThreads, true>(a, t, grid, shmem, stream);
    }
    switch (coop_threads(l.total)) {
      case 1024: return launch_coop_nt<TB, kLDS, 1024>(a, t, grid, shmem, stream);
      case 512: return launch_coop_nt<TB, kLDS, 512>(a, t, grid, shmem, stream);
      case 256: return launch_coop_nt<TB, kLDS, 256>(a, t, grid, shmem, stream);
      default: return launch_coop_nt<TB, kLDS, 64>(a, t, grid, shmem, stream);
    }
  } else {
    a.scratch = static_cast<char*>(d_scratch);
    a.stride = l.total;
    const int64_t nblk = scratch_bytes / l.total;
    if (nblk < 1)
      return fail(GBX_ERR_ARG, "gbx_eri_coop: scratch buffer smaller than one workgroup slice");
    grid = std::min<int64_t>({grid, nblk, coop_global_wgs_per_cu() * coop_num_cu()});
    return launch_coop_nt<TB, kLDS, kCoopGlobalThreads>(a, t, grid, shmem, stream);
  }
}

// Angular momenta of a class in the entries' own order (either shell of a pair may
// carry the larger l: sub-pairs built on their second centre are reversed).
inline bool class_in_range(int la, int lb, int lc, int ld) {
  return la >= 0 && lb >= 0 && lc >= 0 && ld >= 0 && la <= kMaxShellL && lb <= kMaxShellL &&
         lc <= kMaxShellL && ld <= kMaxShellL;
}

// The operator of a launch: Coulomb always; range-separated operators in FP64
// packed launches with omega > 0.
inline int check_potential(const gbx_tables* t, int mode, int precision, const char* who) {
  if (t->potential == GBX_POT_COULOMB || mode == GBX_MODE_SCHWARZ) return GBX_OK;
  if (t->potential != GBX_POT_ERF && t->potential != GBX_POT_ERFC)
    return fail(GBX_ERR_ARG, std::string(who) + ": unknown potential");
  if (!(t->omega > 0.0)) return fail(GBX_ERR_ARG, std::string(who) + ": omega must be > 0");
  if (precision != GBX_FP64)
    return fail(GBX_ERR_UNSUPPORTED,
                std::string(who) + ": range-separated operators are evaluated in FP64 only");
  return GBX_OK;
}

}  // namespace gbx

using namespace gbx;

extern "C" int gbx_coop_workspace(const gbx_program* prog, int32_t la, int32_t lb, int32_t lc,
                                  int32_t ld, int32_t precision, int32_t prim_copies,
                                  int32_t prim_only, int64_t* lds_bytes,
                                  int64_t* global_bytes_per_block) {
  if (!prog || prim_copies < 1 || prim_copies > kCoopThreads)
    return fail(GBX_ERR_ARG, "gbx_coop_workspace: prim_copies must be in [1, 64]");
  (void)la; (void)lb; (void)lc; (void)ld;
  const CoopLayout l =
      coop_layout(prog->header, precision == GBX_FP64 ? 8 : 4, prim_copies, prim_only != 0);
  const bool fits = l.total <= coop_lds_limit();
  if (lds_bytes) *lds_bytes = l.total;
  if (global_bytes_per_block) *global_bytes_per_block = fits ? 0 : l.total;
  return GBX_OK;
}

extern "C" int gbx_eri_coop(const gbx_tables* t, const gbx_program* prog, int32_t la, int32_t lb,
                            int32_t lc, int32_t ld, const int32_t* d_quartets, int64_t nquart,
                            const int32_t* d_count, int32_t precision, int32_t prim_copies,
                            int32_t mode, double* d_out, void* d_scratch, int64_t scratch_bytes,
                            void* stream) {
  if (!t || !prog || !d_out || (nquart > 0 && !d_quartets))
    return fail(GBX_ERR_ARG, "gbx_eri_coop: null argument");
  if (prim_copies < 1 || prim_copies > kCoopThreads || (prim_copies & (prim_copies - 1)))
    return fail(GBX_ERR_ARG, "gbx_eri_coop: prim_copies must be a power of two in [1, 64]");
  if (!class_in_range(la, lb, lc, ld))
    return fail(GBX_ERR_ARG, "gbx_eri_coop: angular momentum out of range");
  if (prog->header[6] != la + lb + lc + ld ||
      prog->header[6] + kTaylorTerms > t->boys_nord - 1)
    return fail(GBX_ERR_ARG, "gbx_eri_coop: program / Boys table do not match the class");
  if (mode != GBX_MODE_PACKED && mode != GBX_MODE_PACKED_ADD && mode != GBX_MODE_SCHWARZ)
    return fail(GBX_ERR_ARG, "gbx_eri_coop: bad mode");
  if (precision != GBX_FP64 && precision != GBX_FP32_BULK)
    return fail(GBX_ERR_ARG, "gbx_eri_coop: bad precision");
  if (int rc = check_potential(t, mode, precision, "gbx_eri_coop")) return rc;
  if (nquart == 0) return GBX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const CoopLayout l = coop_layout(prog->header, precision == GBX_FP64 ? 8 : 4, prim_copies);
  const bool lds = l.total <= coop_lds_limit();
  if (precision == GBX_FP64) {
    return lds ? launch_coop<double, true>(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count,
                                           prim_copies, mode, d_out, d_scratch, scratch_bytes, s)
               : launch_coop<double, false>(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count,
                                            prim_copies, mode, d_out, d_scratch, scratch_bytes, s);
  }
  return lds ? launch_coop<float, true>(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count,
                                        prim_copies, mode, d_out, d_scratch, scratch_bytes, s)
             : launch_coop<float, false>(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count,
                                         prim_copies, mode, d_out, d_scratch, scratch_bytes, s);
}

// ---- entry point that allocates nothing ----
// Global scratch of gbx_eri_coop_noalloc: a file-scope device array, one copy per GPU
// the library runs on, held until the process ends.  Host code never uses its name as
// an address: the device address of each GPU's copy comes from hipGetSymbolAddress,
// taken once per device with that device current.
constexpr int64_t kNoAllocScratchBytes = int64_t(128) << 20;   // 128 MiB
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (looked up on its first use on
// that device).
int noalloc_scratch(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(GBX_ERR_HIP, "gbx_eri_coop_noalloc: no current device (or device index >= 64)");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      return fail(GBX_ERR_HIP, std::string("gbx_eri_coop_noalloc: hipGetSymbolAddress of the "
                                           "scratch array failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return GBX_OK;
}
}  // namespace

extern "C" int gbx_eri_coop_noalloc(const gbx_tables* t, const gbx_program* prog, int32_t la,
                                    int32_t lb, int32_t lc, int32_t ld,
                                    const int32_t* d_quartets, int64_t nquart,
                                    const int32_t* d_count, int32_t precision,
                                    int32_t prim_copies, int32_t mode, double* d_out,
                                    void* stream) {
  if (!prog || prim_copies < 1 || prim_copies > kCoopThreads)
    return fail(GBX_ERR_ARG, "gbx_eri_coop_noalloc: bad program or prim_copies");
  if (precision != GBX_FP64 && precision != GBX_FP32_BULK)
    return fail(GBX_ERR_ARG, "gbx_eri_coop_noalloc: bad precision");
  const CoopLayout l = coop_layout(prog->header, precision == GBX_FP64 ? 8 : 4, prim_copies);
  if (l.total > coop_lds_limit() && l.total > kNoAllocScratchBytes)
    return fail(GBX_ERR_ARG, "gbx_eri_coop_noalloc: the workspace of one workgroup (" +
                                 std::to_string(l.total) + " bytes) exceeds the " +
                                 std::to_string(kNoAllocScratchBytes) + "-byte scratch array");
  void* sp = nullptr;
  if (int rc = noalloc_scratch(&sp)) return rc;
  return gbx_eri_coop(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count, precision,
                      prim_copies, mode, d_out, sp, kNoAllocScratchBytes, stream);
}

extern "C" int gbx_eri_hrr(const gbx_tables* t, const gbx_program* prog, int32_t la, int32_t lb,
                           int32_t lc, int32_t ld, const int32_t* d_quartets, int64_t nquart,
                           const int32_t* d_count, int32_t mode, const double* d_acc,
                           int64_t acc_stride, int32_t acc_parts, double* d_out, void* d_scratch,
                           int64_t scratch_bytes, void* stream) {
  if (!t || !prog || !d_out || (nquart > 0 && (!d_quartets || !d_acc)))
    return fail(GBX_ERR_ARG, "gbx_eri_hrr: null argument");
  if (!class_in_range(la, lb, lc, ld) || prog->header[6] != la + lb + lc + ld)
    return fail(GBX_ERR_ARG, "gbx_eri_hrr: program does not match the class");
  if (mode != GBX_MODE_PACKED && mode != GBX_MODE_PACKED_ADD && mode != GBX_MODE_SCHWARZ)
    return fail(GBX_ERR_ARG, "gbx_eri_hrr: bad mode");
  if (acc_stride < nquart || acc_parts < 1)
    return fail(GBX_ERR_ARG, "gbx_eri_hrr: acc_stride < nquart or acc_parts < 1");
  if (nquart == 0) return GBX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const CoopLayout l = coop_layout(prog->header, 8, 1);
  if (l.total <= coop_lds_limit())
    return launch_coop<double, true>(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count, 1, mode,
                                     d_out, d_scratch, scratch_bytes, s, d_acc, acc_stride,
                                     acc_parts);
  return launch_coop<double, false>(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count, 1, mode,
                                    d_out, d_scratch, scratch_bytes, s, d_acc, acc_stride,
                                    acc_parts);
}

extern "C" int gbx_eri_coop_acc(const gbx_tables* t, const gbx_program* prog, int32_t la,
                                int32_t lb, int32_t lc, int32_t ld, const int32_t* d_quartets,
                                int64_t nquart, const int32_t* d_count, int32_t precision,
                                int32_t prim_copies, int32_t parts, double* d_acc,
                                int64_t acc_stride, void* d_scratch, int64_t scratch_bytes,
                                void* stream) {
  if (!t || !prog || (nquart > 0 && (!d_quartets || !d_acc)))
    return fail(GBX_ERR_ARG, "gbx_eri_coop_acc: null argument");
  if (prim_copies < 1 || prim_copies > kCoopThreads || (prim_copies & (prim_copies - 1)))
    return fail(GBX_ERR_ARG, "gbx_eri_coop_acc: prim_copies must be a power of two in [1, 64]");
  if (parts < 1 || parts > 4096)
    return fail(GBX_ERR_ARG, "gbx_eri_coop_acc: parts must be in [1, 4096]");
  if (acc_stride < nquart) return fail(GBX_ERR_ARG, "gbx_eri_coop_acc: acc_stride < nquart");
  if (!class_in_range(la, lb, lc, ld))
    return fail(GBX_ERR_ARG, "gbx_eri_coop_acc: angular momentum out of range");
  if (prog->header[6] != la + lb + lc + ld ||
      prog->header[6] + kTaylorTerms > t->boys_nord - 1)
    return fail(GBX_ERR_ARG, "gbx_eri_coop_acc: program / Boys table do not match the class");
  if (precision != GBX_FP64 && precision != GBX_FP32_BULK)
    return fail(GBX_ERR_ARG, "gbx_eri_coop_acc: bad precision");
  if (int rc = check_potential(t, GBX_MODE_PACKED, precision, "gbx_eri_coop_acc")) return rc;
  if (nquart == 0) return GBX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const CoopLayout l =
      coop_layout(prog->header, precision == GBX_FP64 ? 8 : 4, prim_copies, true);
  const bool lds = l.total <= coop_lds_limit();
  if (precision == GBX_FP64)
    return lds ? launch_coop<double, true>(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count,
                                           prim_copies, GBX_MODE_PACKED, nullptr, d_scratch,
                                           scratch_bytes, s, nullptr, acc_stride, 1, d_acc, parts)
               : launch_coop<double, false>(t, prog, la, lb, lc, ld, d_quartets, nquart,
                                            d_count, prim_copies, GBX_MODE_PACKED, nullptr,
                                            d_scratch, scratch_bytes, s, nullptr, acc_stride, 1,
                                            d_acc, parts);
  return lds ? launch_coop<float, true>(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count,
                                        prim_copies, GBX_MODE_PACKED, nullptr, d_scratch,
                                        scratch_bytes, s, nullptr, acc_stride, 1, d_acc, parts)
             : launch_coop<float, false>(t, prog, la, lb, lc, ld, d_quartets, nquart, d_count,
                                         prim_copies, GBX_MODE_PACKED, nullptr, d_scratch,
                                         scratch_bytes, s, nullptr, acc_stride, 1, d_acc, parts);
}
