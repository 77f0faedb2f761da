// Shared pieces of the per-class register kernels (one lane = one shell quartet):
// launch arguments, a fixed-order Boys evaluation and the block store.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

constexpr int kRegThreads = 256;

struct RegArgs {
  const int32_t* quart;  // (bra pair, ket pair) per quartet
  int64_t nquart;
  int32_t mode;          // GBX_MODE_PACKED, GBX_MODE_PACKED_ADD or GBX_MODE_SCHWARZ
  double* out;
  double* acc;           // split kernels: partial contracted [e0|f0],
                         //   acc[(f * lanes_per_quartet + r) * acc_stride + q]
  int64_t acc_stride;
  int32_t lanes_per_quartet;  // split kernels: lanes sharing one shell quartet (R)
  const int64_t* seg_off;     // fused kernels, segmented launch: first quartet of each of
                              //   kSegments segments, and the end of the last (capacity of
                              //   segment s: seg_off[s + 1] - seg_off[s])
  const int32_t* seg_cnt;     //   and its count (device); null: plain launch
  const int32_t* count;       // plain launches: device quartet count (<= nquart, the
                              //   capacity the grid is sized for); null: nquart
};

constexpr int kSegments = 16;

// Fused kernels: quartet index of this lane, or -1.  A segmented launch covers the
// primitive-count bins of one class in one grid: lane gid takes quartet
// seg_off[s] + (gid - sum_{s' < s} n[s']), n[s] = min(seg_cnt[s], capacity of s).
__device__ inline int64_t lane_quartet(const RegArgs& a, int64_t gid) {
  if (a.seg_cnt == nullptr) {
    const int64_t n = a.count ? min(a.nquart, static_cast<int64_t>(*a.count)) : a.nquart;
    return gid < n ? gid : -1;
  }
  int64_t base = 0;
#pragma unroll
  for (int s = 0; s < kSegments; ++s) {
    const int64_t c = min(static_cast<int64_t>(a.seg_cnt[s]), a.seg_off[s + 1] - a.seg_off[s]);
    if (gid < base + c) return a.seg_off[s] + (gid - base);
    base += c;
  }
  return -1;
}

// F_m(T), m = 0..L, into F[] (same algorithm as boys_eval, unrolled for fixed L).
template <int L>
__device__ inline void boys_eval_fixed(double T, const gbx_tables& t, double* F) {
  const double tmax = t.boys_h * (t.boys_nt - 1);
  const double inv_h = 1.0 / t.boys_h;  // uniform: computed once per kernel
  if (T <= tmax) {
    const int k = static_cast<int>(T * inv_h + 0.5);
    const double x = k * t.boys_h - T;
    const double* row = t.boys_table + static_cast<int64_t>(k) * t.boys_nord;
    const double emT = L > 0 ? row[0] * exp_small(x) : 0.0;
    row += 1 + L;
    double h = row[7];
    h = row[6] + x * h * (1.0 / 7.0);
    h = row[5] + x * h * (1.0 / 6.0);
    h = row[4] + x * h * (1.0 / 5.0);
    h = row[3] + x * h * (1.0 / 4.0);
    h = row[2] + x * h * (1.0 / 3.0);
    h = row[1] + x * h * (1.0 / 2.0);
    F[L] = row[0] + x * h;
    const double twoT = 2.0 * T;
#pragma unroll
    for (int m = L - 1; m >= 0; --m) F[m] = (twoT * F[m + 1] + emT) * (1.0 / (2 * m + 1));
  } else {
    const double emT = L > 0 ? exp(-T) : 0.0;
    const double r = rsqrt(T);
    const double inv2T = 0.5 * r * r;
    F[0] = kHalfSqrtPi * r;
#pragma unroll
    for (int m = 0; m < L; ++m) F[m + 1] = (static_cast<double>(2 * m + 1) * F[m] - emT) * inv2T;
  }
}

// Boys table staged in LDS by the fused kernels: for the class's L, row k holds
// e^{-k h}, F_L(k h) .. F_{L+7}(k h) and one pad (16-byte aligned rows).
constexpr int kBoysRows = 401;       // T = 0, h, ..., 40 with h = 0.1
constexpr int kBoysLdsStride = 10;

template <int L>
__device__ inline void boys_stage_lds(const gbx_tables& t, double* lds) {
  for (int i = threadIdx.x; i < kBoysRows * 9; i += kRegThreads) {
    const int k = i / 9, j = i - k * 9;
    lds[k * kBoysLdsStride + j] = t.boys_table[static_cast<int64_t>(k) * t.boys_nord +
                                               (j == 0 ? 0 : L + j)];
  }
}

// boys_eval_fixed with the table rows read from the LDS copy.
template <int L>
__device__ inline void boys_eval_lds(double T, const gbx_tables& t, const double* lds,
                                     double* F) {
  const double tmax = t.boys_h * (kBoysRows - 1);
  const double inv_h = 1.0 / t.boys_h;
  if (T <= tmax) {
    const int k = static_cast<int>(T * inv_h + 0.5);
    const double x = k * t.boys_h - T;
    const double* row = lds + k * kBoysLdsStride;
    const double emT = L > 0 ? row[0] * exp_small(x) : 0.0;
    double h = row[8];
    h = row[7] + x * h * (1.0 / 7.0);
    h = row[6] + x * h * (1.0 / 6.0);
    h = row[5] + x * h * (1.0 / 5.0);
    h = row[4] + x * h * (1.0 / 4.0);
    h = row[3] + x * h * (1.0 / 3.0);
    h = row[2] + x * h * (1.0 / 2.0);
    F[L] = row[1] + x * h;
    const double twoT = 2.0 * T;
#pragma unroll
    for (int m = L - 1; m >= 0; --m) F[m] = (twoT * F[m + 1] + emT) * (1.0 / (2 * m + 1));
  } else {
    const double emT = L > 0 ? exp(-T) : 0.0;
    const double r = rsqrt(T);
    const double inv2T = 0.5 * r * r;
    F[0] = kHalfSqrtPi * r;
#pragma unroll
    for (int m = 0; m < L; ++m) F[m + 1] = (static_cast<double>(2 * m + 1) * F[m] - emT) * inv2T;
  }
}

__device__ inline uint64_t tri_u64(uint64_t i, uint64_t j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// Bra and ket entries cover the same two shells (the quartet (ij|ij) of a shell pair, or
// one of its sub-quartets): the packed placement then writes only the ij >= kl half.
__device__ inline bool same_shells(int4 pb, int4 pk) {
  return (pb.x == pk.x && pb.y == pk.y) || (pb.x == pk.y && pb.y == pk.x);
}

// Store or (GBX_MODE_PACKED_ADD) add one packed value.
__device__ __forceinline__ void put_packed(double* p, double v, bool add) {
  if (add)
    *p += v;
  else
    *p = v;
}

// Write one canonical-order Cartesian block (NA x NB x NC x ND, factorial2 norms
// applied) through the per-shell maps (component position, norm_cont / sign scale)
// into the canonical packed slots (stored, or added in GBX_MODE_PACKED_ADD),
// skipping the duplicates of diagonal pairs and of bra and ket on the same shells.  In GBX_MODE_SCHWARZ write, for the bra pair, sqrt(max |block|)
// to out[2 bra] and sqrt(max |contracted [e0|f0]|) (accmax) to out[2 bra + 1].
template <int NA, int NB, int NC, int ND>
__device__ inline void store_block(const gbx_tables& t, const RegArgs& a, int bra, int ket,
                                   int4 pb, int4 pk, const double* out, double accmax) {
  constexpr int N = NA * NB * NC * ND;
  if (a.mode == GBX_MODE_SCHWARZ) {
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) m = fmax(m, fabs(out[i]));
    a.out[2 * bra] = sqrt(m);
    a.out[2 * bra + 1] = sqrt(accmax);
    return;
  }
  const int4* si = reinterpret_cast<const int4*>(t.shell_info);
  const int4 sa = si[2 * pb.x], sb = si[2 * pb.y], sc = si[2 * pk.x], sd = si[2 * pk.y];
  uint32_t fa[NA], fb[NB], fc[NC], fd[ND];
  double ca[NA], cb[NB], cc[NC], cd[ND];
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    fa[i] = static_cast<uint32_t>(sa.x + t.comp_map[sa.w + i]);
    ca[i] = t.comp_scale[sa.w + i];
  }
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    fb[i] = static_cast<uint32_t>(sb.x + t.comp_map[sb.w + i]);
    cb[i] = t.comp_scale[sb.w + i];
  }
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    fc[i] = static_cast<uint32_t>(sc.x + t.comp_map[sc.w + i]);
    cc[i] = t.comp_scale[sc.w + i];
  }
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    fd[i] = static_cast<uint32_t>(sd.x + t.comp_map[sd.w + i]);
    cd[i] = t.comp_scale[sd.w + i];
  }
  const bool bdiag = pb.x == pb.y, kdiag = pk.x == pk.y, same = same_shells(pb, pk);
  const bool add = a.mode == GBX_MODE_PACKED_ADD;
  double* base = a.out + t.pair_out_base[bra];
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
      if (bdiag && fa[ia] < fb[ib]) continue;
      const uint64_t ij = tri_u64(fa[ia], fb[ib]);
#pragma unroll
      for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
          if (kdiag && fc[ic] < fd[id]) continue;
          const uint64_t kl = tri_u64(fc[ic], fd[id]);
          if (same && ij < kl) continue;
          put_packed(base + tri_u64(ij, kl),
                     out[((ia * NB + ib) * NC + ic) * ND + id] * ca[ia] * cb[ib] * cc[ic] * cd[id],
                     add);
        }
    }
}

// Sum of the R partial accumulators f of quartet q written by the split kernels
// (same order as the cooperative horizontal-recursion pass: r = 0, 1, ...).
__device__ __forceinline__ double sum_parts(const RegArgs& a, int f, int64_t q) {
  const int R = a.lanes_per_quartet;
  const double* p = a.acc + static_cast<int64_t>(f) * R * a.acc_stride + q;
  double s = 0.0;
  for (int r = 0; r < R; ++r) s += p[static_cast<int64_t>(r) * a.acc_stride];
  return s;
}

// Ket part of the placement of one column (fixed c, d components).
struct ColumnSlot {
  uint64_t kl;
  double cc, cd;
  bool skip;
};

// Placement of one ket column (fixed c, d components) of a canonical-order block,
// element by element: the same maps, scales and duplicate rules as store_block.
template <int NA, int NB, int NC, int ND>
struct ColumnStore {
  const int32_t* cmap;   // table pointers held by value (a reference to the kernel
  const double* cscale;  // argument would force a private copy of it)
  double* base;
  int4 sc, sd;
  bool bdiag, kdiag, same, add;
  uint32_t fa[NA], fb[NB];
  double ca[NA], cb[NB];

  __device__ __forceinline__ ColumnStore(const gbx_tables& t, const RegArgs& a, int bra, int ket,
                                         int4 pb, int4 pk)
      : cmap(t.comp_map), cscale(t.comp_scale) {
    const int4* si = reinterpret_cast<const int4*>(t.shell_info);
    const int4 sa = si[2 * pb.x], sb = si[2 * pb.y];
    sc = si[2 * pk.x];
    sd = si[2 * pk.y];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      fa[i] = static_cast<uint32_t>(sa.x + cmap[sa.w + i]);
      ca[i] = cscale[sa.w + i];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      fb[i] = static_cast<uint32_t>(sb.x + cmap[sb.w + i]);
      cb[i] = cscale[sb.w + i];
    }
    bdiag = pb.x == pb.y;
    kdiag = pk.x == pk.y;
    same = same_shells(pb, pk);
    add = a.mode == GBX_MODE_PACKED_ADD;
    base = a.out + t.pair_out_base[bra];
  }

  __device__ __forceinline__ ColumnSlot column(int ic, int id) const {
    const uint32_t fc = static_cast<uint32_t>(sc.x + cmap[sc.w + ic]);
    const uint32_t fd = static_cast<uint32_t>(sd.x + cmap[sd.w + id]);
    return ColumnSlot{tri_u64(fc, fd), cscale[sc.w + ic], cscale[sd.w + id],
                      kdiag && fc < fd};
  }

  __device__ __forceinline__ void put(const ColumnSlot& cl, int ia, int ib, double v) const {
    if (cl.skip || (bdiag && fa[ia] < fb[ib])) return;
    const uint64_t ij = tri_u64(fa[ia], fb[ib]);
    if (same && ij < cl.kl) return;
    put_packed(base + tri_u64(ij, cl.kl), v * ca[ia] * cb[ib] * cl.cc * cl.cd, add);
  }
};

int launch_reg_generated(const gbx_tables* t, int la, int lb, int lc, int ld, const RegArgs& a,
                         int precision, hipStream_t s);
int launch_hrr_generated(const gbx_tables* t, int la, int lb, int lc, int ld, const RegArgs& a,
                         hipStream_t s);
bool has_hrr_generated(int la, int lb, int lc, int ld);
int launch_split_generated(const gbx_tables* t, int la, int lb, int lc, int ld, int k,
                           const RegArgs& a, int precision, hipStream_t s);
bool has_reg_generated(int la, int lb, int lc, int ld);
int split_count_generated(int la, int lb, int lc, int ld);

}  // namespace gbx
