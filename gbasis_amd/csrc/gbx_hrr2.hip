// Horizontal recursion, output transforms and placement of the large classes in LDS.
//
// The horizontal recursion of gbasis/integrals/_two_elec_int_improved.py:364-588 turns
// the contracted [e0|f0] (bra component e of degree l_a..l_a+l_b, ket component f of
// degree l_c..l_c+l_d) into (ab|cd).  Its ket part acts on each e independently and its
// bra part on each (c, d) independently, so one workgroup per shell quartet runs it in
// two stages whose working sets fit the LDS even where the whole recursion does not:
//   stage K, rows e in groups of rb: the ket recursion
//     [e0|c, d+1_i] = [e0|c+1_i, d] + CD_i [e0|c, d]
//   level by level (d of degree 1..l_d), the factorial2 norms of c and d and the
//   output transforms of shells C and D (base_four_symm.py:246-557), into the row W[e]
//   of an LDS array W[e][c'd'];
//   stage B, columns c'd' in groups of rb: the bra recursion
//     [a, b+1_i| = [a+1_i, b| + AB_i [a, b|
//   from W[.][c'd'], the norms and transforms of shells A and B, and the placement into
//   the canonical packed slots (base_four_symm.py:160-244), stored or added.
// The recursion steps come from a recipe built on the host per class (the direction of
// each step is the last nonzero one of the component raised, as in the class programs
// of gbasis_amd/programs.py, so the arithmetic is that of the cooperative pass).
//
// Recipe (int32): NA, NC, l_d, l_b, then the level sizes of the ket recursion (l_d
// values) and of the bra recursion (l_b values), then per level its entries
// s1 | s2 << 12 | dir << 24 (indices into the level below; level 0 of the ket is the
// accumulator order of f, that of the bra the order of e).  The last level of each side
// lists (c, d) (or (a, b)) in canonical component order, c major.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

// Workgroup sizes: 1024 lanes where the LDS arrays take most of the 160 KiB (one
// workgroup per CU; 16 wavefronts keep the level loops and the transforms busy), 512 and
// 256 where they are small enough for several workgroups, and so quartets, per CU.
constexpr int kHrr2Threads = 1024;
constexpr int kHrr2MidThreads = 512;
constexpr int kHrr2SmallThreads = 256;
constexpr int64_t kHrr2SmallLds = 32 * 1024;  // LDS bytes up to which the 256-lane form runs
constexpr int64_t kHrr2MidLds = 64 * 1024;    // and the 512-lane form

struct Hrr2Args {
  const int32_t* quart;
  int64_t nquart;
  const int32_t* count;
  const int32_t* rec;
  const double* acc_in;  // acc_in[(f * acc_parts + r) * acc_stride + qi] (qmajor 0) or
                         // acc_in[(qi * acc_parts + r) * acc_stride + f] (qmajor 1),
                         // summed over r
  int64_t acc_stride;
  int32_t acc_parts;
  int32_t qmajor;
  double* out;
  int32_t add;           // 1: add into the packed slots (later sub-quartet phases)
  int32_t la, lb, lc, ld;
  int32_t rb;            // rows (stage K) / columns (stage B) per group
  int32_t half;          // doubles per half of a row buffer
  int32_t w_off;         // doubles before W in the LDS
  int32_t max_dc, max_dd;  // output dimensions W is sized for
};

__device__ inline uint64_t hrr2_tri(uint64_t i, uint64_t j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// The flat index (r * A + a) * B + b of a [nr][A][B] loop, visited by a thread from
// threadIdx.x in steps of kT (the workgroup size): the steps are decomposed once, so the
// loops need no integer division per element (B = 1 or A = 1 for two-level loops).
template <int kT>
struct Hrr2Iter {
  int r, a, b;
  int dr, da, db;
  int A, B;
  __device__ Hrr2Iter(int A_, int B_) : A(A_), B(B_) {
    const int q = threadIdx.x / B;
    b = threadIdx.x - q * B;
    r = q / A;
    a = q - r * A;
    const int qs = kT / B;
    db = kT - qs * B;
    dr = qs / A;
    da = qs - dr * A;
  }
  __device__ void next() {
    b += db;
    int carry = 0;
    if (b >= B) {
      b -= B;
      carry = 1;
    }
    a += da + carry;
    carry = 0;
    if (a >= A) {
      a -= A;
      carry = 1;
    }
    r += dr + carry;
  }
};

// One level of a recursion for rows [0, nr) of the row buffers:
// dst[j] = src[s1] + g_dir src[s2], g = (gx, gy, gz).
template <int kT>
__device__ inline void hrr2_level(double* buf, int srow, int src_half, int dst_half, int half,
                                  int nr, int size, const int32_t* ent, double gx, double gy,
                                  double gz) {
  for (Hrr2Iter<kT> it(size, 1); it.r < nr; it.next()) {
    const int r = it.r, j = it.a;
    const int e = ent[j];
    const int dir = e >> 24;
    const double g = dir == 0 ? gx : (dir == 1 ? gy : gz);
    const double* s = buf + r * srow + src_half * half;
    buf[r * srow + dst_half * half + j] = s[e & 0xFFF] + g * s[(e >> 12) & 0xFFF];
  }
}

template <int kT>
__global__ __launch_bounds__(kT) void eri_hrr2_kernel(gbx_tables t, Hrr2Args a) {
  extern __shared__ __align__(16) double lds[];
  double* buf = lds;
  double* W = lds + a.w_off;
  const int srow = 2 * a.half;
  const int NA = a.rec[0], NC = a.rec[1];
  const int nlk = a.rec[2], nlb = a.rec[3];
  // first entry of the bra levels
  int off_b0 = 4 + nlk + nlb;
  for (int k = 1; k <= nlk; ++k) off_b0 += a.rec[3 + k];
  const int nc[4] = {(a.la + 1) * (a.la + 2) / 2, (a.lb + 1) * (a.lb + 2) / 2,
                     (a.lc + 1) * (a.lc + 2) / 2, (a.ld + 1) * (a.ld + 2) / 2};
  const double* f2[4] = {t.f2norm + f2_offset(a.la), t.f2norm + f2_offset(a.lb),
                         t.f2norm + f2_offset(a.lc), t.f2norm + f2_offset(a.ld)};
  const int4* pinfo = reinterpret_cast<const int4*>(t.pair_info);
  const int4* sinfo4 = reinterpret_cast<const int4*>(t.shell_info);
  const int64_t nquart = a.count ? min(a.nquart, static_cast<int64_t>(*a.count)) : a.nquart;

  for (int64_t qi = xcd_chunked(blockIdx.x, gridDim.x); qi < nquart; qi += gridDim.x) {
    const int bra = a.quart[2 * qi], ket = a.quart[2 * qi + 1];
    const int4 pb = pinfo[bra];
    const int4 pk = pinfo[ket];
    const int sh[4] = {pb.x, pb.y, pk.x, pk.y};
    int4 si0[4], si1[4];
    int dim[4];
    for (int x = 0; x < 4; ++x) {
      si0[x] = sinfo4[2 * sh[x]];
      si1[x] = sinfo4[2 * sh[x] + 1];
      dim[x] = si0[x].z == 1 ? si1[x].y : nc[x];
    }
    // W holds NA rows of max_dc x max_dd values (the largest output dimensions of shells
    // of l_c and l_d in the batch, from the host); a quartet beyond them is never queued
    if (dim[2] > a.max_dc || dim[3] > a.max_dd) continue;
    const double CDx = t.pair_geom[4 * ket], CDy = t.pair_geom[4 * ket + 1],
                 CDz = t.pair_geom[4 * ket + 2];
    const double ABx = t.pair_geom[4 * bra], ABy = t.pair_geom[4 * bra + 1],
                 ABz = t.pair_geom[4 * bra + 2];
    const int NJ = dim[2] * dim[3];

    // ---- stage K: rows e of W ----
    for (int e0 = 0; e0 < NA; e0 += a.rb) {
      const int nr = min(a.rb, NA - e0);
      __syncthreads();  // the buffers of the previous group (or quartet) are free
      for (Hrr2Iter<kT> it(NC, 1); it.r < nr; it.next()) {
        const int r = it.r, c = it.a;
        const int64_t f = static_cast<int64_t>(c) * NA + e0 + r;
        double s = 0.0;
        if (a.qmajor) {
#pragma unroll 4
          for (int p = 0; p < a.acc_parts; ++p)
            s += a.acc_in[(qi * a.acc_parts + p) * a.acc_stride + f];
        } else {
#pragma unroll 4
          for (int p = 0; p < a.acc_parts; ++p)
            s += a.acc_in[(f * a.acc_parts + p) * a.acc_stride + qi];
        }
        buf[r * srow + c] = s;
      }
      __syncthreads();
      for (int k = 1, o = 4 + nlk + nlb; k <= nlk; ++k) {
        const int size = a.rec[3 + k];
        hrr2_level<kT>(buf, srow, (k - 1) & 1, k & 1, a.half, nr, size, a.rec + o, CDx, CDy,
                       CDz);
        o += size;
        __syncthreads();
      }
      // the Cartesian (c, d) block of each row in half h; norms of c and d, transform
      // of C into half 1 - h as [c'][d], then that of D into W
      const int h = nlk & 1;
      {
        const double* M = t.out_mat + si1[2].x;
        for (Hrr2Iter<kT> it(dim[2], nc[3]); it.r < nr; it.next()) {
          const int r = it.r, jc = it.a, id = it.b;
          const int rem = jc * nc[3] + id;
          const double* F = buf + r * srow + h * a.half;
          double v;
          if (si0[2].z == 1) {
            v = 0.0;
            for (int ic = 0; ic < nc[2]; ++ic)
              v += M[jc * nc[2] + ic] * f2[2][ic] * F[ic * nc[3] + id];
          } else {
            v = f2[2][jc] * F[jc * nc[3] + id];
          }
          buf[r * srow + (1 - h) * a.half + rem] = v * f2[3][id];
        }
      }
      __syncthreads();
      {
        const double* M = t.out_mat + si1[3].x;
        for (Hrr2Iter<kT> it(dim[2], dim[3]); it.r < nr; it.next()) {
          const int r = it.r, jc = it.a, jd = it.b;
          const int j = jc * dim[3] + jd;
          const double* T = buf + r * srow + (1 - h) * a.half + jc * nc[3];
          double v;
          if (si0[3].z == 1) {
            v = 0.0;
            for (int id = 0; id < nc[3]; ++id) v += M[jd * nc[3] + id] * T[id];
          } else {
            v = T[jd];
          }
          W[static_cast<int64_t>(e0 + r) * NJ + j] = v;
        }
      }
    }

    // ---- stage B: columns c'd' of W, bra recursion, transforms, placement ----
    const uint64_t obase = t.pair_out_base[bra];
    const bool bra_diag = pb.x == pb.y, ket_diag = pk.x == pk.y;
    const bool same_pair = (pb.x == pk.x && pb.y == pk.y) || (pb.x == pk.y && pb.y == pk.x);
    for (int j0 = 0; j0 < NJ; j0 += a.rb) {
      const int nr = min(a.rb, NJ - j0);
      __syncthreads();  // W is complete; the buffers of the previous group are free
      for (Hrr2Iter<kT> it(nr, 1); it.r < NA; it.next()) {
        const int e = it.r, r = it.a;
        buf[r * srow + e] = W[static_cast<int64_t>(e) * NJ + j0 + r];
      }
      __syncthreads();
      for (int k = 1, o = off_b0; k <= nlb; ++k) {
        const int size = a.rec[3 + nlk + k];
        hrr2_level<kT>(buf, srow, (k - 1) & 1, k & 1, a.half, nr, size, a.rec + o, ABx, ABy,
                       ABz);
        o += size;
        __syncthreads();
      }
      const int h = nlb & 1;
      {
        const double* M = t.out_mat + si1[0].x;
        for (Hrr2Iter<kT> it(dim[0], nc[1]); it.r < nr; it.next()) {
          const int r = it.r, ja = it.a, ib = it.b;
          const int rem = ja * nc[1] + ib;
          const double* F = buf + r * srow + h * a.half;
          double v;
          if (si0[0].z == 1) {
            v = 0.0;
            for (int ia = 0; ia < nc[0]; ++ia)
              v += M[ja * nc[0] + ia] * f2[0][ia] * F[ia * nc[1] + ib];
          } else {
            v = f2[0][ja] * F[ja * nc[1] + ib];
          }
          buf[r * srow + (1 - h) * a.half + rem] = v * f2[1][ib];
        }
      }
      __syncthreads();
      {
        const double* M = t.out_mat + si1[1].x;
        int rr = -1, jc = 0, jd = 0;
        for (Hrr2Iter<kT> it(dim[0], dim[1]); it.r < nr; it.next()) {
          const int r = it.r, ja = it.a, jb = it.b;
          if (r != rr) {  // the column (c', d') of this row group
            rr = r;
            jc = (j0 + r) / dim[3];
            jd = j0 + r - jc * dim[3];
          }
          const double* T = buf + r * srow + (1 - h) * a.half + ja * nc[1];
          double v;
          if (si0[1].z == 1) {
            v = 0.0;
            for (int ib = 0; ib < nc[1]; ++ib) v += M[jb * nc[1] + ib] * T[ib];
          } else {
            v = T[jb];
          }
          const int pos[4] = {ja, jb, jc, jd};
          uint32_t fn[4];
          for (int x = 0; x < 4; ++x) {
            int p = pos[x];
            if (si0[x].z == 0) {
              v *= t.comp_scale[si0[x].w + p];
              p = t.comp_map[si0[x].w + p];
            }
            fn[x] = static_cast<uint32_t>(si0[x].x + p);
          }
          if (bra_diag && fn[0] < fn[1]) continue;
          if (ket_diag && fn[2] < fn[3]) continue;
          const uint64_t ij = hrr2_tri(fn[0], fn[1]), kl = hrr2_tri(fn[2], fn[3]);
          if (same_pair && ij < kl) continue;
          if (a.add)
            a.out[obase + hrr2_tri(ij, kl)] += v;
          else
            a.out[obase + hrr2_tri(ij, kl)] = v;
        }
      }
    }
  }
}

inline int64_t hrr2_lds_limit() {
  static int64_t cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  if (cached[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
      v = 65536;
    cached[dev] = v;
  }
  return cached[dev];
}

}  // namespace gbx

using namespace gbx;

extern "C" int gbx_eri_hrr_lds(const gbx_tables* t, int32_t la, int32_t lb, int32_t lc,
                               int32_t ld, const int32_t* h_recipe, const int32_t* d_recipe,
                               int32_t recipe_len, int32_t max_dim_c, int32_t max_dim_d,
                               const int32_t* d_quartets, int64_t nquart, const int32_t* d_count,
                               int32_t mode, const double* d_acc, int64_t acc_stride,
                               int32_t acc_parts, int32_t acc_layout, double* d_out,
                               void* stream) {
  if (!t || !h_recipe || !d_recipe || !d_out || (nquart > 0 && (!d_quartets || !d_acc)))
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: null argument");
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > kMaxShellL || lb > kMaxShellL ||
      lc > kMaxShellL || ld > kMaxShellL)
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: angular momentum out of range");
  if (mode != GBX_MODE_PACKED && mode != GBX_MODE_PACKED_ADD)
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: packed modes only");
  if (acc_parts < 1 || (acc_layout != 0 && acc_layout != 1))
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: acc_parts < 1 or bad acc_layout");
  auto ncomp = [](int lo, int hi) {
    int n = 0;
    for (int l = lo; l <= hi; ++l) n += (l + 1) * (l + 2) / 2;
    return n;
  };
  auto nct = [](int l) { return (l + 1) * (l + 2) / 2; };
  const int NA = ncomp(la, la + lb), NC = ncomp(lc, lc + ld);
  // the recipe must describe this class; its level sizes bound the row buffers
  if (recipe_len < 4 || h_recipe[0] != NA || h_recipe[1] != NC || h_recipe[2] != ld ||
      h_recipe[3] != lb || recipe_len < 4 + ld + lb)
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: recipe does not match the class");
  int64_t entries = 0, half = std::max(NA, NC);
  for (int k = 0; k < ld + lb; ++k) {
    const int s = h_recipe[4 + k];
    if (s < 1 || s > 4095) return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: bad recipe level size");
    entries += s;
    half = std::max<int64_t>(half, s);
  }
  if (recipe_len != 4 + ld + lb + entries)
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: recipe length does not match its levels");
  if (ld > 0 && h_recipe[4 + ld - 1] != nct(lc) * nct(ld))
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: last ket level is not the (c, d) block");
  if (lb > 0 && h_recipe[4 + ld + lb - 1] != nct(la) * nct(lb))
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: last bra level is not the (a, b) block");
  if (acc_layout == 0 ? acc_stride < nquart : acc_stride < int64_t(NA) * NC)
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: acc_stride below the quartets (layout 0) or "
                             "the accumulators (layout 1)");
  if (max_dim_c < 1 || max_dim_c > nct(lc) || max_dim_d < 1 || max_dim_d > nct(ld))
    return fail(GBX_ERR_ARG, "gbx_eri_hrr_lds: output dimensions out of range");
  // the transform intermediates share the row buffers
  half = std::max<int64_t>({half, int64_t(nct(lc)) * nct(ld), int64_t(nct(la)) * nct(lb)});
  half = (half + 1) & ~int64_t(1);
  const int64_t w_doubles = int64_t(NA) * max_dim_c * max_dim_d;
  const int64_t limit = hrr2_lds_limit();
  const int64_t room = limit / 8 - w_doubles;
  const int64_t rb = std::min<int64_t>(16, room / (2 * half));
  if (rb < 1)
    return fail(GBX_ERR_UNSUPPORTED, "gbx_eri_hrr_lds: the class does not fit the LDS");
  if (nquart == 0) return GBX_OK;
  Hrr2Args a{};
  a.quart = d_quartets;
  a.nquart = nquart;
  a.count = d_count;
  a.rec = d_recipe;
  a.acc_in = d_acc;
  a.acc_stride = acc_stride;
  a.acc_parts = acc_parts;
  a.qmajor = acc_layout;
  a.out = d_out;
  a.add = mode == GBX_MODE_PACKED_ADD ? 1 : 0;
  a.la = la;
  a.lb = lb;
  a.lc = lc;
  a.ld = ld;
  a.rb = static_cast<int>(rb);
  a.half = static_cast<int>(half);
  a.w_off = static_cast<int>(rb * 2 * half);
  a.max_dc = max_dim_c;
  a.max_dd = max_dim_d;
  const size_t shmem = static_cast<size_t>((rb * 2 * half + w_doubles) * 8);
  const unsigned grid = static_cast<unsigned>(std::min<int64_t>(nquart, int64_t(1) << 20));
  if (static_cast<int64_t>(shmem) <= kHrr2SmallLds) {
    hipLaunchKernelGGL(eri_hrr2_kernel<kHrr2SmallThreads>, dim3(grid), dim3(kHrr2SmallThreads),
                       shmem, static_cast<hipStream_t>(stream), *t, a);
    return check_launch("eri_hrr2_kernel");
  }
  if (static_cast<int64_t>(shmem) <= kHrr2MidLds) {  // within the default 64 KiB limit
    hipLaunchKernelGGL(eri_hrr2_kernel<kHrr2MidThreads>, dim3(grid), dim3(kHrr2MidThreads),
                       shmem, static_cast<hipStream_t>(stream), *t, a);
    return check_launch("eri_hrr2_kernel");
  }
  if (shmem > 65536) {
    static bool raised[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(GBX_ERR_HIP, "eri_hrr2_kernel: no device");
    if (dev < 0 || dev >= 64 || !raised[dev]) {
      if (hipFuncSetAttribute(reinterpret_cast<const void*>(&eri_hrr2_kernel<kHrr2Threads>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(limit)) != hipSuccess)
        return fail(GBX_ERR_HIP, "eri_hrr2_kernel: could not raise the LDS limit");
      if (dev >= 0 && dev < 64) raised[dev] = true;
    }
  }
  hipLaunchKernelGGL(eri_hrr2_kernel<kHrr2Threads>, dim3(grid), dim3(kHrr2Threads), shmem,
                     static_cast<hipStream_t>(stream), *t, a);
  return check_launch("eri_hrr2_kernel");
}
