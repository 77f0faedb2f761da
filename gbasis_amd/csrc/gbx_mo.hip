// Four-index transform of the canonical 8-fold packed ERIs of one molecule into a
// new basis (e.g. molecular orbitals), without a dense N^4 intermediate:
//
//   (ab|cd) = sum_ijkl T_ai T_bj T_ck T_dl (ij|kl)
//
// (gbasis/base_four_symm.py:559-616, four tensordot passes over the dense array in
// the reference).  Both halves are the same operation on a batch of symmetric
// matrices, Y = T X T^T with X symmetric n x n and T m x n:
//   pass 1, one matrix per AO pair r = ij (i >= j): X_r[k][l] = (ij|kl) from the packed
//     array; the lower triangle Y_r[c][d] (c >= d) goes to the half-transformed array
//     H[cd][ij] (Mpair x Npair, Npair = n(n+1)/2, Mpair = m(m+1)/2);
//   pass 2, one matrix per new pair r = cd: X_r[i][j] = H[cd][tri(i, j)]; the lower
//     triangle of Y_r gives (ab|cd) for a >= b, written to the four dense positions of
//     (ab|cd) = (ba|cd) = (ab|dc) = (ba|dc) in chemist or physicist order.
// Each pass runs, over chunks of rc matrices, as two GEMMs on the FP64 matrix cores:
//   GEMM 1, rows (k, q):  U[k][d][q] = sum_l X_q[k][l] T[d][l]   (rows gathered from
//           the packed array or H);
//   GEMM 2, rows (d, q):  Z[d][q][c] = sum_k U[k][d][q] T[c][k] = Y_q[c][d],
//           whose epilogue keeps c >= d and places it in H or the dense output.
// The matrix q of the chunk is the fastest index of both GEMMs' rows, so that the
// U traffic, the H rows and the dense output (whose last index is the pair's second
// function) are written and read along consecutive addresses.
// The GEMM tile: 64 x 64 outputs per 256-lane workgroup, K in steps of 16 staged in
// LDS, each wavefront a 32 x 32 block of four 16 x 16 v_mfma_f64_16x16x4f64
// accumulators.  Operand and result lanes of that instruction: A[i][k] in lane
// i + 16 k, B[k][j] in lane j + 16 k, C[i][j] in lane j + 16 (i mod 4), register i / 4.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

constexpr int kMoThreads = 256;
constexpr int kMoBM = 64, kMoBK = 16;
constexpr int kMoPadA = kMoBK + 1;   // LDS row stride of the A tile (doubles), odd against
                                     // bank conflicts of the fragment reads
constexpr int kMoGroup = 8;          // GEMM 1 rows: groups of 8 matrices per row k
constexpr int64_t kMoChunkDoubles = int64_t(1) << 27;  // U of one chunk: at most 1 GiB
constexpr int kMoFill = 32;          // tile edge of the ket-transpose fill

typedef double mo_v4d __attribute__((ext_vector_type(4)));

struct MoArgs {
  const double* packed;  // pass 1 source
  double* half;          // H: pass 1 output, pass 2 source
  double* out;           // dense m^4 output
  const double* TT;      // n x m (T transposed), row-major
  double* u;             // U of the current chunk: [k][d][q]
  int64_t n, m, npair_n;
  int64_t r0, rc;        // matrices [r0, r0 + rc) of this chunk
  int64_t tiles;         // row tiles x column tiles of the GEMM
  int notation;          // 0 chemist, 1 physicist
};

__device__ inline uint64_t mo_tri(uint64_t i, uint64_t j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// t = r (r + 1) / 2 + c with r >= c
__device__ inline void mo_tri_split(uint64_t t, uint64_t& r, uint64_t& c) {
  uint64_t x = static_cast<uint64_t>((sqrt(8.0 * static_cast<double>(t) + 1.0) - 1.0) * 0.5);
  while (x * (x + 1) / 2 > t) --x;
  while ((x + 1) * (x + 2) / 2 <= t) ++x;
  r = x;
  c = t - x * (x + 1) / 2;
}

// GEMM 1 row -> (row k of X, matrix q of the chunk): rows run over groups of kMoGroup
// matrices, all n rows k of a group, the group's matrices fastest.  A row tile thus
// covers 8 matrices x 8 rows, whose packed-array (or H) lines it shares with the next
// row tiles; q >= rc marks padding rows.
__device__ inline void mo_row1(int64_t row, int64_t n, int64_t& k, int64_t& q) {
  const int64_t t = row / kMoGroup;
  const int64_t g = t / n;
  k = t - g * n;
  q = g * kMoGroup + (row - t * kMoGroup);
}

// One GEMM of a pass (kStage 1 or 2); K = n and N = m for both, M = n ceil8(rc) (stage 1)
// or m rc (stage 2).  The workgroup tile is 64 rows x 32 kNT columns: 4 wavefronts,
// 2 x 2, each 32 rows x 16 kNT columns of 16 x 16 MFMA tiles; K steps of 16 through LDS
// with the next step's global loads held in registers during the current step's MFMAs.
// The grid is 1-D, padded to a multiple of the XCD count; the logical tile of a
// workgroup makes the consecutive workgroups of one XCD neighbours, so that they share
// A lines in that XCD's L2.
template <int kPass, int kStage, int kNT>
__global__ __launch_bounds__(kMoThreads, 2) void mo_gemm_kernel(MoArgs a) {
  constexpr int kBN = 32 * kNT;
  constexpr int kPadB = kBN + 1;
  constexpr int kBPer = kMoBK * kBN / kMoThreads;  // B elements per thread and K step
  __shared__ double As[kMoBM * kMoPadA];
  __shared__ double Bs[kMoBK * kPadB];
  const int64_t n = a.n, m = a.m, rc = a.rc;
  const int64_t M = kStage == 1 ? n * ((rc + kMoGroup - 1) / kMoGroup * kMoGroup) : m * rc;
  const int64_t per_xcd = (a.tiles + kXcdCount - 1) / kXcdCount;
  const int64_t tile = (blockIdx.x % kXcdCount) * per_xcd + blockIdx.x / kXcdCount;
  if (tile >= a.tiles) return;  // the whole workgroup leaves before any barrier
  const int64_t col_tiles = (m + kBN - 1) / kBN;
  const int64_t row0 = (tile / col_tiles) * kMoBM, col0 = (tile % col_tiles) * kBN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 16 * kNT;
  // GEMM 2 keeps c >= d only: a 16-column MFMA tile whose columns all lie below the
  // smallest d of the wavefront's rows is skipped (a wave-uniform test)
  bool live[kNT];
  {
    const int64_t d_lo = kStage == 2 ? (row0 + wm) / rc : 0;
#pragma unroll
    for (int y = 0; y < kNT; ++y) live[y] = col0 + wn + 16 * y + 15 >= d_lo;
  }

  // A loader.  Stage 1: 16 lanes along l in each of the rows tid / 16 + 16 u, row
  // (k, q) = X_{r0+q}[k][:].  Stage 2: the 64 lanes along the rows (d, q), which are
  // consecutive in U, K columns 4 (tid / 64) .. + 3.
  bool a_ok[4];
  uint64_t a_r[4], a_k[4];  // stage 1: matrix and row index per loaded row
  if constexpr (kStage == 1) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int64_t k = 0, q = 0;
      mo_row1(row0 + (tid >> 4) + 16 * u, n, k, q);
      a_ok[u] = q < rc;
      a_k[u] = static_cast<uint64_t>(k);
      a_r[u] = static_cast<uint64_t>(a.r0 + (a_ok[u] ? q : 0));
    }
  } else {
    a_ok[0] = row0 + lane < M;  // U[k][d][q] = u[(k m + d) rc + q] = u[k m rc + row]
  }
  auto load_a = [&](int64_t k0, double* v) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      v[u] = 0.0;
      if constexpr (kStage == 1) {
        const int64_t l = k0 + (tid & 15);
        if (a_ok[u] && l < n) {
          const uint64_t kl = mo_tri(a_k[u], static_cast<uint64_t>(l));
          if constexpr (kPass == 1)
            v[u] = a.packed[mo_tri(a_r[u], kl)];
          else
            v[u] = a.half[a_r[u] * static_cast<uint64_t>(a.npair_n) + kl];
        }
      } else {
        const int64_t k = k0 + (tid >> 6) * 4 + u;
        if (a_ok[0] && k < n) v[u] = a.u[k * m * rc + row0 + lane];
      }
    }
  };
  auto store_a = [&](const double* v) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if constexpr (kStage == 1)
        As[((tid >> 4) + 16 * u) * kMoPadA + (tid & 15)] = v[u];
      else
        As[lane * kMoPadA + (tid >> 6) * 4 + u] = v[u];
    }
  };
  // B loader: element e = tid + 256 j of the 16 x kBN tile, row e / kBN, column e % kBN
  auto load_b = [&](int64_t k0, double* v) {
#pragma unroll
    for (int j = 0; j < kBPer; ++j) {
      const int e = tid + kMoThreads * j;
      const int64_t k = k0 + e / kBN, c = col0 + e % kBN;
      v[j] = (k < n && c < m) ? a.TT[k * m + c] : 0.0;
    }
  };
  auto store_b = [&](const double* v) {
#pragma unroll
    for (int j = 0; j < kBPer; ++j) {
      const int e = tid + kMoThreads * j;
      Bs[(e / kBN) * kPadB + e % kBN] = v[j];
    }
  };

  mo_v4d acc[2][kNT];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < kNT; ++y) acc[x][y] = mo_v4d{0.0, 0.0, 0.0, 0.0};

  double va[4], vb[kBPer];
  load_a(0, va);
  load_b(0, vb);
  for (int64_t k0 = 0; k0 < n; k0 += kMoBK) {
    __syncthreads();  // the previous K step's tiles are consumed
    store_a(va);
    store_b(vb);
    __syncthreads();
    if (k0 + kMoBK < n) {  // the next step's loads fly during this step's MFMAs
      load_a(k0 + kMoBK, va);
      load_b(k0 + kMoBK, vb);
    }
#pragma unroll
    for (int ks = 0; ks < kMoBK; ks += 4) {
      double af[2], bf[kNT];
#pragma unroll
      for (int x = 0; x < 2; ++x)
        af[x] = As[(wm + x * 16 + (lane & 15)) * kMoPadA + ks + (lane >> 4)];
#pragma unroll
      for (int y = 0; y < kNT; ++y)
        bf[y] = Bs[(ks + (lane >> 4)) * kPadB + wn + y * 16 + (lane & 15)];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < kNT; ++y)
          if (live[y])
            acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[x], bf[y], acc[x][y], 0, 0, 0);
    }
  }

  // epilogue: element (i, j) of a 16 x 16 tile sits in lane j + 16 (i mod 4), register i / 4
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int reg = 0; reg < 4; ++reg) {
      const int64_t row = row0 + wm + x * 16 + (lane >> 4) + 4 * reg;
      if (row >= M) continue;
      int64_t i = 0, q = 0;  // (k, q) in stage 1, (d, q) in stage 2
      if constexpr (kStage == 1) {
        mo_row1(row, n, i, q);
        if (q >= rc) continue;
      } else {
        i = row / rc;
        q = row - i * rc;
      }
#pragma unroll
      for (int y = 0; y < kNT; ++y) {
        const int64_t col = col0 + wn + y * 16 + (lane & 15);  // d (stage 1) or c (stage 2)
        if (col >= m) continue;
        const double v = acc[x][y][reg];
        if constexpr (kStage == 1) {
          a.u[(i * m + col) * rc + q] = v;
        } else {
          const int64_t c = col, d = i;
          if (c < d) continue;  // Y_q is symmetric: the lower triangle only
          const uint64_t r = static_cast<uint64_t>(a.r0 + q);
          if constexpr (kPass == 1) {
            a.half[mo_tri(c, d) * static_cast<uint64_t>(a.npair_n) + r] = v;
          } else {
            // (ab|ef) with (a, b) = (c, d) of Y_q and (e, f) the new pair r, e >= f; the
            // (f, e) positions are filled by mo_fill_kernel
            uint64_t e = 0, f = 0;
            mo_tri_split(r, e, f);
            const uint64_t mm = m, A = c, B = d;
            if (a.notation == 0) {  // chemist (ab|ef) at [a][b][e][f]
              a.out[((A * mm + B) * mm + e) * mm + f] = v;
              a.out[((B * mm + A) * mm + e) * mm + f] = v;
            } else {                // physicist <ae|bf> = (ab|ef) at [a][e][b][f]
              a.out[((A * mm + e) * mm + B) * mm + f] = v;
              a.out[((B * mm + e) * mm + A) * mm + f] = v;
            }
          }
        }
      }
    }
}

// (ab|fe) = (ab|ef): copies each position of the new pair (e, f), e > f, to its (f, e)
// position.  The pair's two indices are the last two (chemist [a][b][e][f], row stride
// m) or the second and last (physicist [a][e][b][f], row stride m^2) of the dense
// array; per slice (a, b), a 32 x 32 tile below the diagonal is read along f, turned in
// LDS and written along e.  Work items (slice, tile pair) are strided over the grid.
__global__ __launch_bounds__(kMoThreads) void mo_fill_kernel(double* out, int64_t m,
                                                             int notation, int64_t items) {
  __shared__ double tile[kMoFill][kMoFill + 1];
  const int64_t nt = (m + kMoFill - 1) / kMoFill;
  const int64_t npt = nt * (nt + 1) / 2;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 lanes
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int64_t slice = it / npt;
    const int64_t tp = it - slice * npt;
    int64_t te = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(tp) + 1.0) - 1.0) * 0.5);
    while (te * (te + 1) / 2 > tp) --te;
    while ((te + 1) * (te + 2) / 2 <= tp) ++te;
    const int64_t tf = tp - te * (te + 1) / 2;  // tile pair (te, tf), te >= tf
    const int64_t sa = slice / m, sb = slice - sa * m;
    const int64_t base = notation == 0 ? slice * m * m : sa * m * m * m + sb * m;
    const int64_t rs = notation == 0 ? m : m * m;  // stride of e
    const int64_t e0 = te * kMoFill, f0 = tf * kMoFill;
    __syncthreads();  // the previous item's tile is consumed
    for (int yy = ty; yy < kMoFill; yy += 8) {
      const int64_t e = e0 + yy, f = f0 + tx;
      if (e < m && f < m && e > f) tile[yy][tx] = out[base + e * rs + f];
    }
    __syncthreads();
    for (int yy = ty; yy < kMoFill; yy += 8) {
      const int64_t f = f0 + yy, e = e0 + tx;  // write [f][e] along e
      if (e < m && f < m && e > f) out[base + f * rs + e] = tile[tx][yy];
    }
  }
}

// Matrices per chunk: their U takes at most kMoChunkDoubles.
inline int64_t mo_chunk(int64_t n, int64_t m, int64_t nmat) {
  return std::max<int64_t>(1, std::min<int64_t>(nmat, kMoChunkDoubles / (n * m)));
}

// 16-column MFMA tiles per wavefront: one column tile of 32 kNT >= m when m <= 192, else
// the width in 4..6 that pads m least.
inline int mo_nt(int64_t m) {
  if (m <= 192) return static_cast<int>((m + 31) / 32);
  int best = 4;
  int64_t pad = (m + 127) / 128 * 128;
  for (int nt = 5; nt <= 6; ++nt) {
    const int64_t p = (m + 32 * nt - 1) / (32 * nt) * (32 * nt);
    if (p <= pad) { pad = p; best = nt; }
  }
  return best;
}

template <int kPass, int kStage, int kNT>
int mo_gemm_nt(MoArgs a, hipStream_t s) {
  const int64_t rows = kStage == 1 ? a.n * ((a.rc + kMoGroup - 1) / kMoGroup * kMoGroup)
                                   : a.m * a.rc;
  a.tiles = ((rows + kMoBM - 1) / kMoBM) * ((a.m + 32 * kNT - 1) / (32 * kNT));
  const int64_t blocks = (a.tiles + kXcdCount - 1) / kXcdCount * kXcdCount;
  hipLaunchKernelGGL((mo_gemm_kernel<kPass, kStage, kNT>), dim3(static_cast<unsigned>(blocks)),
                     dim3(kMoThreads), 0, s, a);
  return check_launch(kStage == 1 ? "mo_gemm_kernel (stage 1)" : "mo_gemm_kernel (stage 2)");
}

template <int kPass, int kStage>
int mo_gemm(MoArgs a, hipStream_t s) {
  switch (mo_nt(a.m)) {
    case 1: return mo_gemm_nt<kPass, kStage, 1>(a, s);
    case 2: return mo_gemm_nt<kPass, kStage, 2>(a, s);
    case 3: return mo_gemm_nt<kPass, kStage, 3>(a, s);
    case 4: return mo_gemm_nt<kPass, kStage, 4>(a, s);
    case 5: return mo_gemm_nt<kPass, kStage, 5>(a, s);
    default: return mo_gemm_nt<kPass, kStage, 6>(a, s);
  }
}

template <int kPass>
int mo_pass(MoArgs a, int64_t nmat, hipStream_t s) {
  const int64_t rc_max = mo_chunk(a.n, a.m, nmat);
  for (int64_t r0 = 0; r0 < nmat; r0 += rc_max) {
    a.r0 = r0;
    a.rc = std::min(rc_max, nmat - r0);
    if (int rc = mo_gemm<kPass, 1>(a, s)) return rc;
    if (int rc = mo_gemm<kPass, 2>(a, s)) return rc;
  }
  return GBX_OK;
}

inline int mo_fill(double* out, int64_t m, int notation, hipStream_t s) {
  if (m < 2) return GBX_OK;
  const int64_t nt = (m + kMoFill - 1) / kMoFill;
  const int64_t items = m * m * (nt * (nt + 1) / 2);
  const int64_t blocks = std::min<int64_t>(items, int64_t(1) << 16);
  hipLaunchKernelGGL(mo_fill_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kMoThreads), 0,
                     s, out, m, notation, items);
  return check_launch("mo_fill_kernel");
}

// Doubles of the workspace: H, then the U of one chunk (the larger of the two passes').
inline int64_t mo_workspace_doubles(int64_t n, int64_t m, int64_t* half_doubles) {
  const int64_t npn = n * (n + 1) / 2, npm = m * (m + 1) / 2;
  const int64_t u1 = mo_chunk(n, m, npn) * n * m, u2 = mo_chunk(n, m, npm) * n * m;
  if (half_doubles) *half_doubles = npn * npm;
  return npn * npm + std::max(u1, u2);
}

}  // namespace gbx

using namespace gbx;

extern "C" int gbx_mo_workspace(int64_t nfunc, int64_t nnew, int64_t* workspace_elems) {
  if (!workspace_elems) return fail(GBX_ERR_ARG, "gbx_mo_workspace: null argument");
  if (nfunc < 1 || nnew < 1 || nfunc > 32768 || nnew > 32768)
    return fail(GBX_ERR_ARG, "gbx_mo_workspace: sizes must be in [1, 32768]");
  *workspace_elems = mo_workspace_doubles(nfunc, nnew, nullptr);
  return GBX_OK;
}

extern "C" int gbx_mo_transform(const double* d_packed, int64_t nfunc, const double* d_T,
                                const double* d_TT, int64_t nnew, double* d_work,
                                double* d_out, int32_t notation, void* stream) {
  if (!d_packed || !d_T || !d_TT || !d_work || !d_out)
    return fail(GBX_ERR_ARG, "gbx_mo_transform: null argument");
  if (notation != 0 && notation != 1)
    return fail(GBX_ERR_ARG, "gbx_mo_transform: notation must be 0 (chemist) or 1 (physicist)");
  // rows of one chunk stay below 2^27 (or one matrix's n): the 1-D grids stay in range
  if (nfunc < 1 || nfunc > 32768 || nnew < 1 || nnew > 32768)
    return fail(GBX_ERR_ARG, "gbx_mo_transform: nfunc and nnew must be in [1, 32768]");
  int64_t half_doubles = 0;
  mo_workspace_doubles(nfunc, nnew, &half_doubles);
  hipStream_t s = static_cast<hipStream_t>(stream);
  MoArgs a{};
  a.packed = d_packed;
  a.half = d_work;
  a.out = d_out;
  a.TT = d_TT;
  a.u = d_work + half_doubles;
  a.n = nfunc;
  a.m = nnew;
  a.npair_n = nfunc * (nfunc + 1) / 2;
  a.notation = notation;
  if (int rc = mo_pass<1>(a, a.npair_n, s)) return rc;
  if (int rc = mo_pass<2>(a, nnew * (nnew + 1) / 2, s)) return rc;
  return mo_fill(d_out, nnew, notation, s);
}
