// Four-index transform of the canonical 8-fold packed ERIs of one molecule into a
// new basis (e.g. molecular orbitals), without a dense N^4 intermediate:
//
//   (ab|cd) = sum_ijkl T_ai T_bj T_ck T_dl (ij|kl)
//
// (gbasis/base_four_symm.py:559-616, four tensordot passes over the dense array in
// the reference).  Both halves are the same operation on a batch of symmetric
// matrices, Y = T X T^T with X symmetric n x n and T m x n:
//   pass 1, one matrix per AO pair ij (i >= j): X_ij[k][l] = (ij|kl) gathered from the
//     packed array; its lower triangle Y[c][d] (c >= d) goes to the half-transformed
//     array H[cd][ij] (Mpair x Npair, Npair = n(n+1)/2, Mpair = m(m+1)/2);
//   pass 2, one matrix per new pair cd: X_cd[i][j] = H[cd][tri(i, j)]; its lower
//     triangle gives (ab|cd) for a >= b, written to the four dense positions of
//     (ab|cd) = (ba|cd) = (ab|dc) = (ba|dc) in chemist or physicist order.
// A workgroup owns one matrix and a tile of kMoTC rows c of Y: V = T[c-tile] X is
// accumulated over kMoKC-row chunks of X staged in LDS (each element of X is
// gathered once per tile), then Y[c][d] = sum_l V[c][l] T[d][l] with T^T read
// coalesced.  FP64 vector FMAs throughout (FP64 MFMA has the same peak on
// gfx950); the passes cost n^2 m + n m^2 / 2 FMAs per matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

constexpr int kMoThreads = 256;
constexpr int kMoTC = 16;   // rows of Y per workgroup
constexpr int kMoKC = 16;   // rows of X per LDS chunk

struct MoArgs {
  const double* packed;  // pass 1 source
  double* half;          // H: pass 1 output, pass 2 source
  double* out;           // dense m^4 output
  const double* T;       // m x n, row-major
  const double* TT;      // n x m (T transposed), row-major
  int64_t n, m, npair_n;
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

template <int kPass>
__device__ inline double mo_source(const MoArgs& a, uint64_t r, uint64_t k, uint64_t l) {
  if constexpr (kPass == 1) return a.packed[mo_tri(r, mo_tri(k, l))];
  else return a.half[r * static_cast<uint64_t>(a.npair_n) + mo_tri(k, l)];
}

template <int kPass>
__global__ __launch_bounds__(kMoThreads) void mo_sandwich_kernel(MoArgs a) {
  extern __shared__ __align__(16) char smem[];
  const int64_t n = a.n, m = a.m;
  double* V = reinterpret_cast<double*>(smem);  // kMoTC x n
  double* Xs = V + kMoTC * n;                   // kMoKC x n
  double* Ts = Xs + kMoKC * n;                  // kMoTC x kMoKC
  const uint64_t r = blockIdx.x;
  const int64_t c0 = static_cast<int64_t>(blockIdx.y) * kMoTC;
  const int tc = static_cast<int>(std::min<int64_t>(kMoTC, m - c0));
  const int tid = threadIdx.x;
  for (int64_t e = tid; e < kMoTC * n; e += kMoThreads) V[e] = 0.0;
  for (int64_t k0 = 0; k0 < n; k0 += kMoKC) {
    const int kc = static_cast<int>(std::min<int64_t>(kMoKC, n - k0));
    __syncthreads();  // previous chunk consumed (and V zeroed)
    for (int64_t e = tid; e < kc * n; e += kMoThreads)
      Xs[e] = mo_source<kPass>(a, r, static_cast<uint64_t>(k0 + e / n),
                               static_cast<uint64_t>(e % n));
    for (int e = tid; e < kMoTC * kMoKC; e += kMoThreads) {
      const int c = e / kMoKC, kk = e % kMoKC;
      Ts[e] = (c < tc && kk < kc) ? a.T[(c0 + c) * n + k0 + kk] : 0.0;
    }
    __syncthreads();
    for (int64_t e = tid; e < tc * n; e += kMoThreads) {
      const int c = static_cast<int>(e / n);
      const int64_t l = e % n;
      double s = V[e];
      for (int kk = 0; kk < kc; ++kk) s += Ts[c * kMoKC + kk] * Xs[kk * n + l];
      V[e] = s;
    }
  }
  __syncthreads();
  uint64_t pc = 0, pd = 0;
  if constexpr (kPass == 2) mo_tri_split(r, pc, pd);
  // lower triangle of the tile: Y[c][d], d <= c
  for (int64_t e = tid; e < tc * m; e += kMoThreads) {
    const int c = static_cast<int>(e / m);
    const int64_t d = e % m;
    const int64_t cg = c0 + c;
    if (d > cg) continue;
    const double* vr = V + c * n;
    double s = 0.0;
    for (int64_t l = 0; l < n; ++l) s += vr[l] * a.TT[l * m + d];
    if constexpr (kPass == 1) {
      a.half[mo_tri(cg, d) * static_cast<uint64_t>(a.npair_n) + r] = s;
    } else {
      // (ab|cd) with (a, b) = (cg, d) and (c, d) = (pc, pd)
      const uint64_t M = m, A = cg, B = d;
      auto put = [&](uint64_t i, uint64_t j, uint64_t k, uint64_t l) {
        // chemist (ij|kl) at [i][j][k][l]; physicist <ik|jl> = (ij|kl) at [i][k][j][l]
        const uint64_t idx = a.notation == 0 ? ((i * M + j) * M + k) * M + l
                                             : ((i * M + k) * M + j) * M + l;
        a.out[idx] = s;
      };
      put(A, B, pc, pd);
      put(B, A, pc, pd);
      put(A, B, pd, pc);
      put(B, A, pd, pc);
    }
  }
}

size_t mo_lds_bytes(int64_t n) {
  return static_cast<size_t>(2 * kMoTC * n + kMoTC * kMoKC) * sizeof(double);
}

template <int kPass>
int mo_launch(const MoArgs& a, uint64_t nmat, int lds_limit, int dev, hipStream_t s) {
  const size_t shmem = mo_lds_bytes(a.n);
  if (shmem > 65536) {
    static bool raised[64] = {};  // per device: the attribute is per-device state
    if (dev < 0 || dev >= 64 || !raised[dev]) {
      if (hipFuncSetAttribute(reinterpret_cast<const void*>(&mo_sandwich_kernel<kPass>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              lds_limit) != hipSuccess)
        return fail(GBX_ERR_HIP, "gbx_mo_transform: could not raise the LDS limit");
      if (dev >= 0 && dev < 64) raised[dev] = true;
    }
  }
  const dim3 grid(static_cast<unsigned>(nmat), static_cast<unsigned>((a.m + kMoTC - 1) / kMoTC));
  hipLaunchKernelGGL(mo_sandwich_kernel<kPass>, grid, dim3(kMoThreads), shmem, s, a);
  return check_launch(kPass == 1 ? "mo_sandwich_kernel<1>" : "mo_sandwich_kernel<2>");
}

}  // namespace gbx

using namespace gbx;

extern "C" int gbx_mo_workspace(int64_t nfunc, int64_t nnew, int64_t* half_elems) {
  if (!half_elems) return fail(GBX_ERR_ARG, "gbx_mo_workspace: null argument");
  if (nfunc < 1 || nnew < 1) return fail(GBX_ERR_ARG, "gbx_mo_workspace: sizes must be >= 1");
  *half_elems = (nfunc * (nfunc + 1) / 2) * (nnew * (nnew + 1) / 2);
  return GBX_OK;
}

extern "C" int gbx_mo_transform(const double* d_packed, int64_t nfunc, const double* d_T,
                                const double* d_TT, int64_t nnew, double* d_half,
                                double* d_out, int32_t notation, void* stream) {
  if (!d_packed || !d_T || !d_TT || !d_half || !d_out)
    return fail(GBX_ERR_ARG, "gbx_mo_transform: null argument");
  if (notation != 0 && notation != 1)
    return fail(GBX_ERR_ARG, "gbx_mo_transform: notation must be 0 (chemist) or 1 (physicist)");
  // the pair counts index the grid (< 2^31) and every dense index stays below 2^63
  if (nfunc < 1 || nfunc > 32768 || nnew < 1 || nnew > 32768)
    return fail(GBX_ERR_ARG, "gbx_mo_transform: nfunc and nnew must be in [1, 32768]");
  const int64_t npair_n = nfunc * (nfunc + 1) / 2, npair_m = nnew * (nnew + 1) / 2;
  if (npair_n >= (int64_t(1) << 31) || npair_m >= (int64_t(1) << 31))
    return fail(GBX_ERR_ARG, "gbx_mo_transform: too many pairs for one grid");
  int dev = 0, lds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return fail(GBX_ERR_HIP, "gbx_mo_transform: device query failed");
  if (mo_lds_bytes(nfunc) > static_cast<size_t>(lds))
    return fail(GBX_ERR_ARG, "gbx_mo_transform: nfunc too large for the LDS tiles");
  hipStream_t s = static_cast<hipStream_t>(stream);
  MoArgs a{d_packed, d_half, d_out, d_T, d_TT, nfunc, nnew, npair_n, notation};
  int rc = mo_launch<1>(a, static_cast<uint64_t>(npair_n), lds, dev, s);
  if (rc != GBX_OK) return rc;
  return mo_launch<2>(a, static_cast<uint64_t>(npair_m), lds, dev, s);
}
