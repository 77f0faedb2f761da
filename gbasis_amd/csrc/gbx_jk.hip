// Coulomb and exchange matrices digested straight from the canonical 8-fold
// packed ERI array of one molecule, without materialising the dense N^4 array.
//
//   J_mn = sum_ls (mn|ls) D_ls        K_ml = sum_ns (mn|ls) D_ns
//
// for a symmetric density D (N x N, row-major).  A unique slot (ij|kl), i >= j,
// k >= l, ij >= kl, stands for the distinct index tuples of its 8-fold orbit.
// With v' = (ij|kl) / ((1 + d_ij)(1 + d_kl)(1 + d_{ij,kl})) the orbit adds
//   J[i][j] += 2 v' D_kl    J[k][l] += 2 v' D_ij
//   K[i][k] += v' D_jl      K[j][k] += v' D_il    K[i][l] += v' D_jk    K[j][l] += v' D_ik
// to one triangle's worth of accumulators, and J = A + A^T, K = B + B^T at the
// end gives the full symmetric matrices.  This is the digestion a
// consumer of gbasis's ERIs performs (notebooks/tutorial/JCP_paper.ipynb,
// cell 28, forms the energies sum_ij 2 (ii|jj) and sum_ij (ij|ij) over occupied
// MOs from the dense MO-basis array; tr(D J) / 2 and tr(D K) / 4 are the same
// numbers for D = 2 C_occ C_occ^T).
//
// The pass is bound by reading the packed array once (8 bytes per slot) and by
// six accumulations per slot.  When both accumulators fit in one workgroup's LDS
// (N <= 100 on gfx950) every workgroup accumulates privately with LDS atomics
// over a grid-stride range of slots and adds its matrices to the output once;
// larger N accumulate with global FP64 atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

constexpr int kJkThreads = 256;

// t = r (r + 1) / 2 + c with r >= c
__device__ inline void tri_split(uint64_t t, uint64_t& r, uint64_t& c) {
  uint64_t x = static_cast<uint64_t>((sqrt(8.0 * static_cast<double>(t) + 1.0) - 1.0) * 0.5);
  while (x * (x + 1) / 2 > t) --x;
  while ((x + 1) * (x + 2) / 2 <= t) ++x;
  r = x;
  c = t - x * (x + 1) / 2;
}

__device__ inline void digest_slot(uint64_t s, double v, const double* __restrict__ D,
                                   int64_t n, double* Ja, double* Ka) {
  uint64_t ij, kl, i, j, k, l;
  tri_split(s, ij, kl);
  tri_split(ij, i, j);
  tri_split(kl, k, l);
  if (i == j) v *= 0.5;
  if (k == l) v *= 0.5;
  if (ij == kl) v *= 0.5;
  atomicAdd(&Ja[i * n + j], 2.0 * v * D[k * n + l]);
  atomicAdd(&Ja[k * n + l], 2.0 * v * D[i * n + j]);
  atomicAdd(&Ka[i * n + k], v * D[j * n + l]);
  atomicAdd(&Ka[j * n + k], v * D[i * n + l]);
  atomicAdd(&Ka[i * n + l], v * D[j * n + k]);
  atomicAdd(&Ka[j * n + l], v * D[i * n + k]);
}

__global__ __launch_bounds__(kJkThreads) void jk_zero_kernel(double* __restrict__ J,
                                                             double* __restrict__ K, int64_t nn) {
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < nn;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    J[e] = 0.0;
    K[e] = 0.0;
  }
}

// LDS-private accumulators: smem = [A (n*n) | B (n*n)]
__global__ __launch_bounds__(kJkThreads) void jk_digest_lds_kernel(
    const double* __restrict__ packed, uint64_t nslot, const double* __restrict__ D, int64_t n,
    double* __restrict__ J, double* __restrict__ K) {
  extern __shared__ __align__(16) char smem[];
  double* A = reinterpret_cast<double*>(smem);
  double* B = A + n * n;
  const int64_t nn = n * n;
  for (int64_t e = threadIdx.x; e < 2 * nn; e += kJkThreads) A[e] = 0.0;
  __syncthreads();
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kJkThreads;
  for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * kJkThreads + threadIdx.x; s < nslot;
       s += stride) {
    const double v = packed[s];
    if (v != 0.0) digest_slot(s, v, D, n, A, B);
  }
  __syncthreads();
  for (int64_t e = threadIdx.x; e < nn; e += kJkThreads) {
    if (A[e] != 0.0) atomicAdd(&J[e], A[e]);
    if (B[e] != 0.0) atomicAdd(&K[e], B[e]);
  }
}

__global__ __launch_bounds__(kJkThreads) void jk_digest_global_kernel(
    const double* __restrict__ packed, uint64_t nslot, const double* __restrict__ D, int64_t n,
    double* __restrict__ J, double* __restrict__ K) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kJkThreads;
  for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * kJkThreads + threadIdx.x; s < nslot;
       s += stride) {
    const double v = packed[s];
    if (v != 0.0) digest_slot(s, v, D, n, J, K);
  }
}

// J <- J + J^T, K <- K + K^T in place; one thread per (r >= c)
__global__ __launch_bounds__(kJkThreads) void jk_symmetrize_kernel(double* __restrict__ J,
                                                                   double* __restrict__ K,
                                                                   int64_t n) {
  const uint64_t ntri = static_cast<uint64_t>(n) * (n + 1) / 2;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < ntri;
       t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint64_t r, c;
    tri_split(t, r, c);
    const double j = J[r * n + c] + J[c * n + r];
    const double k = K[r * n + c] + K[c * n + r];
    J[r * n + c] = j;
    J[c * n + r] = j;
    K[r * n + c] = k;
    K[c * n + r] = k;
  }
}

}  // namespace gbx

using namespace gbx;

extern "C" int gbx_jk_packed(const double* d_packed, int64_t nfunc, const double* d_dens,
                             double* d_J, double* d_K, void* stream) {
  if (!d_packed || !d_dens || !d_J || !d_K)
    return fail(GBX_ERR_ARG, "gbx_jk_packed: null argument");
  // 65536 functions keep every slot index below 2^62
  if (nfunc < 1 || nfunc > 65536)
    return fail(GBX_ERR_ARG, "gbx_jk_packed: nfunc must be in [1, 65536]");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = nfunc;
  const int64_t nn = n * n;
  const uint64_t npair = static_cast<uint64_t>(n) * (n + 1) / 2;
  const uint64_t nslot = npair * (npair + 1) / 2;

  int dev = 0, ncu = 0, lds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return fail(GBX_ERR_HIP, "gbx_jk_packed: device query failed");

  const unsigned zgrid =
      static_cast<unsigned>(std::min<int64_t>((nn + kJkThreads - 1) / kJkThreads, 4096));
  hipLaunchKernelGGL(jk_zero_kernel, dim3(zgrid), dim3(kJkThreads), 0, s, d_J, d_K, nn);
  int rc = check_launch("jk_zero_kernel");
  if (rc != GBX_OK) return rc;

  // enough workgroups to fill the GPU, each with 512 slots per thread at least
  const uint64_t want = (nslot + 512 * kJkThreads - 1) / (512 * kJkThreads);
  const size_t shmem = static_cast<size_t>(2 * nn) * sizeof(double);
  if (shmem + 256 <= static_cast<size_t>(lds)) {
    const unsigned grid =
        static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(want, 4 * uint64_t(ncu))));
    if (shmem > 65536) {
      static bool raised[64] = {};  // per device: the attribute is per-device state
      if (dev < 0 || dev >= 64 || !raised[dev]) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&jk_digest_lds_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                lds - 256) != hipSuccess)
          return fail(GBX_ERR_HIP, "gbx_jk_packed: could not raise the LDS limit");
        if (dev >= 0 && dev < 64) raised[dev] = true;
      }
    }
    hipLaunchKernelGGL(jk_digest_lds_kernel, dim3(grid), dim3(kJkThreads), shmem, s, d_packed,
                       nslot, d_dens, n, d_J, d_K);
    rc = check_launch("jk_digest_lds_kernel");
  } else {
    const unsigned grid = static_cast<unsigned>(
        std::max<uint64_t>(1, std::min<uint64_t>(want * 16, 64 * uint64_t(ncu))));
    hipLaunchKernelGGL(jk_digest_global_kernel, dim3(grid), dim3(kJkThreads), 0, s, d_packed,
                       nslot, d_dens, n, d_J, d_K);
    rc = check_launch("jk_digest_global_kernel");
  }
  if (rc != GBX_OK) return rc;
  const uint64_t ntri = static_cast<uint64_t>(n) * (n + 1) / 2;
  const unsigned sgrid =
      static_cast<unsigned>(std::min<uint64_t>((ntri + kJkThreads - 1) / kJkThreads, 4096));
  hipLaunchKernelGGL(jk_symmetrize_kernel, dim3(sgrid), dim3(kJkThreads), 0, s, d_J, d_K, n);
  return check_launch("jk_symmetrize_kernel");
}
