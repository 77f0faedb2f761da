// C-ABI plumbing of gbasis_amd (error strings, version) and the device Boys
// function entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(GBX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return GBX_OK;
}

__global__ void boys_kernel(gbx_tables t, const double* __restrict__ T, int64_t n, int mmax,
                            double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* o = out + i * (mmax + 1);
  const double Ti = T[i];
  if (!(Ti >= 0.0)) {  // negative or NaN: no table row; the ERI kernels never form one
    for (int m = 0; m <= mmax; ++m) o[m] = __builtin_nan("");
    return;
  }
  boys_eval(Ti, mmax, t, [&](int m, double v) { o[m] = v; });
}

__device__ inline uint64_t tri_idx(uint64_t i, uint64_t j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// dense[a,b,c,d] from the canonical packed slots; chemist: (ab|cd),
// physicist: <ab|cd> = (ac|bd).
__global__ void unpack_kernel(const double* __restrict__ packed, int64_t n, int physicist,
                              double* __restrict__ dense) {
  const int64_t total = n * n * n * n;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int64_t r = e;
    const uint64_t d = r % n; r /= n;
    const uint64_t c = r % n; r /= n;
    const uint64_t b = r % n;
    const uint64_t a = r / n;
    const uint64_t ij = physicist ? tri_idx(a, c) : tri_idx(a, b);
    const uint64_t kl = physicist ? tri_idx(b, d) : tri_idx(c, d);
    dense[e] = packed[tri_idx(ij, kl)];
  }
}

}  // namespace gbx

extern "C" int gbx_unpack_dense(const double* d_packed, int64_t nfunc, int32_t notation,
                                double* d_dense, void* stream) {
  if (nfunc < 0 || (nfunc > 0 && (!d_packed || !d_dense)) || (notation != 0 && notation != 1))
    return gbx::fail(GBX_ERR_ARG, "gbx_unpack_dense: bad arguments");
  // nfunc^4 elements must fit int64 (and any dense array that large is far beyond memory)
  if (nfunc > 32768) return gbx::fail(GBX_ERR_ARG, "gbx_unpack_dense: nfunc exceeds 32768");
  if (nfunc == 0) return GBX_OK;
  const int64_t total = nfunc * nfunc * nfunc * nfunc;
  const int threads = 256;
  const int64_t blocks = std::min<int64_t>((total + threads - 1) / threads, 1 << 20);
  hipLaunchKernelGGL(gbx::unpack_kernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0,
                     static_cast<hipStream_t>(stream), d_packed, nfunc, notation, d_dense);
  return gbx::check_launch("unpack_kernel");
}

extern "C" int gbx_version(void) { return 1; }

extern "C" const char* gbx_last_error(void) { return gbx::g_last_error.c_str(); }

extern "C" int gbx_boys(const gbx_tables* t, const double* d_T, int64_t n, int32_t mmax,
                        double* d_out, void* stream) {
  if (!t || (n > 0 && (!d_T || !d_out))) return gbx::fail(GBX_ERR_ARG, "gbx_boys: null argument");
  if (mmax < 0 || mmax + gbx::kTaylorTerms > t->boys_nord - 1)
    return gbx::fail(GBX_ERR_ARG, "gbx_boys: mmax exceeds the Boys table");
  if (n == 0) return GBX_OK;
  const int threads = 256;
  const int64_t blocks = (n + threads - 1) / threads;
  hipLaunchKernelGGL(gbx::boys_kernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0,
                     static_cast<hipStream_t>(stream), *t, d_T, n, mmax, d_out);
  return gbx::check_launch("boys_kernel");
}
