// Shared device helpers of the gbasis_amd kernels: primitive-quartet setup and
// the Boys function F_m(T), plus host-side error plumbing of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "gbasis_amd.h"

namespace gbx {

constexpr int kPrimRec = 10;          // doubles per primitive-pair record
constexpr int kTaylorTerms = 8;       // Taylor terms of the Boys table expansion
constexpr int kMaxShellL = 6;         // largest l per shell the host may ask for
constexpr int64_t kCoopLdsLimit = 65536;
constexpr int kXcdCount = 8;          // XCDs of an MI355X; workgroups go round-robin to them
constexpr int kXcdChunk = 16;         // consecutive work items per XCD in xcd_chunked()

// The hardware deals workgroup b of a G-workgroup grid to XCD b mod 8.  This maps b to a
// work item so that each XCD runs chunks of 16 consecutive items (chunk c on XCD c mod 8):
// the items of one chunk share the 128-byte lines of item-minor arrays (accumulators
// [f][quartet]) in that XCD's L2, and the chunks stay interleaved over the XCDs, so a
// cost-sorted queue still loads them evenly.  A bijection on [0, G) (the last G mod 128
// workgroups keep their own index).
__device__ inline int64_t xcd_chunked(int64_t b, int64_t G) {
  const int64_t full = G / (kXcdCount * kXcdChunk) * (kXcdCount * kXcdChunk);
  if (b >= full) return b;
  const int64_t x = b % kXcdCount, i = b / kXcdCount;
  return ((i / kXcdChunk) * kXcdCount + x) * kXcdChunk + i % kXcdChunk;
}

// host: record an error message for gbx_last_error() and return the code
int fail(int code, const std::string& msg);
// host: turn a pending launch error into a C-ABI error code
int check_launch(const char* what);

// offset of the factorial2 norms of angular momentum l in the f2norm table
__host__ __device__ inline int f2_offset(int l) { return l * (l + 1) * (l + 2) / 6; }

// Scalars of one primitive quartet (bra primitive pair x ket primitive pair),
// in the notation of gbasis/integrals/_two_elec_int_improved.py:651-697.
template <typename T>
struct PrimQuartet {
  T PA[3];   // P - A
  T WP[3];   // -(rho/zeta) (P - Q)
  T roz;     // rho / zeta
  T i2z;     // 1 / (2 zeta)
  T QC[3];   // Q - C
  T WQ[3];   // (rho/eta) (P - Q)
  T roe;     // rho / eta
  T i2e;     // 1 / (2 eta)
  T i2s;     // 1 / (2 (zeta + eta))
  double pref;  // 2 pi^(5/2) / (zeta eta sqrt(zeta+eta)) * K_ab K_cd (always FP64)
};

// rb / rk: primitive-pair records {zeta, 1/(2zeta), P(3), P-A(3), K', 0}.
// Returns T = rho |P-Q|^2.
template <typename T>
__device__ inline double prim_quartet_setup(const double* __restrict__ rb,
                                            const double* __restrict__ rk, PrimQuartet<T>& s) {
  const double zeta = rb[0], i2z = rb[1], eta = rk[0], i2e = rk[1];
  // one reciprocal square root gives both 1/(zeta+eta) and its square root
  const double rs = rsqrt(zeta + eta);
  const double inv_s = rs * rs;
  const double rho = zeta * eta * inv_s;
  const double PQx = rb[2] - rk[2], PQy = rb[3] - rk[3], PQz = rb[4] - rk[4];
  const double roz = eta * inv_s;
  const double roe = zeta * inv_s;
  s.PA[0] = static_cast<T>(rb[5]);
  s.PA[1] = static_cast<T>(rb[6]);
  s.PA[2] = static_cast<T>(rb[7]);
  s.WP[0] = static_cast<T>(-roz * PQx);
  s.WP[1] = static_cast<T>(-roz * PQy);
  s.WP[2] = static_cast<T>(-roz * PQz);
  s.roz = static_cast<T>(roz);
  s.i2z = static_cast<T>(i2z);
  s.QC[0] = static_cast<T>(rk[5]);
  s.QC[1] = static_cast<T>(rk[6]);
  s.QC[2] = static_cast<T>(rk[7]);
  s.WQ[0] = static_cast<T>(roe * PQx);
  s.WQ[1] = static_cast<T>(roe * PQy);
  s.WQ[2] = static_cast<T>(roe * PQz);
  s.roe = static_cast<T>(roe);
  s.i2e = static_cast<T>(i2e);
  s.i2s = static_cast<T>(0.5 * inv_s);
  s.pref = rb[8] * rk[8] * rs;
  return rho * (PQx * PQx + PQy * PQy + PQz * PQz);
}

// 1 / (2m + 1), m = 0..31
__constant__ static const double kInvOdd[32] = {
    1.0,       1.0 / 3,  1.0 / 5,  1.0 / 7,  1.0 / 9,  1.0 / 11, 1.0 / 13, 1.0 / 15,
    1.0 / 17,  1.0 / 19, 1.0 / 21, 1.0 / 23, 1.0 / 25, 1.0 / 27, 1.0 / 29, 1.0 / 31,
    1.0 / 33,  1.0 / 35, 1.0 / 37, 1.0 / 39, 1.0 / 41, 1.0 / 43, 1.0 / 45, 1.0 / 47,
    1.0 / 49,  1.0 / 51, 1.0 / 53, 1.0 / 55, 1.0 / 57, 1.0 / 59, 1.0 / 61, 1.0 / 63};

// e^x for |x| <= 0.05 (half a Boys grid step): Taylor polynomial of degree 8,
// truncation error below 6e-18 relative.
__device__ inline double exp_small(double x) {
  double p = 1.0 / 40320.0;
  p = p * x + 1.0 / 5040.0;
  p = p * x + 1.0 / 720.0;
  p = p * x + 1.0 / 120.0;
  p = p * x + 1.0 / 24.0;
  p = p * x + 1.0 / 6.0;
  p = p * x + 0.5;
  p = p * x + 1.0;
  return p * x + 1.0;
}

constexpr double kHalfSqrtPi = 0.88622692545275801365;  // sqrt(pi) / 2

// Boys function F_m(T), m = 0..L, delivered through out(m, value).
//
// For T on the tabulated range [0, (nt-1) h]: F_L from an 8-term Taylor
// expansion about the nearest grid point T_k,
//     F_L(T) = sum_j F_{L+j}(T_k) (T_k - T)^j / j!      (dF_n/dT = -F_{n+1}),
// then the downward recurrence F_m = (2T F_{m+1} + e^{-T}) / (2m+1), which is
// stable, with e^{-T} = e^{-T_k} e^{T_k - T} (tabulated e^{-T_k}, exp_small).
// Beyond the table, F_0 = sqrt(pi/T)/2 (erf(sqrt T) = 1 to double precision
// for T > 36) and the upward recurrence F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T),
// stable for T > m.  Row k of the table is e^{-k h}, then F_n(k h) for
// n < nord - 1, computed on the host in extended precision from the series
// F_n(T) = e^{-T} sum_i (2T)^i / ((2n+1)(2n+3)...(2n+2i+1)).
template <typename Out>
__device__ inline void boys_eval(double T, int L, const gbx_tables& t, Out out) {
  const double tmax = t.boys_h * (t.boys_nt - 1);
  const double inv_h = 1.0 / t.boys_h;  // uniform: computed once per kernel
  if (T <= tmax) {
    const int k = static_cast<int>(T * inv_h + 0.5);
    const double x = k * t.boys_h - T;
    const double* row = t.boys_table + static_cast<int64_t>(k) * t.boys_nord;
    const double emT = row[0] * exp_small(x);
    row += 1 + L;
    // Horner form of sum_j row[j] x^j / j!
    double h = row[7];
    h = row[6] + x * h * (1.0 / 7.0);
    h = row[5] + x * h * (1.0 / 6.0);
    h = row[4] + x * h * (1.0 / 5.0);
    h = row[3] + x * h * (1.0 / 4.0);
    h = row[2] + x * h * (1.0 / 3.0);
    h = row[1] + x * h * (1.0 / 2.0);
    h = row[0] + x * h;
    double fm = h;
    out(L, fm);
    const double twoT = 2.0 * T;
    for (int m = L - 1; m >= 0; --m) {
      fm = (twoT * fm + emT) * kInvOdd[m];
      out(m, fm);
    }
  } else {
    const double emT = exp(-T);
    const double r = rsqrt(T);
    const double inv2T = 0.5 * r * r;
    double fm = kHalfSqrtPi * r;
    out(0, fm);
    for (int m = 0; m < L; ++m) {
      fm = (static_cast<double>(2 * m + 1) * fm - emT) * inv2T;
      out(m + 1, fm);
    }
  }
}

}  // namespace gbx
