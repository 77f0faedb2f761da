// Primitive pass of the high angular momentum classes by Rys quadrature.
//
// For one work item (shell quartet qi, part r of its primitive quartets) a workgroup
// sums, over its primitive quartets and the n = L/2 + 1 Rys nodes u (x = t^2) of each,
// pref * W_u * Ix(ax, cx) Iy(ay, cy) Iz(az, cz) into the contracted [a0|c0] of every
// accumulator of the class program (bra component a of degree la..la+lb, ket
// component c of degree lc..lc+ld), and writes them to
// acc_out[(f * parts + r) * acc_stride + qi] -- the layout the cooperative horizontal
// recursion pass (gbx_eri_hrr) reads.  It replaces the Boys function and the vertical
// and ket-centre recursions of gbasis/integrals/_two_elec_int_improved.py:87-361 (the
// same primitive integrals: the Rys rule integrates the Boys moments F_m(T), m < 2n,
// exactly).
//
// One-dimensional integrals of a node, direction i (I(0, 0) = 1; pref * W_u for z):
//   I(e+1, 0) = C00 I(e, 0) + e B10 I(e-1, 0)
//   I(e, f+1) = D00 I(e, f) + f B01 I(e, f-1) + e B00 I(e-1, f)
//   C00 = (P-A)_i + u (W-P)_i, D00 = (Q-C)_i + u (W-Q)_i,
//   B10 = (1 - u rho/zeta) / (2 zeta), B01 = (1 - u rho/eta) / (2 eta), B00 = u / (2(zeta+eta)).
// A batch of kRysPB primitive quartets has its n * 3 tables built in LDS by one thread
// each; then every thread adds, for each of its accumulators, the products over the
// batch's nodes (the node loop is unrolled, so the table reads use immediate offsets).
// Nodes and weights come from the table of gbasis_amd/rys.py: Chebyshev interpolants
// on unit intervals of T < kRysTmax, the half-range Hermite rule beyond.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

constexpr int kRysThreads = 512;
constexpr int kRysPB = 2;          // primitive quartets per LDS batch
constexpr int kRysKmax = 32;       // accumulators per thread and launch
constexpr int kRysNmax = 9;
constexpr int kRysNseg = 100;
constexpr int kRysDeg = 13;
constexpr int kRysNcoef = kRysDeg + 1;
constexpr double kRysTmax = 100.0;
constexpr int64_t kRysTableSize =
    int64_t(kRysNseg) * kRysNcoef * kRysNmax * (kRysNmax + 1) + kRysNmax * (kRysNmax + 1);

__host__ __device__ inline int64_t rys_off(int n) {
  return int64_t(kRysNseg) * kRysNcoef * n * (n - 1);
}
__host__ __device__ inline int64_t rys_hoff(int n) {
  return int64_t(kRysNseg) * kRysNcoef * kRysNmax * (kRysNmax + 1) + n * (n - 1);
}

// Node x_r = t_r^2 and weight W_r of the n-point rule at T.
__device__ inline void rys_node(const double* __restrict__ tab, int n, int r, double T, double& x,
                                double& w) {
  if (T >= kRysTmax) {
    const double* h = tab + rys_hoff(n);
    x = h[r] / T;
    w = h[n + r] * rsqrt(T);
    return;
  }
  const int k = min(static_cast<int>(T), kRysNseg - 1);
  const double s = 2.0 * (T - k) - 1.0;
  const double s2 = 2.0 * s;
  const double* c = tab + rys_off(n) + static_cast<int64_t>(k) * kRysNcoef * 2 * n;
  double bx1 = 0.0, bx2 = 0.0, bw1 = 0.0, bw2 = 0.0;
  // not unrolled: the coefficients are loaded as they are used, which keeps the
  // registers of the accumulators live across this phase
#pragma unroll 1
  for (int j = kRysDeg; j >= 1; --j) {
    const double tx = c[j * 2 * n + r] + s2 * bx1 - bx2;
    const double tw = c[j * 2 * n + n + r] + s2 * bw1 - bw2;
    bx2 = bx1;
    bx1 = tx;
    bw2 = bw1;
    bw1 = tw;
  }
  x = c[r] + s * bx1 - bx2;
  w = c[n + r] + s * bw1 - bw2;
}

struct RysArgs {
  const int32_t* quart;
  int64_t nquart;
  const int32_t* count;
  const double* rys;      // node table
  const int32_t* acc_idx; // per accumulator: ox | oy << 10 | oz << 20, offsets of
                          //   (ax, cx), (ay, cy), (az, cz) in a (lab+1) x (lcd+1) table
  int32_t lab, lcd, nacc, kacc;
  int32_t acc_lo;         // this launch's accumulators: [acc_lo, acc_lo + kacc * threads)
  int32_t parts;
  double* acc_out;
  int64_t acc_stride;
};

template <int NR, int KMAX>
__global__ __launch_bounds__(kRysThreads) void eri_rys_acc_kernel(gbx_tables t, RysArgs a) {
  extern __shared__ __align__(16) double rtab[];
  const int tid = threadIdx.x;
  const int W = a.lcd + 1;
  const int TS = (a.lab + 1) * W;
  // this thread's accumulators f = tid + k * kRysThreads and their table offsets
  int idx[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int f = a.acc_lo + tid + k * kRysThreads;
    idx[k] = (k < a.kacc && f < a.nacc) ? a.acc_idx[f] : -1;
  }
  const int4* pinfo = reinterpret_cast<const int4*>(t.pair_info);
  const int64_t nquart = a.count ? min(a.nquart, static_cast<int64_t>(*a.count)) : a.nquart;
  for (int64_t wi = blockIdx.x; wi < nquart * a.parts; wi += gridDim.x) {
    const int64_t qi = wi % nquart;
    const int part = static_cast<int>(wi / nquart);
    const int bra = a.quart[2 * qi], ket = a.quart[2 * qi + 1];
    const int4 pb = pinfo[bra];
    const int4 pk = pinfo[ket];
    const int nk = pk.w;
    const int nprimq = pb.w * nk;
    const int chunk = (nprimq + a.parts - 1) / a.parts;
    const int pbeg = min(nprimq, part * chunk);
    const int pend = min(nprimq, pbeg + chunk);
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) acc[k] = 0.0;

    for (int p0 = pbeg; p0 < pend; p0 += kRysPB) {
      const int ncopy = min(kRysPB, pend - p0);
      __syncthreads();  // the previous batch's tables are no longer read
      if (tid < kRysPB * NR * 3) {
        const int p = tid / (3 * NR);
        const int rem = tid - p * 3 * NR;
        const int r = rem / 3;
        const int d = rem - r * 3;
        double* I = rtab + ((p * NR + r) * 3 + d) * TS;
        if (p < ncopy) {
          const int pp = p0 + p;
          const int ib = pp / nk, ik = pp - ib * nk;
          const double* rb = t.prim_pairs + static_cast<int64_t>(pb.z + ib) * kPrimRec;
          const double* rk = t.prim_pairs + static_cast<int64_t>(pk.z + ik) * kPrimRec;
          PrimQuartet<double> s;
          const double T = prim_quartet_setup(rb, rk, s);
          double u, wt;
          rys_node(a.rys, NR, r, T, u, wt);
          const double pa = d == 0 ? s.PA[0] : (d == 1 ? s.PA[1] : s.PA[2]);
          const double wp = d == 0 ? s.WP[0] : (d == 1 ? s.WP[1] : s.WP[2]);
          const double qc = d == 0 ? s.QC[0] : (d == 1 ? s.QC[1] : s.QC[2]);
          const double wq = d == 0 ? s.WQ[0] : (d == 1 ? s.WQ[1] : s.WQ[2]);
          const double c00 = pa + u * wp, d00 = qc + u * wq;
          const double b10 = s.i2z * (1.0 - u * s.roz);
          const double b01 = s.i2e * (1.0 - u * s.roe);
          const double b00 = u * s.i2s;
          I[0] = d == 2 ? s.pref * wt : 1.0;
          if (a.lab > 0) I[W] = c00 * I[0];
          for (int e = 1; e < a.lab; ++e)
            I[(e + 1) * W] = c00 * I[e * W] + static_cast<double>(e) * b10 * I[(e - 1) * W];
          for (int f = 0; f < a.lcd; ++f) {
            const double fb01 = static_cast<double>(f) * b01;
            for (int e = 0; e <= a.lab; ++e) {
              double v = d00 * I[e * W + f];
              if (f > 0) v += fb01 * I[e * W + f - 1];
              if (e > 0) v += static_cast<double>(e) * b00 * I[(e - 1) * W + f];
              I[e * W + f + 1] = v;
            }
          }
        } else {
          for (int e = 0; e < TS; ++e) I[e] = 0.0;  // an absent primitive adds nothing
        }
      }
      __syncthreads();
      // the node stride, opaque to the compiler in each batch: otherwise it hoists the
      // (accumulator, node) addresses out of the batch loop and spills them
      int stride = 3 * TS;
      __asm__ volatile("" : "+s"(stride));
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        if (idx[k] < 0) continue;
        // likewise the table offsets: decoded again in each batch, three registers fewer
        // per accumulator than keeping the decoded addresses
        __asm__ volatile("" : "+v"(idx[k]));
        const double* X = rtab + (idx[k] & 1023);
        const double* Y = rtab + TS + ((idx[k] >> 10) & 1023);
        const double* Z = rtab + 2 * TS + (idx[k] >> 20);
        double s = acc[k];
#pragma unroll
        for (int pr = 0; pr < kRysPB * NR; ++pr) {
          const int o = pr * stride;
          s = fma(X[o] * Y[o], Z[o], s);
        }
        acc[k] = s;
      }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      if (idx[k] < 0) continue;
      const int f = a.acc_lo + tid + k * kRysThreads;
      a.acc_out[(static_cast<int64_t>(f) * a.parts + part) * a.acc_stride + qi] = acc[k];
    }
  }
}

template <int NR>
int launch_rys_nr(const gbx_tables* t, const RysArgs& a, int64_t grid, size_t shmem,
                  hipStream_t s) {
  const unsigned g = static_cast<unsigned>(grid);
  if (a.kacc <= 8)
    hipLaunchKernelGGL((eri_rys_acc_kernel<NR, 8>), dim3(g), dim3(kRysThreads), shmem, s, *t, a);
  else if (a.kacc <= 16)
    hipLaunchKernelGGL((eri_rys_acc_kernel<NR, 16>), dim3(g), dim3(kRysThreads), shmem, s, *t, a);
  else if (a.kacc <= 24)
    hipLaunchKernelGGL((eri_rys_acc_kernel<NR, 24>), dim3(g), dim3(kRysThreads), shmem, s, *t, a);
  else
    hipLaunchKernelGGL((eri_rys_acc_kernel<NR, 32>), dim3(g), dim3(kRysThreads), shmem, s, *t, a);
  return check_launch("eri_rys_acc_kernel");
}

}  // namespace gbx

using namespace gbx;

extern "C" int64_t gbx_rys_table_size(void) { return kRysTableSize; }

extern "C" int gbx_eri_rys_acc(const gbx_tables* t, const double* d_rys, const int32_t* d_acc_idx,
                               int32_t la, int32_t lb, int32_t lc, int32_t ld,
                               const int32_t* d_quartets, int64_t nquart, const int32_t* d_count,
                               int32_t parts, double* d_acc, int64_t acc_stride, void* stream) {
  if (!t || (nquart > 0 && (!d_quartets || !d_acc || !d_rys || !d_acc_idx)))
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: null argument");
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > kMaxShellL || lb > kMaxShellL ||
      lc > kMaxShellL || ld > kMaxShellL)
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: angular momentum out of range");
  if (t->potential != GBX_POT_COULOMB)
    return fail(GBX_ERR_UNSUPPORTED, "gbx_eri_rys_acc: the Rys rule integrates 1/r only");
  const int L = la + lb + lc + ld;
  const int nr = L / 2 + 1;
  if (nr < 5 || nr > kRysNmax)
    return fail(GBX_ERR_UNSUPPORTED, "gbx_eri_rys_acc: classes with 8 <= L <= 17 only");
  if (parts < 1 || parts > 4096)
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: parts must be in [1, 4096]");
  if (acc_stride < nquart) return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: acc_stride < nquart");
  const int lab = la + lb, lcd = lc + ld;
  // accumulators: S(la, lab) x S(lc, lcd) components
  auto ncomp = [](int lo, int hi) {
    int n = 0;
    for (int l = lo; l <= hi; ++l) n += (l + 1) * (l + 2) / 2;
    return n;
  };
  const int nacc = ncomp(la, lab) * ncomp(lc, lcd);
  // at most kRysKmax accumulators per thread (the registers of one lane): classes with
  // more run as several launches over consecutive accumulator ranges
  const int kall = (nacc + kRysThreads - 1) / kRysThreads;
  const int nsplit = (kall + kRysKmax - 1) / kRysKmax;
  const int kacc = (kall + nsplit - 1) / nsplit;
  if (nsplit > 4) return fail(GBX_ERR_UNSUPPORTED, "gbx_eri_rys_acc: too many accumulators");
  if (nquart == 0) return GBX_OK;
  RysArgs a{};
  a.quart = d_quartets;
  a.nquart = nquart;
  a.count = d_count;
  a.rys = d_rys;
  a.acc_idx = d_acc_idx;
  a.lab = lab;
  a.lcd = lcd;
  a.nacc = nacc;
  a.kacc = kacc;
  a.parts = parts;
  a.acc_out = d_acc;
  a.acc_stride = acc_stride;
  const size_t shmem = sizeof(double) * kRysPB * nr * 3 * (lab + 1) * (lcd + 1);
  if (shmem > 65536) return fail(GBX_ERR_UNSUPPORTED, "gbx_eri_rys_acc: node tables exceed 64 KiB");
  const int64_t grid = std::min<int64_t>(nquart * parts, int64_t(1) << 20);
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int sp = 0; sp < nsplit; ++sp) {
    a.acc_lo = sp * kacc * kRysThreads;
    int rc = GBX_OK;
    switch (nr) {
      case 5: rc = launch_rys_nr<5>(t, a, grid, shmem, s); break;
      case 6: rc = launch_rys_nr<6>(t, a, grid, shmem, s); break;
      case 7: rc = launch_rys_nr<7>(t, a, grid, shmem, s); break;
      case 8: rc = launch_rys_nr<8>(t, a, grid, shmem, s); break;
      default: rc = launch_rys_nr<9>(t, a, grid, shmem, s); break;
    }
    if (rc != GBX_OK) return rc;
  }
  return GBX_OK;
}
