// Primitive pass of the high angular momentum classes by Rys quadrature.
//
// For one work item (shell quartet qi, part r of its primitive quartets) a workgroup
// sums, over its primitive quartets and the n = L/2 + 1 Rys nodes u (x = t^2) of each,
// pref * W_u * Ix(ax, cx) Iy(ay, cy) Iz(az, cz) into the contracted [a0|c0] of every
// accumulator of the class program (bra component a of degree la..la+lb, ket
// component c of degree lc..lc+ld), and writes them to
// acc_out[(f * parts + r) * acc_stride + qi] -- the layout the cooperative horizontal
// recursion pass (gbx_eri_hrr) reads -- or acc_out[(qi * parts + r) * acc_stride + f],
// the quartet-major layout of the LDS pass (gbx_eri_hrr_lds), whose writes by
// consecutive lanes are contiguous.  It replaces the Boys function and the vertical
// and ket-centre recursions of gbasis/integrals/_two_elec_int_improved.py:87-361 (the
// same primitive integrals: the Rys rule integrates the Boys moments F_m(T), m < 2n,
// exactly).
//
// One-dimensional integrals of a node, direction i (I(0, 0) = 1; pref * W_u for z):
//   I(e+1, 0) = C00 I(e, 0) + e B10 I(e-1, 0)
//   I(e, f+1) = D00 I(e, f) + f B01 I(e, f-1) + e B00 I(e-1, f)
//   C00 = (P-A)_i + u (W-P)_i, D00 = (Q-C)_i + u (W-Q)_i,
//   B10 = (1 - u rho/zeta) / (2 zeta), B01 = (1 - u rho/eta) / (2 eta), B00 = u / (2(zeta+eta)).
// A batch of primitive quartets (as many as their tables fit in 64 KiB of LDS) has its
// n * 3 tables built in LDS by one thread each, column by column in registers; two
// layouts of the accumulators over the lanes follow, the faster one chosen per class:
// strided (lane tid holds f = tid + 512 k; three table reads and one FMA per node), or
// grouped (every lane adds the batch's products into the accumulators of its group:
// the [a0|c0] that share (ax, ay, cx, cy), and with them one product Ix Iy per node, so
// an accumulator costs one table read and one FMA per node; groups of one shape fill a
// wavefront (laid out on the host), and blockIdx.y selects a block of 512 slots).
// Nodes and weights come from the table of gbasis_amd/rys.py: Chebyshev interpolants
// on unit intervals of T < kRysTmax, the half-range Hermite rule beyond.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

constexpr int kRysThreads = 512;
constexpr int kRysPBmax = 32;      // primitive quartets per LDS batch, at most (the
                                   // default; a caller may ask for fewer)
constexpr int kRysMaxE = 9;        // la + lb + 1 <= 9 (l <= 4 per shell)
constexpr int kRysMaxLb = 4;       // second shell of bra and ket: l <= 4 (groups of <= 5 x 5)
constexpr int kRysKmax = 24;       // strided layout: accumulators per lane and launch
constexpr int kRysStrided = 0, kRysGrouped = 1;
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
  const int32_t* groups;  // grouped layout: per lane slot a group descriptor (below) or -1,
                          //   then per slot MA x MC accumulator rows f of its members (-1:
                          //   none); strided layout: per accumulator f, ox | oy << 10 |
                          //   oz << 20 (its three table offsets)
  int32_t nslots;         // grouped: lane slots; strided: accumulators
  int32_t lab, lcd;
  int32_t pb;             // primitive quartets per batch (their tables fill the LDS)
  int32_t kacc, acc_lo;   // strided layout: accumulators per lane, first of this launch
  int32_t parts;
  double* acc_out;
  int64_t acc_stride;
  int32_t qmajor;         // 1: acc_out[(qi * parts + r) * acc_stride + f]
};

// Address of accumulator f of work item (qi, part) in either layout.
__device__ inline int64_t rys_acc_index(const RysArgs& a, int64_t qi, int part, int f) {
  return a.qmajor ? (qi * a.parts + part) * a.acc_stride + f
                  : (static_cast<int64_t>(f) * a.parts + part) * a.acc_stride + qi;
}

// One thread's part of a batch's node tables: thread tid < PB NR 3 builds the table of
// primitive quartet p0 + tid / (3 NR), node (tid / 3) % NR, direction tid % 3, column by
// column in registers (column f: I(e, f), e <= lab), each column stored once in LDS
// (entries beyond lab are computed and never stored); an absent primitive's tables are
// zero.
__device__ inline void rys_build(const gbx_tables& t, const RysArgs& a, int4 pb, int4 pk, int p0,
                                 int ncopy, int NR, double* rtab) {
  const int tid = threadIdx.x;
  const int W = a.lcd + 1;
  const int TS = (a.lab + 1) * W;
  const int nk = pk.w;
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
    // the quantities of prim_quartet_setup (gbx_common.h) this direction needs, with the
    // same operations, taken before the node evaluation so that few stay live across it
    const double zeta = rb[0], eta = rk[0];
    const double rs = rsqrt(zeta + eta);
    const double inv_s = rs * rs;
    const double rho = zeta * eta * inv_s;
    const double PQx = rb[2] - rk[2], PQy = rb[3] - rk[3], PQz = rb[4] - rk[4];
    const double PQd = d == 0 ? PQx : (d == 1 ? PQy : PQz);
    const double roz = eta * inv_s, roe = zeta * inv_s;
    const double pa = rb[5 + d], qc = rk[5 + d];
    const double wp = -roz * PQd, wq = roe * PQd;
    const double i2z = rb[1], i2e = rk[1], i2s = 0.5 * inv_s;
    const double pref = d == 2 ? rb[8] * rk[8] * rs : 1.0;
    const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
    double u, wt;
    rys_node(a.rys, NR, r, T, u, wt);
    const double c00 = pa + u * wp, d00 = qc + u * wq;
    const double b10 = i2z * (1.0 - u * roz);
    const double b01 = i2e * (1.0 - u * roe);
    const double b00 = u * i2s;
    // the table column by column in registers (column f: I(e, f), e <= lab), each
    // column stored once; entries beyond lab are computed and never stored
    double col[kRysMaxE], prv[kRysMaxE];
    col[0] = d == 2 ? pref * wt : 1.0;
    prv[0] = 0.0;
#pragma unroll
    for (int e = 1; e < kRysMaxE; ++e) {
      col[e] = c00 * col[e - 1] + (e > 1 ? static_cast<double>(e - 1) * b10 * col[e - 2] : 0.0);
      prv[e] = 0.0;
    }
#pragma unroll
    for (int e = 0; e < kRysMaxE; ++e)
      if (e <= a.lab) I[e * W] = col[e];
    for (int f = 0; f < a.lcd; ++f) {
      const double fb01 = static_cast<double>(f) * b01;
      // column f + 1 in place, from the top entry down: entry e reads col[e - 1] of
      // column f, which is still in place
#pragma unroll
      for (int e = kRysMaxE - 1; e >= 0; --e) {
        double v = d00 * col[e] + fb01 * prv[e];
        if (e > 0) v += static_cast<double>(e) * b00 * col[e - 1];
        prv[e] = col[e];
        col[e] = v;
        if (e <= a.lab) I[e * W + f + 1] = v;
      }
    }
  } else {
    for (int e = 0; e < TS; ++e) I[e] = 0.0;  // an absent primitive adds nothing
  }
}

// Group descriptor of a lane: the accumulators [a0|c0] with a = (ax, ay, az_lo + i),
// c = (cx, cy, cz_lo + j), i < alen, j < clen share Ix(ax, cx) Iy(ay, cy):
//   ox = ax (lcd+1) + cx | oy = ay (lcd+1) + cy << 7 | az_lo << 14 | alen << 18 |
//   cz_lo << 22 | clen << 26.
// The lanes of a wavefront hold groups of one shape (alen, clen).
// Both layouts are compiled for 4 waves per SIMD (two 512-lane workgroups per CU, at
// most 128 VGPRs).  The lanes with 24 strided or 25 grouped accumulators then keep part
// of their state in scratch; measured on the g classes, the second workgroup per CU
// still shortens them by 20-40 % (it runs while the first builds its node tables).
constexpr int kRysWaves = 4;

// The nodes of one LDS batch for a group of exactly A x C members: straight-line table
// reads and FMAs (no per-member branch, so the reads of a node are all in flight before
// its FMAs wait on them).
template <int A, int C, int MA, int MC>
__device__ inline void rys_group_nodes(const double* rtab, int nodes, int TS, int W, int ox,
                                       int oy, int oz0, double (&acc)[MA][MC]) {
#pragma unroll 2
  for (int pr = 0; pr < nodes; ++pr) {
    const double* base = rtab + pr * 3 * TS;
    const double xy = base[ox] * base[TS + oy];
    const double* zb = base + 2 * TS + oz0;
#pragma unroll
    for (int i = 0; i < A; ++i)
#pragma unroll
      for (int j = 0; j < C; ++j) acc[i][j] = fma(xy, zb[i * W + j], acc[i][j]);
  }
}

// Shape S = (alen - 1) MC + (clen - 1) of this wavefront's groups -> its exact loop.
template <int MA, int MC, int S = 0>
__device__ inline void rys_group_dispatch(int shape, const double* rtab, int nodes, int TS, int W,
                                          int ox, int oy, int oz0, double (&acc)[MA][MC]) {
  if constexpr (S < MA * MC) {
    if (shape == S) {
      rys_group_nodes<S / MC + 1, S % MC + 1, MA, MC>(rtab, nodes, TS, W, ox, oy, oz0, acc);
      return;
    }
    rys_group_dispatch<MA, MC, S + 1>(shape, rtab, nodes, TS, W, ox, oy, oz0, acc);
  }
}

template <int MA, int MC>
__global__ __launch_bounds__(kRysThreads, kRysWaves) void eri_rys_grouped_kernel(gbx_tables t,
                                                                                RysArgs a) {
  extern __shared__ __align__(16) double rtab[];
  const int tid = threadIdx.x;
  const int W = a.lcd + 1;
  const int TS = (a.lab + 1) * W;
  const int NR = (a.lab + a.lcd) / 2 + 1;
  const int PB = a.pb;
  // the group of this lane: slot blockIdx.y 512 + tid
  const int slot = blockIdx.y * kRysThreads + tid;
  const int g = slot < a.nslots ? a.groups[slot] : -1;
  // shape of this wavefront's groups (the same for every lane of a wavefront; a wavefront
  // of empty slots has none)
  const int gw = __builtin_amdgcn_readfirstlane(g);
  const int alen = gw < 0 ? 0 : (gw >> 18) & 15;
  const int clen = gw < 0 ? 0 : (gw >> 26) & 15;
  const int shape = alen > 0 && clen > 0 ? (alen - 1) * MC + (clen - 1) : -1;
  // an empty slot (g < 0) in a wavefront of groups reads the first entries of the tables
  // and writes nothing
  const int ox = g < 0 ? 0 : g & 127;
  const int oy = g < 0 ? 0 : (g >> 7) & 127;
  const int oz0 = g < 0 ? 0 : ((g >> 14) & 15) * W + ((g >> 22) & 15);
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
    double acc[MA][MC];
#pragma unroll
    for (int i = 0; i < MA; ++i)
#pragma unroll
      for (int j = 0; j < MC; ++j) acc[i][j] = 0.0;

    for (int p0 = pbeg; p0 < pend; p0 += PB) {
      const int ncopy = min(PB, pend - p0);
      __syncthreads();  // the previous batch's tables are no longer read
      if (tid < PB * NR * 3) rys_build(t, a, pb, pk, p0, ncopy, NR, rtab);
      __syncthreads();
      if (shape >= 0)
        rys_group_dispatch<MA, MC>(shape, rtab, ncopy * NR, TS, W, ox, oy, oz0, acc);
    }
    if (g >= 0) {
      const int32_t* rows = a.groups + a.nslots + static_cast<int64_t>(slot) * MA * MC;
#pragma unroll
      for (int i = 0; i < MA; ++i)
#pragma unroll
        for (int j = 0; j < MC; ++j) {
          const int f = (i < alen && j < clen) ? rows[i * MC + j] : -1;
          if (f >= 0) a.acc_out[rys_acc_index(a, qi, part, f)] = acc[i][j];
        }
    }
  }
}

template <int MA, int MC>
int launch_rys_mc(const gbx_tables* t, const RysArgs& a, dim3 grid, size_t shmem,
                  hipStream_t s) {
  hipLaunchKernelGGL((eri_rys_grouped_kernel<MA, MC>), grid, dim3(kRysThreads), shmem, s, *t, a);
  return check_launch("eri_rys_grouped_kernel");
}

template <int MA>
int launch_rys_ma(int mc, const gbx_tables* t, const RysArgs& a, dim3 grid, size_t shmem,
                  hipStream_t s) {
  switch (mc) {
    case 1: return launch_rys_mc<MA, 1>(t, a, grid, shmem, s);
    case 2: return launch_rys_mc<MA, 2>(t, a, grid, shmem, s);
    case 3: return launch_rys_mc<MA, 3>(t, a, grid, shmem, s);
    case 4: return launch_rys_mc<MA, 4>(t, a, grid, shmem, s);
    default: return launch_rys_mc<MA, 5>(t, a, grid, shmem, s);
  }
}

// Strided layout: lane tid holds accumulators f = acc_lo + tid + k 512, k < KMAX, each
// read as three table entries per node.
template <int NR, int KMAX>
__global__ __launch_bounds__(kRysThreads, kRysWaves) void eri_rys_strided_kernel(gbx_tables t,
                                                                                RysArgs a) {
  extern __shared__ __align__(16) double rtab[];
  const int tid = threadIdx.x;
  const int W = a.lcd + 1;
  const int TS = (a.lab + 1) * W;
  const int PB = a.pb;
  int idx[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int f = a.acc_lo + tid + k * kRysThreads;
    idx[k] = (k < a.kacc && f < a.nslots) ? a.groups[f] : -1;
  }
  const int4* pinfo = reinterpret_cast<const int4*>(t.pair_info);
  const int64_t nquart = a.count ? min(a.nquart, static_cast<int64_t>(*a.count)) : a.nquart;
  for (int64_t wi = blockIdx.x; wi < nquart * a.parts; wi += gridDim.x) {
    const int64_t qi = wi % nquart;
    const int part = static_cast<int>(wi / nquart);
    const int bra = a.quart[2 * qi], ket = a.quart[2 * qi + 1];
    const int4 pb = pinfo[bra];
    const int4 pk = pinfo[ket];
    const int nprimq = pb.w * pk.w;
    const int chunk = (nprimq + a.parts - 1) / a.parts;
    const int pbeg = min(nprimq, part * chunk);
    const int pend = min(nprimq, pbeg + chunk);
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) acc[k] = 0.0;
    for (int p0 = pbeg; p0 < pend; p0 += PB) {
      const int ncopy = min(PB, pend - p0);
      __syncthreads();  // the previous batch's tables are no longer read
      if (tid < PB * NR * 3) rys_build(t, a, pb, pk, p0, ncopy, NR, rtab);
      __syncthreads();
      // the node stride, opaque to the compiler in each batch: otherwise it hoists the
      // (accumulator, node) addresses out of the batch loop and spills them
      int stride = 3 * TS;
      __asm__ volatile("" : "+s"(stride));
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        if (idx[k] < 0) continue;
        // decoded again in each batch (three registers fewer per accumulator than
        // keeping the addresses across batches)
        __asm__ volatile("" : "+v"(idx[k]));
        const double* X = rtab + (idx[k] & 1023);
        const double* Y = rtab + TS + ((idx[k] >> 10) & 1023);
        const double* Z = rtab + 2 * TS + (idx[k] >> 20);
        double s = acc[k];
        for (int p = 0; p < ncopy; ++p) {
          // the nodes of one primitive quartet unrolled: NR independent reads in flight
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const int o = (p * NR + r) * stride;
            s = fma(X[o] * Y[o], Z[o], s);
          }
        }
        acc[k] = s;
      }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      if (idx[k] < 0) continue;
      const int f = a.acc_lo + tid + k * kRysThreads;
      a.acc_out[rys_acc_index(a, qi, part, f)] = acc[k];
    }
  }
}

template <int NR>
int launch_rys_strided_nr(const gbx_tables* t, const RysArgs& a, dim3 grid, size_t shmem,
                          hipStream_t s) {
  if (a.kacc <= 8)
    hipLaunchKernelGGL((eri_rys_strided_kernel<NR, 8>), grid, dim3(kRysThreads), shmem, s, *t, a);
  else if (a.kacc <= 16)
    hipLaunchKernelGGL((eri_rys_strided_kernel<NR, 16>), grid, dim3(kRysThreads), shmem, s, *t, a);
  else
    hipLaunchKernelGGL((eri_rys_strided_kernel<NR, 24>), grid, dim3(kRysThreads), shmem, s, *t, a);
  return check_launch("eri_rys_strided_kernel");
}

int launch_rys_strided(int nr, const gbx_tables* t, const RysArgs& a, dim3 grid, size_t shmem,
                       hipStream_t s) {
  switch (nr) {
    case 4: return launch_rys_strided_nr<4>(t, a, grid, shmem, s);
    case 5: return launch_rys_strided_nr<5>(t, a, grid, shmem, s);
    case 6: return launch_rys_strided_nr<6>(t, a, grid, shmem, s);
    case 7: return launch_rys_strided_nr<7>(t, a, grid, shmem, s);
    case 8: return launch_rys_strided_nr<8>(t, a, grid, shmem, s);
    default: return launch_rys_strided_nr<9>(t, a, grid, shmem, s);
  }
}

}  // namespace gbx

using namespace gbx;

extern "C" int64_t gbx_rys_table_size(void) { return kRysTableSize; }

extern "C" int gbx_eri_rys_acc(const gbx_tables* t, const double* d_rys, int32_t layout,
                               int32_t batch_max, const int32_t* d_index, int32_t n_index,
                               int32_t la, int32_t lb, int32_t lc, int32_t ld,
                               const int32_t* d_quartets, int64_t nquart, const int32_t* d_count,
                               int32_t parts, double* d_acc, int64_t acc_stride,
                               int32_t acc_layout, void* stream) {
  if (!t || (nquart > 0 && (!d_quartets || !d_acc || !d_rys || !d_index)))
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: null argument");
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la + lb >= kRysMaxE || lb > kRysMaxLb ||
      lc + ld >= kRysMaxE || ld > kRysMaxLb)
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: angular momentum out of range");
  if (t->potential != GBX_POT_COULOMB)
    return fail(GBX_ERR_UNSUPPORTED, "gbx_eri_rys_acc: the Rys rule integrates 1/r only");
  const int L = la + lb + lc + ld;
  const int nr = L / 2 + 1;
  if (nr < 4 || nr > kRysNmax)
    return fail(GBX_ERR_UNSUPPORTED, "gbx_eri_rys_acc: classes with 6 <= L <= 17 only");
  if (parts < 1 || parts > 4096)
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: parts must be in [1, 4096]");
  if (batch_max < 0 || batch_max > kRysPBmax)
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: batch_max must be in [0, 32]");
  if (acc_layout != 0 && acc_layout != 1)
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: acc_layout must be 0 or 1");
  if ((la + lb + 1) * (lc + ld + 1) > 128)
    return fail(GBX_ERR_UNSUPPORTED, "gbx_eri_rys_acc: tables too large for the index encoding");
  auto ncomp = [](int lo, int hi) {
    int n = 0;
    for (int l = lo; l <= hi; ++l) n += (l + 1) * (l + 2) / 2;
    return n;
  };
  const int nacc = ncomp(la, la + lb) * ncomp(lc, lc + ld);
  if (acc_layout == 0 ? acc_stride < nquart : acc_stride < nacc)
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: acc_stride below the quartets (layout 0) or "
                             "the accumulators (layout 1)");
  if (layout == kRysGrouped) {
    if (n_index < 1 || n_index % kRysThreads != 0 || n_index > 64 * kRysThreads)
      return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: grouped slots must be a multiple of 512");
  } else if (layout == kRysStrided) {
    if (n_index != nacc)
      return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: strided index must hold one entry per accumulator");
  } else {
    return fail(GBX_ERR_ARG, "gbx_eri_rys_acc: unknown layout");
  }
  if (nquart == 0) return GBX_OK;
  RysArgs a{};
  a.quart = d_quartets;
  a.nquart = nquart;
  a.count = d_count;
  a.rys = d_rys;
  a.groups = d_index;
  a.nslots = n_index;
  a.lab = la + lb;
  a.lcd = lc + ld;
  a.parts = parts;
  a.acc_out = d_acc;
  a.acc_stride = acc_stride;
  a.qmajor = acc_layout;
  const int64_t per_prim = int64_t(sizeof(double)) * nr * 3 * (a.lab + 1) * (a.lcd + 1);
  a.pb = static_cast<int>(std::min<int64_t>({batch_max > 0 ? batch_max : kRysPBmax,
                                             65536 / per_prim, kRysThreads / (3 * nr)}));
  if (a.pb < 1) return fail(GBX_ERR_UNSUPPORTED, "gbx_eri_rys_acc: node tables exceed 64 KiB");
  const size_t shmem = static_cast<size_t>(per_prim * a.pb);
  const unsigned gx =
      static_cast<unsigned>(std::min<int64_t>(nquart * parts, int64_t(1) << 20));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (layout == kRysGrouped) {
    const dim3 grid(gx, static_cast<unsigned>(n_index / kRysThreads));
    switch (lb + 1) {
      case 1: return launch_rys_ma<1>(ld + 1, t, a, grid, shmem, s);
      case 2: return launch_rys_ma<2>(ld + 1, t, a, grid, shmem, s);
      case 3: return launch_rys_ma<3>(ld + 1, t, a, grid, shmem, s);
      case 4: return launch_rys_ma<4>(ld + 1, t, a, grid, shmem, s);
      default: return launch_rys_ma<5>(ld + 1, t, a, grid, shmem, s);
    }
  }
  // strided: at most kRysKmax accumulators per lane and launch; classes with more run as
  // several launches over consecutive accumulator ranges
  const int kall = (nacc + kRysThreads - 1) / kRysThreads;
  const int nsplit = (kall + kRysKmax - 1) / kRysKmax;
  a.kacc = (kall + nsplit - 1) / nsplit;
  for (int sp = 0; sp < nsplit; ++sp) {
    a.acc_lo = sp * a.kacc * kRysThreads;
    const int rc = launch_rys_strided(nr, t, a, dim3(gx), shmem, s);
    if (rc != GBX_OK) return rc;
  }
  return GBX_OK;
}
