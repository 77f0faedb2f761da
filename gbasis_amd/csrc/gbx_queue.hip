// Device-side construction of the screened, class-bucketed shell-quartet work
// queue for a batch of molecules.
//
// Candidate quartets are the unique pairs-of-pairs (p >= q) of every molecule,
// the enumeration of gbasis/base_four_symm.py:189-194.  A candidate is kept
// iff Q_p * Q_q >= threshold (gbasis/integrals/_schwarz_screening.py:184-191;
// threshold <= 0 keeps everything), oriented so the "heavier" pair is the
// bra, and marked FP32-bulk iff P_p * P_q < fp32_cutoff (P: bound of the
// contracted [e0|f0], the input of the horizontal recursion) and fp32_ok allows
// the class, checked again for each sub-quartet's own class.  A kept quartet
// becomes up to four sub-quartets, one per phase, from the pairs' sub-pair
// entries pair_sub[p] = (first-centre part A_p, second-centre part B_p or -1):
//   phase 0 (A_p|A_q), always; phase 1 (B_p|A_q); phase 2 (A_p|B_q); phase 3 (B_p|B_q)
// (the primitive pairs of a pair split by the centre they are built on; the
// quartet's integrals are the sum of its sub-quartets').  Each goes to a bucket
//   key = ((((phase*4096 + cls)*2 + dense)*2 + fp32)*16 + bin,
//   cls = ((la*8+lb)*8+lc)*8+ld of the sub-quartet's own entries,
//   bin = min(ceil(log2(nprim_bra*nprim_ket)), 15)
// (the bin only groups sub-quartets of similar cost; kernels read the true
// primitive counts from the pair table, so the last bin holds every count above
// 2^14).  Pass 1 counts sub-quartets per bucket (optionally also packed slots,
// counted once per quartet in phase 0, and primitive quartets per bucket); pass 2
// writes (bra, ket) entries at bucket_offset[key] + an atomically reserved
// position.  Both passes aggregate atomics per wavefront and key.  The order of
// sub-quartets inside a bucket is not deterministic; the integral of every packed
// slot is, since a slot gets exactly one sub-quartet per phase, with arithmetic
// that does not depend on its neighbours, and the phases run in order.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gbasis_amd.h"
#include "gbx_common.h"

namespace gbx {

constexpr int kQueueThreads = 256;
constexpr int kQueueSlots = 64;  // LDS keys per workgroup and phase (a power of two)

struct QueueArgs {
  int32_t n_mol;
  const int64_t* mol_cand_off;  // n_mol+1 prefix sums of P_m (P_m + 1) / 2
  const int32_t* mol_pair_off;  // n_mol first global pair of each molecule
  const int32_t* pair_key;      // per pair entry: la | lb << 3 | dense << 6 | nprim << 8
  const int32_t* pair_sub;      // per pair: (first-centre entry, second-centre entry | -1)
  const int32_t* pair_npf;      // per pair: number of unique function pairs
  const double* pair_bound;     // per pair: Schwarz bound Q of its original shell pair
  const double* pair_pbound;    // per pair: bound of the contracted [e0|f0] (precision choice)
  double threshold, fp32_cutoff;
  const double* mol_fp32_cutoff;  // per molecule (overrides fp32_cutoff); null: fp32_cutoff
  const int8_t* fp32_ok;        // per (class * 2 + dense): FP32 bulk allowed; null = all
  int32_t* bucket_count;        // [GBX_NUM_BUCKETS]
  unsigned long long* bucket_eris;    // [GBX_NUM_BUCKETS] or null
  unsigned long long* bucket_primq;   // [GBX_NUM_BUCKETS] or null
  const int64_t* bucket_offset;       // pass 2: first slot of each bucket in out, and
                                      //   the end of the last ([GBX_NUM_BUCKETS + 1])
  int32_t* bucket_cursor;             // pass 2: [GBX_NUM_BUCKETS], zeroed
  int32_t* out;                       // pass 2: (bra, ket) pairs
  int64_t n_cand;
};

// primitive-count bin of the bucket key: ceil(log2(x)), at most kMaxBin
constexpr int kMaxBin = 15;
__device__ inline int prim_bin(int x) {
  int b = 0;
  while (b < kMaxBin && (1 << b) < x) ++b;
  return b;
}

// The slot of key in an LDS table of kQueueSlots keys (-1 = free), claimed by open
// addressing; -1 when the table is full.
__device__ inline int queue_slot(int* s_key, int key) {
  int h = key & (kQueueSlots - 1);
  for (int probe = 0; probe < kQueueSlots; ++probe) {
    const int prev = atomicCAS(&s_key[h], -1, key);
    if (prev == -1 || prev == key) return h;
    h = (h + 1) & (kQueueSlots - 1);
  }
  return -1;
}

// candidate t of a molecule with P pairs -> (p >= q), t = p (p + 1) / 2 + q
__device__ inline void tri_decode(int64_t t, int64_t& p, int64_t& q) {
  int64_t r = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(t) + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > t) --r;
  while ((r + 1) * (r + 2) / 2 <= t) ++r;
  p = r;
  q = t - r * (r + 1) / 2;
}

template <bool kFill>
__global__ __launch_bounds__(kQueueThreads) void quartet_queue_kernel(QueueArgs a) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kQueueThreads + threadIdx.x;
  bool valid = t < a.n_cand;
  int bra = 0, ket = 0, dense = 0, fp32 = 0;
  int sb[2] = {-1, -1}, sk[2] = {-1, -1};
  uint64_t eris = 0;
  if (valid) {
    // molecule by binary search over the candidate prefix sums
    int lo = 0, hi = a.n_mol - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.mol_cand_off[mid] <= t) lo = mid; else hi = mid - 1;
    }
    int64_t p, q;
    tri_decode(t - a.mol_cand_off[lo], p, q);
    const int gp = a.mol_pair_off[lo] + static_cast<int>(p);
    const int gq = a.mol_pair_off[lo] + static_cast<int>(q);
    const double bound = a.pair_bound[gp] * a.pair_bound[gq];
    if (a.threshold > 0.0 && !(bound >= a.threshold)) {
      valid = false;
    } else {
      const int kp = a.pair_key[gp], kq = a.pair_key[gq];
      const int lap = kp & 7, lbp = (kp >> 3) & 7, laq = kq & 7, lbq = (kq >> 3) & 7;
      const bool swap = ((lap + lbp) * 16 + lap) < ((laq + lbq) * 16 + laq);
      bra = swap ? gq : gp;
      ket = swap ? gp : gq;
      const int kb = swap ? kq : kp, kk = swap ? kp : kq;
      dense = ((kb >> 6) & 1) | ((kk >> 6) & 1);
      const double pbound = a.pair_pbound[gp] * a.pair_pbound[gq];
      const int cls = (((kb & 7) * 8 + ((kb >> 3) & 7)) * 8 + (kk & 7)) * 8 + ((kk >> 3) & 7);
      const double cut = a.mol_fp32_cutoff ? a.mol_fp32_cutoff[lo] : a.fp32_cutoff;
      fp32 = (cut > 0.0 && pbound < cut &&
              (a.fp32_ok == nullptr || a.fp32_ok[cls * 2 + dense])) ? 1 : 0;
      sb[0] = a.pair_sub[2 * bra];
      sb[1] = a.pair_sub[2 * bra + 1];
      sk[0] = a.pair_sub[2 * ket];
      sk[1] = a.pair_sub[2 * ket + 1];
      if (!kFill && a.bucket_eris) {
        const uint64_t nb = static_cast<uint64_t>(a.pair_npf[bra]);
        const uint64_t nk = static_cast<uint64_t>(a.pair_npf[ket]);
        eris = bra == ket ? nb * (nb + 1) / 2 : nb * nk;
      }
    }
  }
  const int lane = threadIdx.x & 63;
  // Counts and positions go to the buckets per workgroup and key, not per wavefront and
  // key: the wavefronts add into an LDS table of up to kQueueSlots keys, then one lane
  // per occupied slot makes the workgroup's device atomics (count pass: the count, ERIs
  // and primitive quartets; fill pass: the reservation of the workgroup's block).  A
  // key that finds no slot goes to the device atomics per wavefront.
  __shared__ int s_key[kQueueSlots];
  __shared__ int s_cnt[kQueueSlots];
  __shared__ int s_base[kQueueSlots];
  __shared__ unsigned long long s_eris[kQueueSlots];
  __shared__ unsigned long long s_pq[kQueueSlots];
  for (int phase = 0; phase < 4; ++phase) {
    const int eb = sb[phase & 1], ek = sk[phase >> 1];
    const bool emit = valid && eb >= 0 && ek >= 0;
    int key = -1;
    uint64_t primq = 0;
    if (emit) {
      const int kb = a.pair_key[eb], kk = a.pair_key[ek];
      const int cls = (((kb & 7) * 8 + ((kb >> 3) & 7)) * 8 + (kk & 7)) * 8 + ((kk >> 3) & 7);
      const int np = (kb >> 8) * (kk >> 8);
      // a sub-quartet of a later phase has its own class (an entry built on the second
      // centre swaps its shells): it runs FP32 only if that class may too, so it never
      // lands in an FP32 bucket the host gave no room
      const int fp32s = (fp32 && (a.fp32_ok == nullptr || a.fp32_ok[cls * 2 + dense])) ? 1 : 0;
      key = (((phase * 4096 + cls) * 2 + dense) * 2 + fp32s) * 16 + prim_bin(np);
      primq = static_cast<uint64_t>(np);
    }
    if (threadIdx.x < kQueueSlots) {
      s_key[threadIdx.x] = -1;
      s_cnt[threadIdx.x] = 0;
      s_eris[threadIdx.x] = 0ull;
      s_pq[threadIdx.x] = 0ull;
    }
    __syncthreads();
    int my_slot = -1, my_off = 0;  // fill: this lane's slot and offset in its block
    // wavefront-aggregated reservation per distinct key
    uint64_t pending = __ballot(emit);
    while (pending) {
      const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
      const int lkey = __shfl(key, leader, 64);
      const bool mine = emit && key == lkey;
      const uint64_t same = __ballot(mine);
      const int n = __popcll(same);
      const int rank = __popcll(same & ((uint64_t(1) << lane) - 1));
      if (kFill) {
        int slot = -1, off = 0;
        if (lane == leader) {
          slot = queue_slot(s_key, lkey);
          if (slot >= 0)
            off = atomicAdd(&s_cnt[slot], n);
          else
            off = atomicAdd(a.bucket_cursor + lkey, n);  // the table is full
        }
        slot = __shfl(slot, leader, 64);
        off = __shfl(off, leader, 64);
        if (mine) {
          my_slot = slot;
          my_off = off + rank;
        }
      } else {
        unsigned long long e = 0ull, pq = 0ull;
        if (a.bucket_eris) {
          // sums over the lanes of this key; the packed slots of a quartet count once,
          // in phase 0
          e = (mine && phase == 0) ? static_cast<unsigned long long>(eris) : 0ull;
          pq = mine ? static_cast<unsigned long long>(primq) : 0ull;
          for (int off = 32; off > 0; off >>= 1) {
            e += __shfl_xor(e, off, 64);
            pq += __shfl_xor(pq, off, 64);
          }
        }
        if (lane == leader) {
          const int slot = queue_slot(s_key, lkey);
          if (slot >= 0) {
            atomicAdd(&s_cnt[slot], n);
            if (a.bucket_eris) {
              atomicAdd(&s_eris[slot], e);
              atomicAdd(&s_pq[slot], pq);
            }
          } else {  // the table is full
            atomicAdd(a.bucket_count + lkey, n);
            if (a.bucket_eris) {
              atomicAdd(a.bucket_eris + lkey, e);
              atomicAdd(a.bucket_primq + lkey, pq);
            }
          }
        }
      }
      pending &= ~same;
    }
    if (kFill) {
      __syncthreads();  // the table holds the workgroup's counts
      if (threadIdx.x < kQueueSlots && s_key[threadIdx.x] >= 0)
        s_base[threadIdx.x] = atomicAdd(a.bucket_cursor + s_key[threadIdx.x],
                                        s_cnt[threadIdx.x]);
      __syncthreads();
      if (emit) {
        // a bucket sized too small never writes into its neighbour: its cursor then
        // exceeds its capacity, which the host can detect
        const int64_t pos =
            a.bucket_offset[key] + (my_slot >= 0 ? s_base[my_slot] : 0) + my_off;
        if (pos < a.bucket_offset[key + 1]) {
          a.out[2 * pos] = eb;
          a.out[2 * pos + 1] = ek;
        }
      }
      __syncthreads();  // s_base is read before the next phase clears the table
    } else {
      __syncthreads();  // the table holds the workgroup's sums
      if (threadIdx.x < kQueueSlots && s_key[threadIdx.x] >= 0) {
        const int k = s_key[threadIdx.x];
        atomicAdd(a.bucket_count + k, s_cnt[threadIdx.x]);
        if (a.bucket_eris) {
          atomicAdd(a.bucket_eris + k, s_eris[threadIdx.x]);
          atomicAdd(a.bucket_primq + k, s_pq[threadIdx.x]);
        }
      }
      __syncthreads();  // the table is read before the next phase clears it
    }
  }
}

int launch_queue(const QueueArgs& a, bool fill, hipStream_t s) {
  if (a.n_cand == 0) return GBX_OK;
  const int64_t blocks = (a.n_cand + kQueueThreads - 1) / kQueueThreads;
  if (blocks > 0x7fffffff) return fail(GBX_ERR_ARG, "gbx_quartets: too many candidates");
  if (fill)
    hipLaunchKernelGGL(quartet_queue_kernel<true>, dim3(static_cast<unsigned>(blocks)),
                       dim3(kQueueThreads), 0, s, a);
  else
    hipLaunchKernelGGL(quartet_queue_kernel<false>, dim3(static_cast<unsigned>(blocks)),
                       dim3(kQueueThreads), 0, s, a);
  return check_launch(fill ? "quartet_queue_kernel<fill>" : "quartet_queue_kernel<count>");
}

}  // namespace gbx

extern "C" int gbx_quartets_count(int32_t n_mol, const int64_t* d_mol_cand_off,
                                  const int32_t* d_mol_pair_off, const int32_t* d_pair_key,
                                  const int32_t* d_pair_sub, const int32_t* d_pair_npf,
                                  const double* d_pair_bound, const double* d_pair_pbound,
                                  double threshold,
                                  double fp32_cutoff, const double* d_mol_fp32_cutoff,
                                  const int8_t* d_fp32_ok, int64_t n_cand,
                                  int32_t* d_bucket_count, uint64_t* d_bucket_eris,
                                  uint64_t* d_bucket_primq, void* stream) {
  if (n_mol < 1 || !d_mol_cand_off || !d_mol_pair_off || !d_pair_key || !d_pair_sub ||
      !d_pair_bound || !d_pair_pbound || !d_bucket_count || n_cand < 0 ||
      (d_bucket_eris && (!d_bucket_primq || !d_pair_npf)))
    return gbx::fail(GBX_ERR_ARG, "gbx_quartets_count: bad arguments");
  gbx::QueueArgs a{};
  a.n_mol = n_mol;
  a.mol_cand_off = d_mol_cand_off;
  a.mol_pair_off = d_mol_pair_off;
  a.pair_key = d_pair_key;
  a.pair_sub = d_pair_sub;
  a.pair_npf = d_pair_npf;
  a.pair_bound = d_pair_bound;
  a.pair_pbound = d_pair_pbound;
  a.threshold = threshold;
  a.fp32_cutoff = fp32_cutoff;
  a.mol_fp32_cutoff = d_mol_fp32_cutoff;
  a.fp32_ok = d_fp32_ok;
  a.bucket_count = d_bucket_count;
  a.bucket_eris = reinterpret_cast<unsigned long long*>(d_bucket_eris);
  a.bucket_primq = reinterpret_cast<unsigned long long*>(d_bucket_primq);
  a.n_cand = n_cand;
  return gbx::launch_queue(a, false, static_cast<hipStream_t>(stream));
}

extern "C" int gbx_quartets_fill(int32_t n_mol, const int64_t* d_mol_cand_off,
                                 const int32_t* d_mol_pair_off, const int32_t* d_pair_key,
                                 const int32_t* d_pair_sub, const double* d_pair_bound,
                                 const double* d_pair_pbound, double threshold,
                                 double fp32_cutoff,
                                 const double* d_mol_fp32_cutoff, const int8_t* d_fp32_ok,
                                 int64_t n_cand, const int64_t* d_bucket_offset,
                                 int32_t* d_bucket_cursor, int32_t* d_quartets, void* stream) {
  if (n_mol < 1 || !d_mol_cand_off || !d_mol_pair_off || !d_pair_key || !d_pair_sub ||
      !d_pair_bound || !d_pair_pbound || !d_bucket_offset || !d_bucket_cursor || !d_quartets ||
      n_cand < 0)
    return gbx::fail(GBX_ERR_ARG, "gbx_quartets_fill: bad arguments");
  gbx::QueueArgs a{};
  a.n_mol = n_mol;
  a.mol_cand_off = d_mol_cand_off;
  a.mol_pair_off = d_mol_pair_off;
  a.pair_key = d_pair_key;
  a.pair_sub = d_pair_sub;
  a.pair_bound = d_pair_bound;
  a.pair_pbound = d_pair_pbound;
  a.threshold = threshold;
  a.fp32_cutoff = fp32_cutoff;
  a.mol_fp32_cutoff = d_mol_fp32_cutoff;
  a.fp32_ok = d_fp32_ok;
  a.bucket_offset = d_bucket_offset;
  a.bucket_cursor = d_bucket_cursor;
  a.out = d_quartets;
  a.n_cand = n_cand;
  return gbx::launch_queue(a, true, static_cast<hipStream_t>(stream));
}
