/*
 * gbasis_amd — MI355X (gfx950) electron-repulsion-integral library, C ABI.
 *
 * Every entry point takes device pointers, plain sizes and a HIP stream passed
 * as void* (NULL = the null stream), returns 0 on success and a negative code on
 * failure, and never throws.  gbx_last_error() returns the message of the last
 * failure on the calling thread.  All buffers are owned by the caller.
 *
 * The Python host layer (gbasis_amd/electron_repulsion.py) builds the tables
 * described below with NumPy, places them in device memory and binds these
 * symbols with ctypes, in the style of the reference's own ctypes binding
 * (gbasis/integrals/libcint.py:275-304).
 */
#ifndef GBASIS_AMD_H
#define GBASIS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBX_OK 0
#define GBX_ERR_ARG -1
#define GBX_ERR_HIP -2
#define GBX_ERR_UNSUPPORTED -3

/* precision of the recursion bulk (Boys, prefactors, accumulation and the
 * horizontal recursion stay FP64) */
#define GBX_FP64 0
#define GBX_FP32_BULK 1

/* output mode of the ERI kernels */
#define GBX_MODE_PACKED 0     /* write canonical 8-fold packed slots        */
#define GBX_MODE_SCHWARZ 1    /* per bra pair p: d_out[2p] = sqrt(max|(ij|ij)|),
                                 d_out[2p+1] = sqrt(max|contracted [e0|f0]|)  */
#define GBX_MODE_PACKED_ADD 2 /* add into the canonical 8-fold packed slots
                                 (sub-quartets of phases 1-3, see the work queue) */

/* two-electron operator (gbx_tables.potential); the range-separated ones use
 * gbx_tables.omega, as gbasis/integrals/boys_functions.py:145-223 */
#define GBX_POT_COULOMB 0 /* 1/r                                              */
#define GBX_POT_ERF 1     /* erf(omega r)/r: G_m = s^(m+1/2) F_m(s T),
                             s = omega^2 / (omega^2 + rho)                  */
#define GBX_POT_ERFC 2    /* erfc(omega r)/r: F_m(T) - G_m                   */

/* Device-resident basis tables of one batch of molecules. */
typedef struct gbx_tables {
  const int32_t *shell_info;     /* 8 x int32 per shell: func_off, l, kind, map_off,
                                    mat_off, nout, 0, 0 (kind 0 = scaled permutation,
                                    kind 1 = dense Cartesian->pure matrix)            */
  const int32_t *comp_map;       /* per Cartesian component: position in its shell  */
  const double *comp_scale;      /* per Cartesian component: output scale           */
  const double *out_mat;         /* dense per-shell matrices, row-major nout x ncart */
  const int32_t *pair_info;      /* 4 x int32 per pair entry: shell_a, shell_b,
                                    prim_off, nprim.  Entries below the batch's pair
                                    count are the shell pairs (la >= lb) with all
                                    their primitive pairs; after them come the
                                    sub-pair entries of split pairs: the primitive
                                    pairs built on the first centre (a, b) and those
                                    built on the second (b, a, records with P-B in
                                    place of P-A)                                    */
  const double *pair_geom;       /* 4 x double per pair entry: A-B (x,y,z), 0        */
  const uint64_t *pair_out_base; /* per pair: first packed slot of its molecule      */
  const double *prim_pairs;      /* 10 x double per primitive pair: zeta, 1/(2 zeta),
                                    P(3), P-A(3) (A: the entry's first shell),
                                    K', 0; K' = c_a c_b N_a N_b
                                    exp(-ab/zeta |AB|^2) sqrt(2) pi^(5/4) / zeta     */
  const double *f2norm;          /* 1/sqrt(prod (2a_i-1)!!) by l, then component     */
  const double *boys_table;      /* boys_nt x boys_nord, row k: e^{-k h}, then
                                    F_n(k h) for n = 0 .. boys_nord - 2              */
  int32_t boys_nt;
  int32_t boys_nord;
  double boys_h;
  int32_t potential;             /* GBX_POT_*: operator of GBX_MODE_PACKED launches;
                                    GBX_MODE_SCHWARZ bounds are always Coulomb, as
                                    the reference's screener
                                    (gbasis/integrals/electron_repulsion.py:475).
                                    Range-separated operators run on the
                                    cooperative kernels in FP64 only; the
                                    register kernels reject them.                 */
  int32_t reserved;
  double omega;                  /* range-separation parameter (> 0)            */
} gbx_tables;

/* A class program (gbasis_amd/programs.py) resident on the device. */
typedef struct gbx_program {
  const int32_t *items;     /* 4 x int32 per item: recursion items pack 16-bit
                               slots (dst | s0 << 16, s1 | s2 << 16, s3 | s4 << 16,
                               code; 0xFFFF = no source); accumulation items
                               (acc, slot, 0, 0); HRR items (dst, src1, src2, dir) */
  const int32_t *phases;    /* 3 x int32 per phase: kind, first item, end item  */
  const int32_t *out_slots; /* per canonical output component: its H slot     */
  int32_t header[16];    /* prim_slots, nacc, h_slots, nout, n_prim_phases,
                            n_phases, L, n_items                              */
} gbx_program;

int gbx_version(void);
const char *gbx_last_error(void);

/* Boys function F_m(T), m = 0..mmax, for n values of T:
 * out[i*(mmax+1)+m].  Replaces boys_function_standard
 * (gbasis/integrals/boys_functions.py:26-63) evaluated on the device.
 * mmax + 8 must be below t->boys_nord (orders up to 24 with the shipped
 * table); a negative or NaN T gives NaN for every order. */
int gbx_boys(const gbx_tables *t, const double *d_T, int64_t n, int32_t mmax, double *d_out,
             void *stream);

/* Workspace of the cooperative kernel for one class program and prim_copies
 * (1..64) primitive quartets in flight (prim_only != 0: of the primitive pass
 * gbx_eri_coop_acc): *lds_bytes = bytes per workgroup; *global_bytes_per_block
 * = the same when it exceeds the LDS one workgroup may allocate on the current
 * device (160 KiB on gfx950; the kernel then works in a slice of d_scratch per
 * workgroup), else 0. */
int gbx_coop_workspace(const gbx_program *prog, int32_t la, int32_t lb, int32_t lc, int32_t ld,
                       int32_t precision, int32_t prim_copies, int32_t prim_only,
                       int64_t *lds_bytes, int64_t *global_bytes_per_block);

/* One angular-momentum class of shell quartets with the cooperative
 * (one workgroup of 256-1024 lanes per shell quartet) kernel.  d_quartets holds
 * nquart (bra, ket) int32 pair entries whose angular momenta are (la, lb) and
 * (lc, ld) in the entries' own order (a sub-pair built on its second centre has
 * la < lb).  Replaces, for a batch of quartets,
 * ElectronRepulsionIntegralImproved.construct_array_contraction
 * (gbasis/integrals/electron_repulsion.py:305-412) ->
 * compute_two_electron_integrals_os_hgp (gbasis/integrals/_two_elec_int_improved.py:591-732)
 * plus the norm_cont / Cartesian->pure / 8-fold placement of
 * gbasis/base_four_symm.py:160-389.  In GBX_MODE_SCHWARZ (bra == ket pair) it
 * writes d_out[2 bra] = sqrt(max|(ij|ij)|) without norm_cont, as
 * compute_schwarz_bound_shell_pair (gbasis/integrals/_schwarz_screening.py:14-59),
 * and d_out[2 bra + 1] = sqrt of the largest contracted input of the horizontal
 * recursion, the bound that selects the FP32 recursion bulk.  GBX_MODE_PACKED_ADD
 * adds the block to the packed slots instead of storing it.  Packed placement
 * writes each unique slot once: a quartet whose bra and ket entries cover the same
 * two shells writes only its ij >= kl half.
 * d_scratch (scratch_bytes) is only used when the class does not fit in LDS.
 * d_count: NULL, or a device int32 holding the number of quartets to evaluate
 * (at most nquart, which is then only the capacity the launch is sized for), so
 * a launch can follow the device queue builder with no host round trip; the
 * same parameter of gbx_eri_reg, gbx_eri_split and gbx_eri_hrr means the same. */
int gbx_eri_coop(const gbx_tables *t, const gbx_program *prog, int32_t la, int32_t lb,
                 int32_t lc, int32_t ld, const int32_t *d_quartets, int64_t nquart,
                 const int32_t *d_count, int32_t precision, int32_t prim_copies, int32_t mode,
                 double *d_out, void *d_scratch, int64_t scratch_bytes, void *stream);

/* gbx_eri_coop without the scratch arguments: allocates nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Classes whose
 * workgroup workspace exceeds the LDS work in a file-scope device array of 128 MiB
 * (one copy per GPU, held until the process ends; its device address is looked up
 * once per device), with as many workgroups as slices of it fit; a workspace larger
 * than the array is rejected with GBX_ERR_ARG.  It calls no hipMalloc / hipFree.  The
 * array needs no clearing (every workgroup initialises the parts of its slice it
 * reads), but all calls on one GPU share it: calls that overlap on one GPU (other
 * streams, other host threads) race and must be serialised by the caller. */
int gbx_eri_coop_noalloc(const gbx_tables *t, const gbx_program *prog, int32_t la, int32_t lb,
                         int32_t lc, int32_t ld, const int32_t *d_quartets, int64_t nquart,
                         const int32_t *d_count, int32_t precision, int32_t prim_copies,
                         int32_t mode, double *d_out, void *stream);

/* Primitive pass of the cooperative kernel, split over `parts` workgroups per
 * shell quartet: work item (q, r) sums the primitive quartets
 * [r c, (r + 1) c), c = ceil(nprimq / parts), of quartet q -- Boys, vertical
 * recursion, ket build, FP64 accumulation (gbasis/integrals/_two_elec_int_improved.py:651-716)
 * -- and writes the partial contracted [e0|f0] to d_acc[(f * parts + r) * acc_stride + q].
 * gbx_eri_hrr with acc_parts = parts finishes the quartets.  Spreads the
 * primitive work of classes with few quartets and many primitive quartets (f
 * and g shells, K = 8) over the GPU, and needs only the primitive part of the
 * workspace.  prog may be a primitive-pass program that produces a subset of the
 * class's accumulators (its accumulation items name their rows f); several such
 * programs together fill d_acc for classes whose primitive store exceeds the LDS. */
int gbx_eri_coop_acc(const gbx_tables *t, const gbx_program *prog, int32_t la, int32_t lb,
                     int32_t lc, int32_t ld, const int32_t *d_quartets, int64_t nquart,
                     const int32_t *d_count, int32_t precision, int32_t prim_copies,
                     int32_t parts, double *d_acc, int64_t acc_stride, void *d_scratch,
                     int64_t scratch_bytes, void *stream);

/* Primitive pass by Rys quadrature for classes with 6 <= L <= 17 (n = L/2 + 1 <= 9
 * nodes), l_a + l_b <= 8, l_c + l_d <= 8, l_b <= 4, l_d <= 4, Coulomb operator, FP64:
 * work item (q, r) sums the primitive quartets [r c, (r + 1) c) of quartet q as sum over
 * nodes of pref W_u Ix Iy Iz and writes the partial contracted [e0|f0] to
 * d_acc[(f * parts + r) * acc_stride + q] (acc_layout 0, acc_stride >= nquart), the
 * layout of gbx_eri_coop_acc, which gbx_eri_hrr with acc_parts = parts finishes, or
 * d_acc[(q * parts + r) * acc_stride + f] (acc_layout 1, acc_stride >= the number of
 * accumulators), which gbx_eri_hrr_lds reads.  Replaces the Boys function, vertical
 * recursion and transfer to the ket of gbasis/integrals/_two_elec_int_improved.py:87-361
 * with the same integrals (the rule integrates F_m(T), m < 2n, exactly).  d_rys is the
 * node table (gbx_rys_table_size doubles, built by gbasis_amd/rys.py).  Accumulator f is
 * the class program's [a0|c0] (a of degree l_a..l_a+l_b, c of l_c..l_c+l_d); with
 * ox = a_x (l_c + l_d + 1) + c_x and oy, oz likewise:
 *  layout 0 (strided): d_index holds n_index = (number of accumulators) entries
 *    ox | oy << 10 | oz << 20, one per f;
 *  layout 1 (grouped): d_index holds n_index (a multiple of 512, at most 32768) lane
 *    descriptors, each -1 or a group of the accumulators sharing (a_x, a_y, c_x, c_y):
 *    ox | oy << 7 | az_lo << 14 | alen << 18 | cz_lo << 22 | clen << 26 (members
 *    a_z = az_lo + i, i < alen, c_z = cz_lo + j, j < clen), every descriptor of a 64-lane
 *    wavefront of the same (alen, clen); then, per slot, (l_b + 1)(l_d + 1) rows f
 *    (member (i, j) at i (l_d + 1) + j), -1 where none.
 * batch_max (0: 32, else 1..32) caps the primitive quartets whose node tables a
 * workgroup builds at once in LDS (further capped by 64 KiB of tables and 512 build
 * threads); larger batches amortise the table build of the small classes. */
int gbx_eri_rys_acc(const gbx_tables *t, const double *d_rys, int32_t layout,
                    int32_t batch_max, const int32_t *d_index, int32_t n_index, int32_t la,
                    int32_t lb, int32_t lc, int32_t ld, const int32_t *d_quartets,
                    int64_t nquart, const int32_t *d_count, int32_t parts, double *d_acc,
                    int64_t acc_stride, int32_t acc_layout, void *stream);

/* Number of doubles of the Rys node table gbx_eri_rys_acc reads. */
int64_t gbx_rys_table_size(void);

/* 1 if a generated register kernel (one lane per shell quartet) exists for the
 * class (kernel orientation), else 0. */
int gbx_has_reg_kernel(int32_t la, int32_t lb, int32_t lc, int32_t ld);

/* One class of shell quartets with the generated register kernel: same
 * contract as gbx_eri_coop (the reference interfaces it replaces are listed
 * there), restricted to quartets whose four shells all have a scaled-
 * permutation output map (Cartesian shells, pure shells with l <= 1). */
int gbx_eri_reg(const gbx_tables *t, int32_t la, int32_t lb, int32_t lc, int32_t ld,
                const int32_t *d_quartets, int64_t nquart, const int32_t *d_count,
                int32_t precision, int32_t mode, double *d_out, void *stream);

/* Work-queue construction on the device for a batch of molecules: the unique
 * shell-quartet loop of gbasis/base_four_symm.py:189-194 with the Schwarz test
 * of SchwarzScreener.is_significant (gbasis/integrals/_schwarz_screening.py:169-191),
 * kept iff bound_p * bound_q >= threshold (threshold <= 0 keeps all), run with
 * the FP32 recursion bulk iff pbound_p * pbound_q < fp32_cutoff (pbound: the
 * second value GBX_MODE_SCHWARZ writes per pair) and d_fp32_ok (NULL, or one
 * int8 per class*2+dense, class = ((la*8+lb)*8+lc)*8+ld) allows it for the
 * quartet's class and for the class of the sub-quartet (below) itself;
 * d_mol_fp32_cutoff (NULL, or one cutoff per molecule) replaces fp32_cutoff
 * per molecule, so FP64 and mixed-precision requests can share one batch.  A kept
 * quartet (p|q) becomes the sub-quartets (A_p|A_q) (phase 0), (B_p|A_q) (1),
 * (A_p|B_q) (2), (B_p|B_q) (3) of its pairs' entries d_pair_sub[2p] = A_p (the
 * pair itself if unsplit), d_pair_sub[2p+1] = B_p or -1; their integrals sum to the
 * quartet's.  Buckets are keyed
 * ((((phase*4096+cls)*2+dense)*2+fp32)*16 + min(ceil(log2 nprimq), 15),
 * cls = ((la*8+lb)*8+lc)*8+ld of the sub-quartet's entries, GBX_NUM_BUCKETS of
 * them.  d_pair_key per pair entry: la | lb << 3 | dense << 6 | nprim << 8.
 * count: sub-quartets per bucket (and, if d_bucket_eris is not NULL, packed slots,
 * counted in phase 0, and primitive quartets per bucket).  fill: writes the
 * (bra, ket) entries at d_bucket_offset[key] + position (d_bucket_cursor zeroed;
 * d_bucket_offset has GBX_NUM_BUCKETS + 1 entries, the last the end of the queue, and
 * entries beyond a bucket's end, d_bucket_offset[key + 1], are not written: the
 * bucket's cursor then exceeds its capacity).
 * Launch phase 0 with GBX_MODE_PACKED and phases 1-3 after it, in order, with
 * GBX_MODE_PACKED_ADD. */
#define GBX_NUM_BUCKETS 1048576
int gbx_quartets_count(int32_t n_mol, const int64_t *d_mol_cand_off, const int32_t *d_mol_pair_off,
                       const int32_t *d_pair_key, const int32_t *d_pair_sub,
                       const int32_t *d_pair_npf,
                       const double *d_pair_bound, const double *d_pair_pbound,
                       double threshold, double fp32_cutoff, const double *d_mol_fp32_cutoff,
                       const int8_t *d_fp32_ok, int64_t n_cand, int32_t *d_bucket_count,
                       uint64_t *d_bucket_eris,
                       uint64_t *d_bucket_primq, void *stream);
int gbx_quartets_fill(int32_t n_mol, const int64_t *d_mol_cand_off, const int32_t *d_mol_pair_off,
                      const int32_t *d_pair_key, const int32_t *d_pair_sub,
                      const double *d_pair_bound,
                      const double *d_pair_pbound, double threshold, double fp32_cutoff,
                      const double *d_mol_fp32_cutoff, const int8_t *d_fp32_ok, int64_t n_cand,
                      const int64_t *d_bucket_offset,
                      int32_t *d_bucket_cursor, int32_t *d_quartets, void *stream);

/* Segmented form of gbx_eri_reg: one grid of max_quartets lanes covers up to 16
 * quartet segments of one class (the primitive-count bins of the device work
 * queue); segment s holds min(d_seg_cnt[s], d_seg_off[s + 1] - d_seg_off[s])
 * quartets starting at d_quartets[2 * d_seg_off[s]] (d_seg_off: 17 entries).  Counts are read on the device, so the launch
 * needs no host round trip (max_quartets: an upper bound of their sum). */
int gbx_eri_reg_segments(const gbx_tables *t, int32_t la, int32_t lb, int32_t lc, int32_t ld,
                         const int32_t *d_quartets, const int64_t *d_seg_off,
                         const int32_t *d_seg_cnt, int64_t max_quartets, int32_t precision,
                         int32_t mode, double *d_out, void *stream);

/* Number of split register kernels for a class too large for one lane's
 * registers (0 if the class has a fused kernel or none). */
int gbx_split_count(int32_t la, int32_t lb, int32_t lc, int32_t ld);

/* Split kernel `split` of a large class: the primitive loop (Boys, vertical
 * recursion, ket build, FP64 accumulation -- the part of
 * compute_two_electron_integrals_os_hgp, gbasis/integrals/_two_elec_int_improved.py:651-716,
 * before the horizontal recursion) for one subset of the contracted [e0|f0].
 * lanes_per_quartet (R) lanes share each quartet, lane r summing primitive
 * quartets r, r+R, ...; partial sums go to d_acc[(f * R + r) * acc_stride + q]
 * (f: accumulator index of the class program, q: quartet index). */
int gbx_eri_split(const gbx_tables *t, int32_t la, int32_t lb, int32_t lc, int32_t ld,
                  int32_t split, const int32_t *d_quartets, int64_t nquart,
                  const int32_t *d_count, int32_t precision, int32_t lanes_per_quartet,
                  double *d_acc, int64_t acc_stride, void *stream);

/* Horizontal recursion, norms, output transform and placement (or the Schwarz
 * bounds in GBX_MODE_SCHWARZ) from contracted [e0|f0] produced by the split
 * kernels (acc_parts partial sums per accumulator, summed here): the rest of
 * gbx_eri_coop (gbasis/integrals/_two_elec_int_improved.py:364-588,
 * gbasis/base_four_symm.py:160-389). */
int gbx_eri_hrr(const gbx_tables *t, const gbx_program *prog, int32_t la, int32_t lb, int32_t lc,
                int32_t ld, const int32_t *d_quartets, int64_t nquart, const int32_t *d_count,
                int32_t mode, const double *d_acc, int64_t acc_stride, int32_t acc_parts,
                double *d_out,
                void *d_scratch, int64_t scratch_bytes, void *stream);

/* The same pass as gbx_eri_hrr in GBX_MODE_PACKED / GBX_MODE_PACKED_ADD, for any class
 * whose row buffers and W fit the LDS, including those whose whole recursion workspace
 * exceeds it (gbx_eri_hrr runs them on global scratch): one workgroup per quartet (256
 * lanes when the LDS arrays take at most 32 KiB, 512 up to 64 KiB, else 1024) runs the
 * ket recursion, the norms of c and d and the transforms of shells C and D row by row
 * (bra component e) into an LDS array W[e][c'd'], then the bra recursion, the norms and transforms of A and B and the
 * placement column by column (gbasis/integrals/_two_elec_int_improved.py:364-588,
 * gbasis/base_four_symm.py:160-557).  h_recipe / d_recipe: the class's recursion
 * recipe (gbasis_amd.programs.hrr_recipe, recipe_len int32) on the host and on the
 * device; max_dim_c / max_dim_d: the largest output dimension of the batch's shells of
 * l_c / l_d (W is sized for them; a quartet with larger ones is skipped, so the caller
 * passes the batch's maxima).  acc_layout 0 reads d_acc as gbx_eri_hrr does, 1 as
 * d_acc[(q * acc_parts + r) * acc_stride + f] (the quartet-major layout of
 * gbx_eri_rys_acc).  GBX_ERR_UNSUPPORTED when W and one row buffer exceed the LDS. */
int gbx_eri_hrr_lds(const gbx_tables *t, int32_t la, int32_t lb, int32_t lc, int32_t ld,
                    const int32_t *h_recipe, const int32_t *d_recipe, int32_t recipe_len,
                    int32_t max_dim_c, int32_t max_dim_d, const int32_t *d_quartets,
                    int64_t nquart, const int32_t *d_count, int32_t mode, const double *d_acc,
                    int64_t acc_stride, int32_t acc_parts, int32_t acc_layout, double *d_out,
                    void *stream);

/* Register form of gbx_eri_hrr for the split classes gbx_has_hrr_kernel lists, packed output
 * (mode GBX_MODE_PACKED or GBX_MODE_PACKED_ADD) and shells with a scaled-permutation output map (Cartesian, pure l <= 1): one
 * lane per quartet sums the acc_parts partials and runs the horizontal
 * recursion one ket column at a time (gbasis/integrals/_two_elec_int_improved.py:364-588,
 * gbasis/base_four_symm.py:160-244). */
int gbx_has_hrr_kernel(int32_t la, int32_t lb, int32_t lc, int32_t ld);
int gbx_eri_hrr_reg(const gbx_tables *t, int32_t la, int32_t lb, int32_t lc, int32_t ld,
                    const int32_t *d_quartets, int64_t nquart, const int32_t *d_count,
                    const double *d_acc, int64_t acc_stride, int32_t acc_parts, int32_t mode,
                    double *d_out, void *stream);

/* Expand the canonical 8-fold packed array of one molecule (nfunc output
 * functions) to the dense (nfunc)^4 tensor; notation 0 = chemist (ab|cd),
 * 1 = physicist <ab|cd> = (ac|bd).  Replaces the 8-fold block placement and
 * concatenation of gbasis/base_four_symm.py:215-244 and the notation transpose
 * of gbasis/integrals/electron_repulsion.py:279-280.  nfunc <= 32768. */
int gbx_unpack_dense(const double *d_packed, int64_t nfunc, int32_t notation, double *d_dense,
                     void *stream);

/* Coulomb and exchange matrices of one molecule straight from its canonical
 * packed ERIs (nfunc <= 65536): d_J[m*n+n'] = sum_ls (mn'|ls) D_ls and
 * d_K[m*n+l] = sum_ns (mn|ls) D_ns for a symmetric row-major density d_dens
 * (n x n, n = nfunc); d_J and d_K (n x n each) are overwritten.  No reference
 * entry point computes these: they replace the digestion its users write on top
 * of the dense array, e.g. the Coulomb / exchange energies of
 * notebooks/tutorial/JCP_paper.ipynb (cell 28) formed from
 * electron_repulsion_integral(basis, transform=C.T, notation="chemist"). */
int gbx_jk_packed(const double *d_packed, int64_t nfunc, const double *d_dens, double *d_J,
                  double *d_K, void *stream);

/* Size (in doubles) of the workspace gbx_mo_transform needs for nfunc basis
 * functions transformed to nnew new functions: the half-transformed array,
 * nfunc(nfunc+1)/2 * nnew(nnew+1)/2, followed by the first GEMM's output for one
 * chunk of matrices (at most max(2^27, nfunc * nnew) doubles). */
int gbx_mo_workspace(int64_t nfunc, int64_t nnew, int64_t *workspace_elems);

/* Four-index transform of one molecule's canonical packed ERIs into a new basis,
 * (ab|cd) = sum_ijkl T_ai T_bj T_ck T_dl (ij|kl), written to the dense nnew^4
 * array d_out in chemist (notation 0) or physicist (notation 1) order.  d_T is the
 * row-major nnew x nfunc transform and d_TT its transpose (nfunc x nnew); d_work is
 * a caller-owned workspace of gbx_mo_workspace doubles (the half-transformed
 * array, indexed [cd][ij], then GEMM staging).  No dense nfunc^4 array is formed;
 * both halves run as GEMMs on the FP64 matrix cores.  nfunc and nnew in [1, 32768].
 * Replaces the four tensordot passes of
 * BaseFourIndexSymmetric.construct_array_lincomb (gbasis/base_four_symm.py:559-616)
 * that electron_repulsion_integral(..., transform=T) runs
 * (gbasis/integrals/electron_repulsion.py:270-271). */
int gbx_mo_transform(const double *d_packed, int64_t nfunc, const double *d_T,
                     const double *d_TT, int64_t nnew, double *d_work, double *d_out,
                     int32_t notation, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GBASIS_AMD_H */
