// C-ABI entry points of the per-class register kernels (the kernels themselves
// are generated into gbx_reg_gen.hip by gbasis_amd/codegen.py).
#include "gbx_reg.h"

extern "C" int gbx_has_reg_kernel(int32_t la, int32_t lb, int32_t lc, int32_t ld) {
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > 7 || lb > 7 || lc > 7 || ld > 7) return 0;
  return gbx::has_reg_generated(la, lb, lc, ld) ? 1 : 0;
}

extern "C" int gbx_eri_reg(const gbx_tables* t, int32_t la, int32_t lb, int32_t lc, int32_t ld,
                           const int32_t* d_quartets, int64_t nquart, const int32_t* d_count,
                           int32_t precision, int32_t mode, double* d_out, void* stream) {
  if (!t || !d_out || (nquart > 0 && !d_quartets))
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_reg: null argument");
  if (precision != GBX_FP64 && precision != GBX_FP32_BULK)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_reg: bad precision");
  if (mode != GBX_MODE_PACKED && mode != GBX_MODE_PACKED_ADD && mode != GBX_MODE_SCHWARZ)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_reg: bad mode");
  if (gbx_has_reg_kernel(la, lb, lc, ld) == 0)
    return gbx::fail(GBX_ERR_UNSUPPORTED, "gbx_eri_reg: no register kernel for this class");
  if (t->potential != GBX_POT_COULOMB && mode != GBX_MODE_SCHWARZ)
    return gbx::fail(GBX_ERR_UNSUPPORTED, "gbx_eri_reg: register kernels evaluate 1/r only");
  if (la + lb + lc + ld + gbx::kTaylorTerms > t->boys_nord - 1 || t->boys_nt != gbx::kBoysRows)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_reg: Boys table does not fit this class / kernel");
  if (nquart == 0) return GBX_OK;
  gbx::RegArgs a{d_quartets, nquart, mode, d_out, nullptr, 0, 1, nullptr, nullptr, d_count};
  return gbx::launch_reg_generated(t, la, lb, lc, ld, a, precision,
                                   static_cast<hipStream_t>(stream));
}

extern "C" int gbx_eri_reg_segments(const gbx_tables* t, int32_t la, int32_t lb, int32_t lc,
                                    int32_t ld, const int32_t* d_quartets,
                                    const int64_t* d_seg_off, const int32_t* d_seg_cnt,
                                    int64_t max_quartets, int32_t precision, int32_t mode,
                                    double* d_out, void* stream) {
  if (!t || !d_out || !d_quartets || !d_seg_off || !d_seg_cnt || max_quartets < 0)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_reg_segments: bad argument");
  if (precision != GBX_FP64 && precision != GBX_FP32_BULK)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_reg_segments: bad precision");
  if (mode != GBX_MODE_PACKED && mode != GBX_MODE_PACKED_ADD && mode != GBX_MODE_SCHWARZ)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_reg_segments: bad mode");
  if (gbx_has_reg_kernel(la, lb, lc, ld) == 0)
    return gbx::fail(GBX_ERR_UNSUPPORTED, "gbx_eri_reg_segments: no register kernel for class");
  if (t->potential != GBX_POT_COULOMB && mode != GBX_MODE_SCHWARZ)
    return gbx::fail(GBX_ERR_UNSUPPORTED,
                     "gbx_eri_reg_segments: register kernels evaluate 1/r only");
  if (la + lb + lc + ld + gbx::kTaylorTerms > t->boys_nord - 1 || t->boys_nt != gbx::kBoysRows)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_reg_segments: Boys table does not fit this class");
  if (max_quartets == 0) return GBX_OK;
  gbx::RegArgs a{d_quartets, max_quartets, mode, d_out, nullptr, 0, 1, d_seg_off, d_seg_cnt,
                 nullptr};
  return gbx::launch_reg_generated(t, la, lb, lc, ld, a, precision,
                                   static_cast<hipStream_t>(stream));
}

extern "C" int gbx_split_count(int32_t la, int32_t lb, int32_t lc, int32_t ld) {
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > 7 || lb > 7 || lc > 7 || ld > 7) return 0;
  return gbx::split_count_generated(la, lb, lc, ld);
}

extern "C" int gbx_eri_split(const gbx_tables* t, int32_t la, int32_t lb, int32_t lc, int32_t ld,
                             int32_t split, const int32_t* d_quartets, int64_t nquart,
                             const int32_t* d_count, int32_t precision,
                             int32_t lanes_per_quartet, double* d_acc,
                             int64_t acc_stride, void* stream) {
  if (!t || (nquart > 0 && (!d_quartets || !d_acc)))
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_split: null argument");
  if (precision != GBX_FP64 && precision != GBX_FP32_BULK)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_split: bad precision");
  if (t->potential != GBX_POT_COULOMB)
    return gbx::fail(GBX_ERR_UNSUPPORTED, "gbx_eri_split: register kernels evaluate 1/r only");
  const int n = gbx_split_count(la, lb, lc, ld);
  if (split < 0 || split >= n)
    return gbx::fail(GBX_ERR_UNSUPPORTED, "gbx_eri_split: no such split for this class");
  if (acc_stride < nquart) return gbx::fail(GBX_ERR_ARG, "gbx_eri_split: acc_stride < nquart");
  if (lanes_per_quartet < 1 || lanes_per_quartet > 4096)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_split: lanes_per_quartet must be in [1, 4096]");
  if (la + lb + lc + ld + gbx::kTaylorTerms > t->boys_nord - 1)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_split: Boys table too short for this class");
  if (nquart == 0) return GBX_OK;
  gbx::RegArgs a{d_quartets, nquart, GBX_MODE_PACKED, nullptr, d_acc, acc_stride,
                 lanes_per_quartet, nullptr, nullptr, d_count};
  return gbx::launch_split_generated(t, la, lb, lc, ld, split, a, precision,
                                     static_cast<hipStream_t>(stream));
}

extern "C" int gbx_has_hrr_kernel(int32_t la, int32_t lb, int32_t lc, int32_t ld) {
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > 7 || lb > 7 || lc > 7 || ld > 7) return 0;
  return gbx::has_hrr_generated(la, lb, lc, ld) ? 1 : 0;
}

extern "C" int gbx_eri_hrr_reg(const gbx_tables* t, int32_t la, int32_t lb, int32_t lc,
                               int32_t ld, const int32_t* d_quartets, int64_t nquart,
                               const int32_t* d_count, const double* d_acc, int64_t acc_stride,
                               int32_t acc_parts, int32_t mode, double* d_out, void* stream) {
  if (!t || !d_out || (nquart > 0 && (!d_quartets || !d_acc)))
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_hrr_reg: null argument");
  if (gbx_has_hrr_kernel(la, lb, lc, ld) == 0)
    return gbx::fail(GBX_ERR_UNSUPPORTED, "gbx_eri_hrr_reg: no register HRR kernel for class");
  if (acc_stride < nquart || acc_parts < 1 || acc_parts > 4096)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_hrr_reg: acc_stride < nquart or bad acc_parts");
  if (mode != GBX_MODE_PACKED && mode != GBX_MODE_PACKED_ADD)
    return gbx::fail(GBX_ERR_ARG, "gbx_eri_hrr_reg: bad mode");
  if (nquart == 0) return GBX_OK;
  gbx::RegArgs a{d_quartets, nquart, mode, d_out, const_cast<double*>(d_acc),
                 acc_stride, acc_parts, nullptr, nullptr, d_count};
  return gbx::launch_hrr_generated(t, la, lb, lc, ld, a, static_cast<hipStream_t>(stream));
}
