// Complex matrix products, FP64 MFMA instantiations and the zgemm dispatcher
// (FP64, FP32 and split-bf16 arithmetic).
#include "gemm_impl.h"

namespace ptn {

int zgemm_fp32(const ZGemm& z, int ta, int tb, int tc, hipStream_t s, void* mws, size_t mws_bytes);
int zgemm_bf16x3(const ZGemm& z, int ta, int tb, int tc, hipStream_t s, void* mws,
                 size_t mws_bytes);
int zgemm_bf16x6(const ZGemm& z, int ta, int tb, int tc, hipStream_t s, void* mws,
                 size_t mws_bytes);

int zgemm(const ZGemm& z, int compute, int ta, int tb, int tc, void* ws, size_t ws_bytes,
          hipStream_t s) {
  if (z.M < 0 || z.N < 0 || z.K < 0 || z.batch < 1) return PTN_ERR_ARG;
  if (z.M == 0 || z.N == 0) return PTN_OK;
  if (compute == PTN_FP64) {
    if (ta == PTN_T_C128 && tb == PTN_T_C128 && tc == PTN_T_C128)
      return launch_typed<double, double2, double2, double2>(z, s, ws, ws_bytes);
    return PTN_ERR_UNSUPPORTED;
  }
  if (compute == PTN_FP32) return zgemm_fp32(z, ta, tb, tc, s, ws, ws_bytes);
  if (compute == PTN_BF16X3) return zgemm_bf16x3(z, ta, tb, tc, s, ws, ws_bytes);
  if (compute == PTN_BF16X6) return zgemm_bf16x6(z, ta, tb, tc, s, ws, ws_bytes);
  return PTN_ERR_UNSUPPORTED;
}

}  // namespace ptn
