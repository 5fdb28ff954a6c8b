// Complex matrix products, FP32 MFMA instantiations. Operands of either
// width are converted to FP32 when staged in LDS; the combinations built are
// the ones the contraction layer uses.
#include "gemm_impl.h"

namespace ptn {

int zgemm_fp32(const ZGemm& z, int ta, int tb, int tc, hipStream_t s, void* mws, size_t mws_bytes) {
  const int key = (ta << 2) | (tb << 1) | tc;
  switch (key) {
    case 1: return launch_typed<float, double2, double2, float2>(z, s, mws, mws_bytes);
    case 2: return launch_typed<float, double2, float2, double2>(z, s, mws, mws_bytes);
    case 3: return launch_typed<float, double2, float2, float2>(z, s, mws, mws_bytes);
    case 4: return launch_typed<float, float2, double2, double2>(z, s, mws, mws_bytes);
    case 5: return launch_typed<float, float2, double2, float2>(z, s, mws, mws_bytes);
    case 6: return launch_typed<float, float2, float2, double2>(z, s, mws, mws_bytes);
    case 7: return launch_typed<float, float2, float2, float2>(z, s, mws, mws_bytes);
  }
  return PTN_ERR_UNSUPPORTED;
}

}  // namespace ptn
