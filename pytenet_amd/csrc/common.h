// Shared definitions for the gfx950 kernels of the MPS sweep library.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>
#include "../../include/ptn.h"

namespace ptn {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

struct cd { double x, y; };   // complex128, interleaved (matches torch/NumPy)
struct cf { float x, y; };    // complex64

__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cconj(cd a) { return {a.x, -a.y}; }

template <typename T> struct cplx_of;
template <> struct cplx_of<double> { typedef double2 type; };
template <> struct cplx_of<float> { typedef float2 type; };

// load a complex element of either precision as (re, im) in the compute type
template <typename Tc> __device__ __forceinline__ void load_c(const double2* p, Tc& re, Tc& im) {
  double2 v = *p; re = (Tc)v.x; im = (Tc)v.y;
}
template <typename Tc> __device__ __forceinline__ void load_c(const float2* p, Tc& re, Tc& im) {
  float2 v = *p; re = (Tc)v.x; im = (Tc)v.y;
}
__device__ __forceinline__ void store_c(double2* p, double re, double im) { *p = make_double2(re, im); }
__device__ __forceinline__ void store_c(float2* p, double re, double im) {
  *p = make_float2((float)re, (float)im);
}
__device__ __forceinline__ void store_c(float2* p, float re, float im) { *p = make_float2(re, im); }

#define PTN_HIP_CHECK(expr)                                  \
  do {                                                       \
    hipError_t e_ = (expr);                                  \
    if (e_ != hipSuccess) { ptn_set_last_hip_error(e_); return PTN_ERR_HIP; } \
  } while (0)

#define PTN_LAUNCH_CHECK()                                   \
  do {                                                       \
    hipError_t e_ = hipGetLastError();                       \
    if (e_ != hipSuccess) { ptn_set_last_hip_error(e_); return PTN_ERR_HIP; } \
  } while (0)

void ptn_set_last_hip_error(hipError_t e);

// live launch timing (timing.hip): timing_acquire hands out a start / stop
// event pair for the next launch on s (slot -1 and no events when timing is
// off or s is capturing); timing_commit attaches the kernel name and the
// launch's algorithmic flops / bytes. tlaunch launches a kernel with the pair
// attached to its dispatch (hipExtLaunchKernelGGL), or plainly.
int timing_acquire(hipStream_t s, hipEvent_t* e0, hipEvent_t* e1);
void timing_commit(int slot, const char* name, double flop, double bytes);
template <typename F, typename... Args>
inline void tlaunch(const char* name, double flop, double bytes, F kernel, dim3 grid, dim3 block,
                    hipStream_t s, Args... args) {
  hipEvent_t e0, e1;
  const int slot = timing_acquire(s, &e0, &e1);
  if (slot >= 0) {
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, args...);
    timing_commit(slot, name, flop, bytes);
  } else {
    kernel<<<grid, block, 0, s>>>(args...);
  }
}

// Device-side predication of the matvec products: while set (on the calling
// host thread), the product and mode-mix kernels of batch problem z return at
// once when p[z * stride] != 0 (e.g. the Lanczos "stopped" word of that
// problem), so no work is done after a breakdown.
struct SkipFlag {
  const int* p = nullptr;
  int64_t stride = 0;
};
SkipFlag& skip_flag();
struct SkipGuard {
  SkipGuard(const int* p, int64_t stride) { skip_flag() = SkipFlag{p, stride}; }
  ~SkipGuard() { skip_flag() = SkipFlag{}; }
};

// set (on the calling host thread) while an allocation-free entry point runs:
// the launch timing and the tile autotuner (which create and wait on events)
// stand aside
bool& noalloc_mode();
struct NoAllocGuard {
  NoAllocGuard() { noalloc_mode() = true; }
  ~NoAllocGuard() { noalloc_mode() = false; }
};

template <typename T> struct type_tag;
template <> struct type_tag<double> { static constexpr const char* s = "double"; };
template <> struct type_tag<float> { static constexpr const char* s = "float"; };
template <> struct type_tag<double2> { static constexpr const char* s = "c128"; };
template <> struct type_tag<float2> { static constexpr const char* s = "c64"; };

// ---------------------------------------------------------------------------
// Complex GEMM descriptor. Element (m, k) of op(A) is A[m*a_m + k*a_k]
// (conjugated if conjA), element (k, n) of op(B) is B[k*b_k + n*b_n]; C[m*c_m + n*c_n].
// Strides are in complex elements. bz_* are batch strides (blockIdx.z).
struct ZGemm {
  int64_t M, N, K;
  const void* A; int64_t a_m, a_k; int conjA;
  const void* B; int64_t b_k, b_n; int conjB;
  void* C; int64_t c_m, c_n;
  int accumulate;      // C += op(A) op(B) when 1
  int64_t batch;       // number of independent problems
  int64_t bz_a, bz_b, bz_c;
  // optional scratch for the split-bf16 operand planes of large products
  // (bf16x3_split_bytes); without it the split happens in the product kernel
  void* sws = nullptr;
  size_t sws_bytes = 0;
};
// scratch bytes of the pre-split split-bf16 product of an M x N x K (batched) product
size_t bf16x3_split_bytes(int64_t M, int64_t N, int64_t K, int64_t batch);
// the same for the three-term split (PTN_BF16X6)
size_t bf16x6_split_bytes(int64_t M, int64_t N, int64_t K, int64_t batch);

// element type codes for the GEMM / mix operands
enum { PTN_T_C128 = 0, PTN_T_C64 = 1 };

// compute: PTN_FP64, PTN_FP32, PTN_BF16X3 or PTN_BF16X6
int zgemm(const ZGemm& g, int compute, int ta, int tb, int tc,
          void* ws, size_t ws_bytes, hipStream_t s);

// Y[l, o, r] = sum_i Wm[o, i] X[l, i, r]   (complex; Wm complex128 on device);
// `batch` independent problems with contiguous X and Y and Wm stride bw
int mode_mix(int64_t nl, int64_t ni, int64_t no, int64_t nr, const void* wm, const void* x, int tx,
             void* y, int ty, hipStream_t s, int64_t batch = 1, int64_t bw = 0);

}  // namespace ptn
