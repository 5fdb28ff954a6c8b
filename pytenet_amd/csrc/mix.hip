// Mode mix of the small MPO index pairs (see ops.hip).
#include "common.h"

namespace ptn {

// ---------------------------------------------------------------------------
// Mode mix: Y[l, o, r] = sum_i Wm[o, i] X[l, i, r] for the small MPO index
// pairs (ni, no <= 32). One thread per (l, r); the complex mixing matrix lives
// in LDS and its zero entries (most of an XXZ/Ising MPO) are skipped with a
// workgroup-uniform branch. HBM-bound: reads ni and writes no complex values
// per (l, r), so the bytes in flight per CU set its speed: the thread's ni
// inputs are held in registers sized to ni (CAP = 16 or 32), and complex64
// operands (the reduced-precision modes, whose products round T1 to FP32
// anyway) are mixed in FP32 arithmetic (A = float); both raise the waves per
// SIMD.
constexpr int MIX_MAX = 32;

// blockIdx.z selects the problem of a batch (strides in complex elements).
template <typename TX, typename TY, typename A, int CAP>
__global__ __launch_bounds__(256) void mode_mix_kernel(int64_t nl, int ni, int no, int64_t nr,
                                                       const double2* __restrict__ wm,
                                                       const TX* __restrict__ x,
                                                       TY* __restrict__ y, int64_t bw,
                                                       int64_t bx, int64_t by,
                                                       const int* __restrict__ skip,
                                                       int64_t skip_stride) {
  __shared__ double2 W[MIX_MAX * MIX_MAX];
  if (skip && skip[blockIdx.z * skip_stride]) return;
  wm += blockIdx.z * bw;
  x += blockIdx.z * bx;
  y += blockIdx.z * by;
  for (int t = threadIdx.x; t < ni * no; t += blockDim.x) W[t] = wm[t];
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t l = blockIdx.y;
  if (r >= nr) return;
  A xr[CAP], xi[CAP];
#pragma unroll
  for (int i = 0; i < CAP; ++i) {
    xr[i] = 0; xi[i] = 0;
    if (i < ni) load_c<A>(x + (l * ni + i) * nr + r, xr[i], xi[i]);
  }
  for (int o = 0; o < no; ++o) {
    A yr = 0, yi = 0;
#pragma unroll
    for (int i = 0; i < CAP; ++i) {
      if (i < ni) {
        const double2 w = W[o * ni + i];
        if (w.x != 0.0 || w.y != 0.0) {
          const A wr = (A)w.x, wi = (A)w.y;
          yr += wr * xr[i] - wi * xi[i];
          yi += wr * xi[i] + wi * xr[i];
        }
      }
    }
    store_c(y + (l * no + o) * nr + r, yr, yi);
  }
}

template <int CAP>
static void mix_launch(dim3 grid, int64_t nl, int ni, int no, int64_t nr, const void* wm,
                       const void* x, int tx, void* y, int ty, int64_t bw, int64_t bx, int64_t by,
                       const SkipFlag& sk, hipStream_t s) {
  if (tx == PTN_T_C128 && ty == PTN_T_C128)
    mode_mix_kernel<double2, double2, double, CAP><<<grid, 256, 0, s>>>(
        nl, ni, no, nr, (const double2*)wm, (const double2*)x, (double2*)y, bw, bx, by, sk.p,
        sk.stride);
  else if (tx == PTN_T_C64 && ty == PTN_T_C64)
    mode_mix_kernel<float2, float2, float, CAP><<<grid, 256, 0, s>>>(
        nl, ni, no, nr, (const double2*)wm, (const float2*)x, (float2*)y, bw, bx, by, sk.p,
        sk.stride);
  else if (tx == PTN_T_C128 && ty == PTN_T_C64)
    mode_mix_kernel<double2, float2, double, CAP><<<grid, 256, 0, s>>>(
        nl, ni, no, nr, (const double2*)wm, (const double2*)x, (float2*)y, bw, bx, by, sk.p,
        sk.stride);
  else
    mode_mix_kernel<float2, double2, double, CAP><<<grid, 256, 0, s>>>(
        nl, ni, no, nr, (const double2*)wm, (const float2*)x, (double2*)y, bw, bx, by, sk.p,
        sk.stride);
}

int mode_mix(int64_t nl, int64_t ni, int64_t no, int64_t nr, const void* wm, const void* x, int tx,
             void* y, int ty, hipStream_t s, int64_t batch, int64_t bw) {
  if (ni > MIX_MAX || no > MIX_MAX || ni < 1 || no < 1) return PTN_ERR_UNSUPPORTED;
  if (nl == 0 || nr == 0 || batch == 0) return PTN_OK;
  if (nl > 65535 || batch > 65535) return PTN_ERR_ARG;
  dim3 grid((unsigned)((nr + 255) / 256), (unsigned)nl, (unsigned)batch);
  const int64_t bx = nl * ni * nr, by = nl * no * nr;
  const SkipFlag sk = skip_flag();
  if (ni <= 16)
    mix_launch<16>(grid, nl, (int)ni, (int)no, nr, wm, x, tx, y, ty, bw, bx, by, sk, s);
  else
    mix_launch<MIX_MAX>(grid, nl, (int)ni, (int)no, nr, wm, x, tx, y, ty, bw, bx, by, sk, s);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

}  // namespace ptn
