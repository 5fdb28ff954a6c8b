// Complex matrix products in split-bf16 arithmetic ("bf16x3") on the CDNA4
// MFMA v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Every operand value x (rounded to FP32 first) is split into two bf16 terms
//   hi = bf16(x),  lo = bf16(x - hi)
// and a product is formed from the three leading terms
//   a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi
// (a_lo b_lo, of relative size 2^-16, is dropped), which carries about 16
// significant bits per product, accumulated in FP32.
//
// The complex product uses the same real embedding as the FP32/FP64 kernels,
//   [Re C, Im C] = [Re A, Im A] . [[Re B, Im B], [-Im B, Re B]],
// with real k = 2 k_complex + (0: real part, 1: imaginary part). For the bf16
// MFMA a lane (i = lane % 16, g = lane / 16) holds A[i][8g .. 8g+7] and
// B[8g .. 8g+7][i] in real k, i.e. four consecutive complex k values of one
// complex row of A (interleaved re/im: the natural complex layout) and of one
// complex column of B, expanded on the fly: real column 2n takes
// (Re B, -Im B), real column 2n+1 takes (Im B, Re B) per complex k. The 16x16
// FP32 result has row 4*(lane/16) + v and column lane % 16.
#include "gemm_impl.h"

namespace ptn {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// hi/lo bf16 split of eight FP32 values
__device__ __forceinline__ void split8(const float (&x)[8], bf16x8_t& hi, bf16x8_t& lo) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    __bf16 h = (__bf16)x[e];
    hi[e] = h;
    lo[e] = (__bf16)(x[e] - (float)h);
  }
}

// Workgroup = 4 waves as 2 x 2; each wave owns (16*TM) rows x (8*TN) complex columns.
template <int TM, int TN> struct TileCfgB {
  static constexpr int WM = 16 * TM;
  static constexpr int WN = 8 * TN;
  static constexpr int BM = 2 * WM;
  static constexpr int BN = 2 * WN;
  static constexpr int BK = 32;             // complex k per LDS stage = 2 MFMA k-steps
  static constexpr int LDK = 2 * BK + 4;    // floats per LDS row (padded)
  static constexpr int NA = BM * BK / 256;
  static constexpr int NB = BN * BK / 256;
};

template <typename TA, typename TB, typename TC, int TM, int TN, bool A_KCONTIG, bool B_NCONTIG>
__global__ __launch_bounds__(256) void zgemm_bf16x3_kernel(GemmKArgs g) {
  typedef TileCfgB<TM, TN> T;
  constexpr int BM = T::BM, BN = T::BN, BK = T::BK, LDK = T::LDK, NA = T::NA, NB = T::NB;

  // A: row-major, real k contiguous; B: column-major (complex column n), real k contiguous
  __shared__ __align__(16) float As[BM * LDK];
  __shared__ __align__(16) float Bs[BN * LDK];
  if (g.skip && g.skip[blockIdx.z * g.skip_stride]) return;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * T::WM;
  const int wn0 = (wave & 1) * T::WN;

  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n, tm, tn);
  const int64_t m0 = (int64_t)tm * BM;
  const int64_t n0 = (int64_t)tn * BN;
  const int64_t bz = blockIdx.z;
  const TA* __restrict__ A = (const TA*)g.A + bz * g.bz_a;
  const TB* __restrict__ B = (const TB*)g.B + bz * g.bz_b;
  TC* __restrict__ C = (TC*)g.C + bz * g.bz_c;
  const float sa = g.conjA ? -1.f : 1.f;
  const float sb = g.conjB ? -1.f : 1.f;

  float ra[NA][2], rb[NB][2];
  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      int e = tid + 256 * i;
      int row, kc;
      if (A_KCONTIG) { row = e / BK; kc = e % BK; } else { kc = e / BM; row = e % BM; }
      int64_t m = m0 + row, k = k0 + kc;
      float re = 0, im = 0;
      if (m < g.M && k < g.K) load_c<float>(A + m * g.a_m + k * g.a_k, re, im);
      ra[i][0] = re; ra[i][1] = sa * im;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      int e = tid + 256 * i;
      int col, kc;
      if (B_NCONTIG) { kc = e / BN; col = e % BN; } else { col = e / BK; kc = e % BK; }
      int64_t n = n0 + col, k = k0 + kc;
      float re = 0, im = 0;
      if (n < g.N && k < g.K) load_c<float>(B + k * g.b_k + n * g.b_n, re, im);
      rb[i][0] = re; rb[i][1] = sb * im;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      int e = tid + 256 * i;
      int row, kc;
      if (A_KCONTIG) { row = e / BK; kc = e % BK; } else { kc = e / BM; row = e % BM; }
      *(float2*)&As[row * LDK + 2 * kc] = make_float2(ra[i][0], ra[i][1]);
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      int e = tid + 256 * i;
      int col, kc;
      if (B_NCONTIG) { kc = e / BN; col = e % BN; } else { col = e / BK; kc = e % BK; }
      *(float2*)&Bs[col * LDK + 2 * kc] = make_float2(rb[i][0], rb[i][1]);
    }
  };

  f4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4_t{0, 0, 0, 0};

  const int grp = lane >> 4;          // k group: real k 8*grp .. 8*grp+7
  const int pb = lane & 1;            // real column parity: 0 -> Re C, 1 -> Im C

  const int nk = (int)((g.K + BK - 1) / BK);
  int t = next_stage(g, tm, tn, 0, nk);
  if (t < nk) load_tiles((int64_t)t * BK);
  while (t < nk) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    const int tnext = next_stage(g, tm, tn, t + 1, nk);
    if (tnext < nk) load_tiles((int64_t)tnext * BK);
    t = tnext;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int kof = ks * 32 + 8 * grp;   // real k offset within the stage
      bf16x8_t ahi[TM], alo[TM], bhi[TN], blo[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const float* p = &As[(wm0 + 16 * i + (lane & 15)) * LDK + kof];
        float4 u0 = *(const float4*)p, u1 = *(const float4*)(p + 4);
        float x[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
        split8(x, ahi[i], alo[i]);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const float* p = &Bs[(wn0 + 8 * j + ((lane & 15) >> 1)) * LDK + kof];
        float4 u0 = *(const float4*)p, u1 = *(const float4*)(p + 4);
        // (re, im) of four complex k; expand to the real-embedding column
        float x[8];
        if (pb == 0) {
          x[0] = u0.x; x[1] = -u0.y; x[2] = u0.z; x[3] = -u0.w;
          x[4] = u1.x; x[5] = -u1.y; x[6] = u1.z; x[7] = -u1.w;
        } else {
          x[0] = u0.y; x[1] = u0.x; x[2] = u0.w; x[3] = u0.z;
          x[4] = u1.y; x[5] = u1.x; x[6] = u1.w; x[7] = u1.z;
        }
        split8(x, bhi[j], blo[j]);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[i], bhi[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[i], blo[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[i], bhi[j], acc[i][j], 0, 0, 0);
        }
    }
  }

  // epilogue as in the FP32 kernel: the partner lane (lane ^ 1) holds the
  // other part of the same complex column; even lanes store the rows of
  // accumulator elements 0, 1 and odd lanes those of elements 2, 3
#pragma unroll
  for (int i = 0; i < TM; ++i) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      float mine[4], other[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        mine[v] = acc[i][j][v];
        other[v] = __shfl_xor(mine[v], 1);
      }
      const int64_t n = n0 + wn0 + 8 * j + ((lane & 15) >> 1);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int v = 2 * pb + h;
        const int64_t m = m0 + wm0 + 16 * i + 4 * grp + v;
        double re = pb ? (double)other[2 + h] : (double)mine[h];
        double im = pb ? (double)mine[2 + h] : (double)other[h];
        if (m < g.M && n < g.N) {
          TC* p = C + m * g.c_m + n * g.c_n;
          if (g.accumulate) {
            double pr, pi;
            load_c<double>(p, pr, pi);
            re += pr; im += pi;
          }
          store_c(p, re, im);
        }
      }
    }
  }
}

template <typename TA, typename TB, typename TC, int TM, int TN>
inline int launch_cfg_bf16(const ZGemm& z, hipStream_t s, void* mws, size_t mws_bytes) {
  typedef TileCfgB<TM, TN> T;
  GemmKArgs g;
  g.M = z.M; g.N = z.N; g.K = z.K;
  g.A = z.A; g.a_m = z.a_m; g.a_k = z.a_k; g.conjA = z.conjA;
  g.B = z.B; g.b_k = z.b_k; g.b_n = z.b_n; g.conjB = z.conjB;
  g.C = z.C; g.c_m = z.c_m; g.c_n = z.c_n;
  g.accumulate = z.accumulate;
  g.bz_a = z.bz_a; g.bz_b = z.bz_b; g.bz_c = z.bz_c;
  g.tiles_m = (int)((z.M + T::BM - 1) / T::BM);
  g.tiles_n = (int)((z.N + T::BN - 1) / T::BN);
  g.skip = skip_flag().p;
  g.skip_stride = skip_flag().stride;
  int64_t nwg = (int64_t)g.tiles_m * g.tiles_n;
  if (nwg > 0x7fffffff || z.batch > 65535) return PTN_ERR_ARG;
  int st = build_masks<TA, TB>(z, g, T::BM, T::BN, T::BK, mws, mws_bytes, s);
  if (st) return st;
  dim3 grid((unsigned)nwg, 1, (unsigned)z.batch);
  const bool ak = (z.a_k == 1) || (z.a_m != 1);
  const bool bn = (z.b_n == 1) || (z.b_k != 1);
  const int slot = timing_pre(s);
  if (ak && bn)
    zgemm_bf16x3_kernel<TA, TB, TC, TM, TN, true, true><<<grid, 256, 0, s>>>(g);
  else if (ak && !bn)
    zgemm_bf16x3_kernel<TA, TB, TC, TM, TN, true, false><<<grid, 256, 0, s>>>(g);
  else if (!ak && bn)
    zgemm_bf16x3_kernel<TA, TB, TC, TM, TN, false, true><<<grid, 256, 0, s>>>(g);
  else
    zgemm_bf16x3_kernel<TA, TB, TC, TM, TN, false, false><<<grid, 256, 0, s>>>(g);
  PTN_LAUNCH_CHECK();
  if (slot >= 0)
    timing_post(slot, s,
                gemm_kernel_name<float, TA, TB, TC, TM, TN>(ak, bn, "zgemm_bf16x3_kernel"),
                gemm_flop(z), gemm_bytes<TA, TB, TC>(z));
  return PTN_OK;
}

template <typename TA, typename TB, typename TC>
inline int launch_typed_bf16(const ZGemm& z, hipStream_t s, void* mws, size_t mws_bytes) {
  int64_t big = ((z.M + 127) / 128) * ((z.N + 63) / 64) * z.batch;
  if (big >= 512) return launch_cfg_bf16<TA, TB, TC, 4, 4>(z, s, mws, mws_bytes);
  return launch_cfg_bf16<TA, TB, TC, 2, 2>(z, s, mws, mws_bytes);
}

int zgemm_bf16x3(const ZGemm& z, int ta, int tb, int tc, hipStream_t s, void* mws,
                 size_t mws_bytes) {
  const int key = (ta << 2) | (tb << 1) | tc;
  switch (key) {
    case 1: return launch_typed_bf16<double2, double2, float2>(z, s, mws, mws_bytes);
    case 2: return launch_typed_bf16<double2, float2, double2>(z, s, mws, mws_bytes);
    case 3: return launch_typed_bf16<double2, float2, float2>(z, s, mws, mws_bytes);
    case 4: return launch_typed_bf16<float2, double2, double2>(z, s, mws, mws_bytes);
    case 5: return launch_typed_bf16<float2, double2, float2>(z, s, mws, mws_bytes);
    case 6: return launch_typed_bf16<float2, float2, double2>(z, s, mws, mws_bytes);
  }
  return PTN_ERR_UNSUPPORTED;
}

}  // namespace ptn
