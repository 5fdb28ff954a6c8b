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
#include <algorithm>
#include <string>
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
  const char* nm = gemm_kernel_name<float, TA, TB, TC, TM, TN>(ak, bn, "zgemm_bf16x3_kernel");
  const double fl = gemm_flop(z), by = gemm_bytes<TA, TB, TC>(z);
  if (ak && bn)
    tlaunch(nm, fl, by, zgemm_bf16x3_kernel<TA, TB, TC, TM, TN, true, true>, grid, dim3(256), s, g);
  else if (ak && !bn)
    tlaunch(nm, fl, by, zgemm_bf16x3_kernel<TA, TB, TC, TM, TN, true, false>, grid, dim3(256), s, g);
  else if (!ak && bn)
    tlaunch(nm, fl, by, zgemm_bf16x3_kernel<TA, TB, TC, TM, TN, false, true>, grid, dim3(256), s, g);
  else
    tlaunch(nm, fl, by, zgemm_bf16x3_kernel<TA, TB, TC, TM, TN, false, false>, grid, dim3(256), s,
            g);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// ---------------------------------------------------------------------------
// Pre-split form for large products. The two operands are split once, in a
// pass of their own, into bf16 hi / lo planes of the real embedding:
//   A_r (Mp x K2p, row-major): A_r[m][2k] = Re A(m, k), A_r[m][2k+1] = Im A(m, k)
//   B_e (N2p x K2p, row-major, i.e. K-contiguous columns of the real B):
//     B_e[2n][2k] = Re B(k, n),   B_e[2n][2k+1] = -Im B(k, n)
//     B_e[2n+1][2k] = Im B(k, n), B_e[2n+1][2k+1] = Re B(k, n)
// (conjugation flips the signs of the Im terms), zero-padded to the tile
// multiples (Mp, N2p: 256; K2p: 32), so that C_r = A_r B_e^T gives
// C_r[m][2n] = Re C(m, n) and C_r[m][2n+1] = Im C(m, n). The product itself
// is then a plain bf16 GEMM with FP32 accumulation and three MFMAs per pair of
// operand fragments (hi hi + hi lo + lo hi): v_mfma_f32_32x32x16_bf16, 256 x
// 256 real tiles, 8 waves (2 x 4) of 128 x 64, 32-deep k-stages double
// buffered in LDS (4 planes x 256 rows x 40 bf16 per stage; the 80-byte row
// pitch keeps the 16-byte fragment reads of a wave on distinct banks).
constexpr int V2_BM = 256, V2_BN = 256, V2_BK = 32, V2_LD = 40;

inline int64_t v2_round(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// k-chunks of a single product whose tiles alone do not fill the chip (each
// workgroup of the product holds a whole CU): 1, 2 or 4, keeping at least
// 256 real k per chunk
inline int v2_splitk(int64_t M, int64_t N, int64_t K, int64_t batch) {
  if (batch != 1) return 1;
  const int64_t tiles = ((M + V2_BM - 1) / V2_BM) * ((2 * N + V2_BN - 1) / V2_BN);
  const int64_t K2p = v2_round(2 * K, V2_BK);
  int sk = 1;
  while (sk < 4 && tiles * sk < 256 && K2p / (2 * sk) >= 256) sk *= 2;
  return sk;
}

// bytes of the split planes of one product (per batch problem: times batch),
// plus the FP32 partial sums of a k-chunked product
inline size_t v2_split_bytes(int64_t M, int64_t N, int64_t K, int64_t batch) {
  const int64_t Mp = v2_round(M, V2_BM), N2p = v2_round(2 * N, V2_BN), K2p = v2_round(2 * K, V2_BK);
  const int sk = v2_splitk(M, N, K, batch);
  const size_t part = sk > 1 ? (size_t)sk * (size_t)Mp * (size_t)N2p * sizeof(float) + 256 : 0;
  return (size_t)batch * (size_t)(Mp + N2p) * (size_t)K2p * 2 * sizeof(__bf16) + 512 + part;
}

__device__ __forceinline__ void split1(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// plane `plane` (0 hi, 1 lo) of eight values
__device__ __forceinline__ bf16x8_t plane8(const float (&x)[8], int plane) {
  bf16x8_t o;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    __bf16 h, l;
    split1(x[e], h, l);
    o[e] = plane ? l : h;
  }
  return o;
}

// A (M x K complex, element (m, k) at A[m a_m + k a_k], conjugated if conj)
// -> A_r hi / lo planes. One workgroup per 64 x 32 complex block, staged
// through LDS so that both the reads (along the operand's contiguous index)
// and the writes (along k) are coalesced: each store is a 16-byte chunk
// (eight real k of one plane), a row's 64 real k of a plane being eight
// consecutive chunks written by consecutive threads; problem blockIdx.z at
// A + z bz.
template <typename TA>
__global__ __launch_bounds__(256) void k_split_a(int64_t M, int64_t K, const TA* __restrict__ A,
                                                 int64_t a_m, int64_t a_k, int conj, int64_t bz,
                                                 int64_t K2p, int64_t pz, __bf16* __restrict__ hi,
                                                 __bf16* __restrict__ lo) {
  __shared__ float2 t[64][33];
  const int64_t m0 = (int64_t)blockIdx.y * 64, k0 = (int64_t)blockIdx.x * 32;
  A += blockIdx.z * bz;
  hi += blockIdx.z * pz;
  lo += blockIdx.z * pz;
  const bool kc = (a_k == 1) || (a_m != 1);
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int r = kc ? e / 32 : e % 64, c = kc ? e % 32 : e / 64;
    const int64_t m = m0 + r, k = k0 + c;
    float re = 0.f, im = 0.f;
    if (m < M && k < K) load_c<float>(A + m * a_m + k * a_k, re, im);
    t[r][c] = make_float2(re, conj ? -im : im);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 64 * 16; q += 256) {
    const int r = q / 16, plane = (q % 16) >> 3, c8 = q % 8;
    const int64_t k2 = 2 * k0 + 8 * c8;   // first real k of the chunk
    if (k2 >= K2p) continue;
    float x[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float2 v = t[r][4 * c8 + j];
      x[2 * j] = v.x;
      x[2 * j + 1] = v.y;
    }
    *(bf16x8_t*)((plane ? lo : hi) + (m0 + r) * K2p + k2) = plane8(x, plane);
  }
}

// B (K x N complex, element (k, n) at B[k b_k + n b_n]) -> B_e hi / lo planes
template <typename TB>
__global__ __launch_bounds__(256) void k_split_b(int64_t K, int64_t N, const TB* __restrict__ B,
                                                 int64_t b_k, int64_t b_n, int conj, int64_t bz,
                                                 int64_t K2p, int64_t pz, __bf16* __restrict__ hi,
                                                 __bf16* __restrict__ lo) {
  __shared__ float2 t[64][33];
  const int64_t n0 = (int64_t)blockIdx.y * 64, k0 = (int64_t)blockIdx.x * 32;
  B += blockIdx.z * bz;
  hi += blockIdx.z * pz;
  lo += blockIdx.z * pz;
  const bool kc = (b_k == 1) || (b_n != 1);
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int r = kc ? e / 32 : e % 64, c = kc ? e % 32 : e / 64;
    const int64_t n = n0 + r, k = k0 + c;
    float re = 0.f, im = 0.f;
    if (n < N && k < K) load_c<float>(B + k * b_k + n * b_n, re, im);
    t[r][c] = make_float2(re, conj ? -im : im);
  }
  __syncthreads();
  // real rows 2 n (Re, -Im per complex k) and 2 n + 1 (Im, Re), 16-byte chunks
  // as in k_split_a
  for (int q = threadIdx.x; q < 128 * 16; q += 256) {
    const int rr = q / 16, plane = (q % 16) >> 3, c8 = q % 8;
    const int64_t k2 = 2 * k0 + 8 * c8;
    if (k2 >= K2p) continue;
    const int r = rr >> 1;
    const bool odd = rr & 1;
    float x[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float2 v = t[r][4 * c8 + j];
      x[2 * j] = odd ? v.y : v.x;
      x[2 * j + 1] = odd ? v.x : -v.y;
    }
    *(bf16x8_t*)((plane ? lo : hi) + (2 * n0 + rr) * K2p + k2) = plane8(x, plane);
  }
}

typedef float f16v_t __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

struct V2Args {
  const __bf16 *ahi, *alo, *bhi, *blo;
  int64_t K2p, pza, pzb;      // k extent, batch strides of the A and B planes
  int64_t kchunk;             // real k per chunk of a k-chunked product (blockIdx.z = chunk)
  float* part;                // FP32 partial sums (Mp x N2p per chunk), or null
  int64_t N2p;
  int64_t M, N;               // complex output extent
  int tiles_m, tiles_n;
  void* C;
  int64_t c_m, c_n, bz_c;
  int accumulate;
  const int* skip;
  int64_t skip_stride;
};

template <typename TC>
__global__ __launch_bounds__(512) void zgemm_bf16x3_v2(V2Args g) {
  __shared__ __align__(16) __bf16 sm[2][4][V2_BM][V2_LD];
  // predication by the problem's flag (blockIdx.z is the k-chunk of a chunked
  // product, which has a single problem)
  if (g.skip && g.skip[(g.part ? 0 : blockIdx.z) * g.skip_stride]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n, tm, tn);
  const int64_t m0 = (int64_t)tm * V2_BM, c0 = (int64_t)tn * V2_BN;   // c0: real column
  // blockIdx.z: the problem of a batch, or the k-chunk of a chunked product
  const int64_t z = g.part ? 0 : blockIdx.z;
  const int64_t kbeg = g.part ? (int64_t)blockIdx.z * g.kchunk : 0;
  const __bf16* src[4] = {g.ahi + z * g.pza + m0 * g.K2p + kbeg, g.alo + z * g.pza + m0 * g.K2p + kbeg,
                          g.bhi + z * g.pzb + c0 * g.K2p + kbeg, g.blo + z * g.pzb + c0 * g.K2p + kbeg};
  // global -> register staging: per plane 256 rows x 32 bf16 = 1024 chunks of
  // 16 bytes, two per thread
  u32x4_t pre[4][2];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int c = tid + 512 * u, row = c >> 2, part = c & 3;
        pre[p][u] = *(const u32x4_t*)(src[p] + (int64_t)row * g.K2p + k0 + 8 * part);
      }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int c = tid + 512 * u, row = c >> 2, part = c & 3;
        *(u32x4_t*)&sm[buf][p][row][8 * part] = pre[p][u];
      }
  };
  f16v_t acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
  // real k of this workgroup: the whole K2p, or its chunk (the last one may be shorter)
  const int64_t klen = g.part ? (g.K2p - kbeg < g.kchunk ? g.K2p - kbeg : g.kchunk) : g.K2p;
  const int nst = (int)(klen / V2_BK);
  const int r32 = lane & 31, kg = 8 * (lane >> 5);
  gload(0);
  lstore(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int buf = st & 1;
    const bool more = st + 1 < nst;
    if (more) gload((int64_t)(st + 1) * V2_BK);
#pragma unroll
    for (int ks = 0; ks < V2_BK / 16; ++ks) {
      const int kk = 16 * ks + kg;
      bf16x8_t ah[4], al[4], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 128 * wm + 32 * i + r32;
        ah[i] = *(const bf16x8_t*)&sm[buf][0][row][kk];
        al[i] = *(const bf16x8_t*)&sm[buf][1][row][kk];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = 64 * wn + 32 * j + r32;
        bh[j] = *(const bf16x8_t*)&sm[buf][2][col][kk];
        bl[j] = *(const bf16x8_t*)&sm[buf][3][col][kk];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    if (more) lstore(buf ^ 1);
    __syncthreads();
  }
  // epilogue: element v of a 32 x 32 result sits at column lane % 32 and row
  // 8 (v / 4) + 4 (lane / 32) + v % 4; real column 2n is Re, 2n + 1 Im, held by
  // the lane pair (2l, 2l + 1); the even lane writes the complex value
  if (g.part) {
    // k-chunk partial sums, real layout (summed by k_v2_reduce)
    float* P = g.part + (int64_t)blockIdx.z * (int64_t)g.tiles_m * V2_BM * g.N2p;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t c = c0 + 64 * wn + 32 * j + r32;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int64_t m = m0 + 128 * wm + 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
          P[m * g.N2p + c] = acc[i][j][v];
        }
      }
    return;
  }
  TC* C = (TC*)g.C + z * g.bz_c;
  const int pb = lane & 1;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t n = (c0 + 64 * wn + 32 * j + r32) >> 1;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float mine = acc[i][j][v];
        const float other = __shfl_xor(mine, 1);
        const int64_t m = m0 + 128 * wm + 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (!pb && m < g.M && n < g.N) {
          double re = mine, im = other;
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

// C (complex, strides c_m, c_n; += when accumulating) = sum over the sk
// k-chunk partials P[s] (Mp x N2p real) in a fixed order
template <typename TC>
__global__ void k_v2_reduce(int64_t M, int64_t N, int64_t Mp, int64_t N2p, int sk,
                            const float* __restrict__ P, TC* __restrict__ C, int64_t c_m,
                            int64_t c_n, int accumulate, const int* __restrict__ skip) {
  if (skip && skip[0]) return;   // the single problem is predicated off
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < M * N;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = t / N, n = t % N;
    double re = 0.0, im = 0.0;
    for (int s = 0; s < sk; ++s) {
      const float2 v = *(const float2*)(P + (int64_t)s * Mp * N2p + m * N2p + 2 * n);
      re += v.x;
      im += v.y;
    }
    TC* p = C + m * c_m + n * c_n;
    if (accumulate) {
      double pr, pi;
      load_c<double>(p, pr, pi);
      re += pr; im += pi;
    }
    store_c(p, re, im);
  }
}

// the pre-split path: split passes into `sws`, then the v2 product
template <typename TA, typename TB, typename TC>
inline int launch_bf16_v2(const ZGemm& z, hipStream_t s) {
  const int64_t Mp = v2_round(z.M, V2_BM), N2p = v2_round(2 * z.N, V2_BN),
                K2p = v2_round(2 * z.K, V2_BK);
  const int64_t pza = Mp * K2p, pzb = N2p * K2p;
  char* base = (char*)(((uintptr_t)z.sws + 255) & ~(uintptr_t)255);
  __bf16* ahi = (__bf16*)base;
  __bf16* alo = ahi + z.batch * pza;
  __bf16* bhi = alo + z.batch * pza;
  __bf16* blo = bhi + z.batch * pzb;
  const unsigned bt = (unsigned)z.batch;
  const unsigned kblk = (unsigned)((K2p + 63) / 64);   // 32 complex k per block
  k_split_a<TA><<<dim3(kblk, (unsigned)(Mp / 64), bt), 256, 0, s>>>(
      z.M, z.K, (const TA*)z.A, z.a_m, z.a_k, z.conjA, z.bz_a, K2p, pza, ahi, alo);
  k_split_b<TB><<<dim3(kblk, (unsigned)(N2p / 128), bt), 256, 0, s>>>(
      z.K, z.N, (const TB*)z.B, z.b_k, z.b_n, z.conjB, z.bz_b, K2p, pzb, bhi, blo);
  PTN_LAUNCH_CHECK();
  const int sk = v2_splitk(z.M, z.N, z.K, z.batch);
  V2Args g;
  g.ahi = ahi; g.alo = alo; g.bhi = bhi; g.blo = blo;
  g.K2p = K2p; g.pza = pza; g.pzb = pzb;
  g.N2p = N2p;
  g.kchunk = v2_round((K2p + sk - 1) / sk, V2_BK);
  g.part = nullptr;
  if (sk > 1)
    g.part = (float*)(((uintptr_t)(blo + z.batch * pzb) + 255) & ~(uintptr_t)255);
  g.M = z.M; g.N = z.N;
  g.tiles_m = (int)(Mp / V2_BM);
  g.tiles_n = (int)(N2p / V2_BN);
  g.C = z.C; g.c_m = z.c_m; g.c_n = z.c_n; g.bz_c = z.bz_c;
  g.accumulate = z.accumulate;
  g.skip = skip_flag().p;
  g.skip_stride = skip_flag().stride;
  // the chunks cover K2p: every chunk but the last is kchunk deep
  const unsigned nz = sk > 1 ? (unsigned)((K2p + g.kchunk - 1) / g.kchunk) : bt;
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), 1, nz);
  static const std::string v2name = std::string("zgemm_bf16x3_v2<") + type_tag<TC>::s + ">";
  tlaunch(v2name.c_str(), gemm_flop(z), gemm_bytes<TA, TB, TC>(z), zgemm_bf16x3_v2<TC>, grid,
          dim3(512), s, g);
  if (sk > 1)
    k_v2_reduce<TC><<<(unsigned)std::min<int64_t>((z.M * z.N + 255) / 256, 4096), 256, 0, s>>>(
        z.M, z.N, Mp, N2p, (int)nz, g.part, (TC*)z.C, z.c_m, z.c_n, z.accumulate, g.skip);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// the pre-split product for shapes that fill the chip with 256 x 256 tiles and
// have room for the split planes in the caller's workspace
inline bool use_bf16_v2(const ZGemm& z, void* mws) {
  if (mws || !z.sws) return false;
  const int64_t tiles = ((z.M + V2_BM - 1) / V2_BM) * ((2 * z.N + V2_BN - 1) / V2_BN) * z.batch;
  if (tiles * v2_splitk(z.M, z.N, z.K, z.batch) < 128 || z.K < 16 || z.batch > 65535) return false;
  return z.sws_bytes >= v2_split_bytes(z.M, z.N, z.K, z.batch);
}

template <typename TA, typename TB, typename TC>
inline int launch_typed_bf16(const ZGemm& z, hipStream_t s, void* mws, size_t mws_bytes) {
  if (use_bf16_v2(z, mws)) return launch_bf16_v2<TA, TB, TC>(z, s);
  int64_t big = ((z.M + 127) / 128) * ((z.N + 63) / 64) * z.batch;
  if (big >= 512) return launch_cfg_bf16<TA, TB, TC, 4, 4>(z, s, mws, mws_bytes);
  return launch_cfg_bf16<TA, TB, TC, 2, 2>(z, s, mws, mws_bytes);
}

size_t bf16x3_split_bytes(int64_t M, int64_t N, int64_t K, int64_t batch) {
  return v2_split_bytes(M, N, K, batch);
}

// ---------------------------------------------------------------------------
// Three-term split ("bf16x6", precision PTN_BF16X6): every FP32-rounded
// operand value is written as
//   x = hi + mid + lo,  hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid)
// (each difference is exact in FP32, and the three terms hold the whole 24-bit
// significand), and a product is formed from the six leading terms
//   a b ~ a_lo b_hi + a_hi b_lo + a_mid b_mid + a_mid b_hi + a_hi b_mid + a_hi b_hi
// (smallest first; the dropped a_mid b_lo, a_lo b_mid, a_lo b_lo are at most
// 2^-23 |a b|: |mid| <= 2^-8 |x|, |lo| <= 2^-16 |x|), accumulated in FP32: the
// accuracy of an FP32 product on the bf16 MFMA, with twice the MFMA work of
// bf16x3 and 1.5 times its operand bytes. The three terms of a value are stored interleaved per 16 real k
// ([row][k / 16][term][k % 16]), so that one k-stage of a row is one
// contiguous 96-byte read. Products too small for the pre-split kernel run on
// the FP32 MFMA (the same accuracy).
constexpr int V3_BK = 16, V3_LD = 24, V3_NT = 3;

inline int64_t v3_kpad(int64_t K) { return v2_round(2 * K, V3_BK); }

inline size_t v3_split_bytes(int64_t M, int64_t N, int64_t K, int64_t batch) {
  const int64_t Mp = v2_round(M, V2_BM), N2p = v2_round(2 * N, V2_BN), K2p = v3_kpad(K);
  const int sk = v2_splitk(M, N, K, batch);
  const size_t part = sk > 1 ? (size_t)sk * (size_t)Mp * (size_t)N2p * sizeof(float) + 256 : 0;
  return (size_t)batch * (size_t)(Mp + N2p) * (size_t)K2p * V3_NT * sizeof(__bf16) + 512 + part;
}

__device__ __forceinline__ void split3(float x, __bf16& hi, __bf16& mid, __bf16& lo) {
  hi = (__bf16)x;
  const float r = x - (float)hi;
  mid = (__bf16)r;
  lo = (__bf16)(r - (float)mid);
}

// term `term` (0 hi, 1 mid, 2 lo) of eight values
__device__ __forceinline__ bf16x8_t term8(const float (&x)[8], int term) {
  bf16x8_t o;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    __bf16 h, d, l;
    split3(x[e], h, d, l);
    o[e] = term == 0 ? h : (term == 1 ? d : l);
  }
  return o;
}

// The split kernels stage a 64 x 32 complex block in LDS and write it as
// 16-byte chunks (eight real k of one term): a row's four k-blocks of 16 real
// k are 4 x 3 x 2 = 24 consecutive chunks (384 contiguous bytes), written by
// consecutive threads, so every store is a whole chunk of a contiguous run.
constexpr int V3_CHUNKS = 4 * V3_NT * 2;

// element offset of chunk (term, half) of k-block kb of a row
__device__ __forceinline__ int64_t v3_chunk(int64_t row, int64_t kb, int term, int half,
                                            int64_t K2p) {
  return row * (V3_NT * K2p) + kb * (V3_NT * V3_BK) + term * V3_BK + 8 * half;
}

// A (M x K complex, as k_split_a) -> interleaved hi / mid / lo terms of A_r
template <typename TA>
__global__ __launch_bounds__(256) void k_split3_a(int64_t M, int64_t K, const TA* __restrict__ A,
                                                  int64_t a_m, int64_t a_k, int conj, int64_t bz,
                                                  int64_t K2p, int64_t pz, __bf16* __restrict__ P) {
  __shared__ float2 t[64][33];
  const int64_t m0 = (int64_t)blockIdx.y * 64, k0 = (int64_t)blockIdx.x * 32;
  A += blockIdx.z * bz;
  P += blockIdx.z * pz;
  const bool kc = (a_k == 1) || (a_m != 1);
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int r = kc ? e / 32 : e % 64, c = kc ? e % 32 : e / 64;
    const int64_t m = m0 + r, k = k0 + c;
    float re = 0.f, im = 0.f;
    if (m < M && k < K) load_c<float>(A + m * a_m + k * a_k, re, im);
    t[r][c] = make_float2(re, conj ? -im : im);
  }
  __syncthreads();
  // A_r row m: real k 2c + 0 / 1 = Re / Im of complex k c
  for (int q = threadIdx.x; q < 64 * V3_CHUNKS; q += 256) {
    const int r = q / V3_CHUNKS, w = q % V3_CHUNKS;
    const int kb = w / (2 * V3_NT), term = (w % (2 * V3_NT)) >> 1, half = w & 1;
    const int64_t kbg = 2 * k0 / V3_BK + kb;   // k-block in the row
    if (kbg * V3_BK >= K2p) continue;
    float x[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float2 v = t[r][8 * kb + 4 * half + j];
      x[2 * j] = v.x;
      x[2 * j + 1] = v.y;
    }
    *(bf16x8_t*)(P + v3_chunk(m0 + r, kbg, term, half, K2p)) = term8(x, term);
  }
}

// B (K x N complex, as k_split_b) -> interleaved terms of the real embedding
// B_e: row 2n holds (Re, -Im), row 2n + 1 holds (Im, Re) per complex k
template <typename TB>
__global__ __launch_bounds__(256) void k_split3_b(int64_t K, int64_t N, const TB* __restrict__ B,
                                                  int64_t b_k, int64_t b_n, int conj, int64_t bz,
                                                  int64_t K2p, int64_t pz, __bf16* __restrict__ P) {
  __shared__ float2 t[64][33];
  const int64_t n0 = (int64_t)blockIdx.y * 64, k0 = (int64_t)blockIdx.x * 32;
  B += blockIdx.z * bz;
  P += blockIdx.z * pz;
  const bool kc = (b_k == 1) || (b_n != 1);
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int r = kc ? e / 32 : e % 64, c = kc ? e % 32 : e / 64;
    const int64_t n = n0 + r, k = k0 + c;
    float re = 0.f, im = 0.f;
    if (n < N && k < K) load_c<float>(B + k * b_k + n * b_n, re, im);
    t[r][c] = make_float2(re, conj ? -im : im);
  }
  __syncthreads();
  // real rows 2 n (Re, -Im per complex k) and 2 n + 1 (Im, Re)
  for (int q = threadIdx.x; q < 128 * V3_CHUNKS; q += 256) {
    const int rr = q / V3_CHUNKS, w = q % V3_CHUNKS;
    const int kb = w / (2 * V3_NT), term = (w % (2 * V3_NT)) >> 1, half = w & 1;
    const int64_t kbg = 2 * k0 / V3_BK + kb;
    if (kbg * V3_BK >= K2p) continue;
    const int r = rr >> 1;
    const bool odd = rr & 1;
    float x[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float2 v = t[r][8 * kb + 4 * half + j];
      x[2 * j] = odd ? v.y : v.x;
      x[2 * j + 1] = odd ? v.x : -v.y;
    }
    *(bf16x8_t*)(P + v3_chunk(2 * n0 + rr, kbg, term, half, K2p)) = term8(x, term);
  }
}

// 256 x 256 real tiles, 8 waves (2 x 4) of 128 x 64 as in zgemm_bf16x3_v2,
// 16-deep k-stages double buffered in LDS (6 term planes x 256 rows x 24 bf16
// per stage = 72 KiB; the 48-byte row pitch keeps the 16-byte fragment reads
// of 16 consecutive rows on distinct banks). The operand pointers of V2Args
// are used as: ahi = A terms, bhi = B terms (alo, blo unused).
template <typename TC>
__global__ __launch_bounds__(512) void zgemm_bf16x6_v3(V2Args g) {
  __shared__ __align__(16) __bf16 sm[2][2 * V3_NT][V2_BM][V3_LD];
  if (g.skip && g.skip[(g.part ? 0 : blockIdx.z) * g.skip_stride]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n, tm, tn);
  const int64_t m0 = (int64_t)tm * V2_BM, c0 = (int64_t)tn * V2_BN;   // c0: real column
  const int64_t z = g.part ? 0 : blockIdx.z;
  const int64_t kbeg = g.part ? (int64_t)blockIdx.z * g.kchunk : 0;
  const int64_t pitch = V3_NT * g.K2p;   // bf16 per row
  const __bf16* srcA = g.ahi + z * g.pza + m0 * pitch + V3_NT * kbeg;
  const __bf16* srcB = g.bhi + z * g.pzb + c0 * pitch + V3_NT * kbeg;
  // global -> register staging: per operand and stage 256 rows x 96 bytes =
  // 1536 chunks of 16 bytes, three per thread (six consecutive threads per row)
  u32x4_t pa[3], pb[3];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int c = tid + 512 * u, row = c / 6, part = c % 6;
      const int64_t o = (int64_t)row * pitch + V3_NT * k0 + 8 * part;
      pa[u] = *(const u32x4_t*)(srcA + o);
      pb[u] = *(const u32x4_t*)(srcB + o);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int c = tid + 512 * u, row = c / 6, part = c % 6;
      const int term = part >> 1, half = 8 * (part & 1);
      *(u32x4_t*)&sm[buf][term][row][half] = pa[u];
      *(u32x4_t*)&sm[buf][V3_NT + term][row][half] = pb[u];
    }
  };
  f16v_t acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
  const int64_t klen = g.part ? (g.K2p - kbeg < g.kchunk ? g.K2p - kbeg : g.kchunk) : g.K2p;
  const int nst = (int)(klen / V3_BK);
  const int r32 = lane & 31, kg = 8 * (lane >> 5);
  gload(0);
  lstore(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int buf = st & 1;
    const bool more = st + 1 < nst;
    if (more) gload((int64_t)(st + 1) * V3_BK);
    bf16x8_t ah[4], am[4], al[4], bh[2], bm[2], bl[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = 128 * wm + 32 * i + r32;
      ah[i] = *(const bf16x8_t*)&sm[buf][0][row][kg];
      am[i] = *(const bf16x8_t*)&sm[buf][1][row][kg];
      al[i] = *(const bf16x8_t*)&sm[buf][2][row][kg];
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = 64 * wn + 32 * j + r32;
      bh[j] = *(const bf16x8_t*)&sm[buf][V3_NT][col][kg];
      bm[j] = *(const bf16x8_t*)&sm[buf][V3_NT + 1][col][kg];
      bl[j] = *(const bf16x8_t*)&sm[buf][V3_NT + 2][col][kg];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[i], bm[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bm[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    if (more) lstore(buf ^ 1);
    __syncthreads();
  }
  // epilogue of zgemm_bf16x3_v2 (the same accumulator layout)
  if (g.part) {
    float* P = g.part + (int64_t)blockIdx.z * (int64_t)g.tiles_m * V2_BM * g.N2p;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t c = c0 + 64 * wn + 32 * j + r32;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int64_t m = m0 + 128 * wm + 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
          P[m * g.N2p + c] = acc[i][j][v];
        }
      }
    return;
  }
  TC* C = (TC*)g.C + z * g.bz_c;
  const int pb_ = lane & 1;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t n = (c0 + 64 * wn + 32 * j + r32) >> 1;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float mine = acc[i][j][v];
        const float other = __shfl_xor(mine, 1);
        const int64_t m = m0 + 128 * wm + 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (!pb_ && m < g.M && n < g.N) {
          double re = mine, im = other;
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

template <typename TA, typename TB, typename TC>
inline int launch_bf16_v3(const ZGemm& z, hipStream_t s) {
  const int64_t Mp = v2_round(z.M, V2_BM), N2p = v2_round(2 * z.N, V2_BN), K2p = v3_kpad(z.K);
  const int64_t pza = Mp * K2p * V3_NT, pzb = N2p * K2p * V3_NT;
  char* base = (char*)(((uintptr_t)z.sws + 255) & ~(uintptr_t)255);
  __bf16* pa = (__bf16*)base;
  __bf16* pb = pa + z.batch * pza;
  const unsigned bt = (unsigned)z.batch;
  const unsigned kblk = (unsigned)((K2p + 63) / 64);   // 32 complex k per block
  k_split3_a<TA><<<dim3(kblk, (unsigned)(Mp / 64), bt), 256, 0, s>>>(
      z.M, z.K, (const TA*)z.A, z.a_m, z.a_k, z.conjA, z.bz_a, K2p, pza, pa);
  k_split3_b<TB><<<dim3(kblk, (unsigned)(N2p / 128), bt), 256, 0, s>>>(
      z.K, z.N, (const TB*)z.B, z.b_k, z.b_n, z.conjB, z.bz_b, K2p, pzb, pb);
  PTN_LAUNCH_CHECK();
  const int sk = v2_splitk(z.M, z.N, z.K, z.batch);
  V2Args g;
  g.ahi = pa; g.alo = nullptr; g.bhi = pb; g.blo = nullptr;
  g.K2p = K2p; g.pza = pza; g.pzb = pzb;
  g.N2p = N2p;
  g.kchunk = v2_round((K2p + sk - 1) / sk, V3_BK);
  g.part = nullptr;
  if (sk > 1) g.part = (float*)(((uintptr_t)(pb + z.batch * pzb) + 255) & ~(uintptr_t)255);
  g.M = z.M; g.N = z.N;
  g.tiles_m = (int)(Mp / V2_BM);
  g.tiles_n = (int)(N2p / V2_BN);
  g.C = z.C; g.c_m = z.c_m; g.c_n = z.c_n; g.bz_c = z.bz_c;
  g.accumulate = z.accumulate;
  g.skip = skip_flag().p;
  g.skip_stride = skip_flag().stride;
  const unsigned nz = sk > 1 ? (unsigned)((K2p + g.kchunk - 1) / g.kchunk) : bt;
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), 1, nz);
  static const std::string v3name = std::string("zgemm_bf16x6_v3<") + type_tag<TC>::s + ">";
  tlaunch(v3name.c_str(), gemm_flop(z), gemm_bytes<TA, TB, TC>(z), zgemm_bf16x6_v3<TC>, grid,
          dim3(512), s, g);
  if (sk > 1)
    k_v2_reduce<TC><<<(unsigned)std::min<int64_t>((z.M * z.N + 255) / 256, 4096), 256, 0, s>>>(
        z.M, z.N, Mp, N2p, (int)nz, g.part, (TC*)z.C, z.c_m, z.c_n, z.accumulate, g.skip);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

inline bool use_bf16_v3(const ZGemm& z, void* mws) {
  if (mws || !z.sws) return false;
  const int64_t tiles = ((z.M + V2_BM - 1) / V2_BM) * ((2 * z.N + V2_BN - 1) / V2_BN) * z.batch;
  if (tiles * v2_splitk(z.M, z.N, z.K, z.batch) < 128 || z.K < 16 || z.batch > 65535) return false;
  return z.sws_bytes >= v3_split_bytes(z.M, z.N, z.K, z.batch);
}

size_t bf16x6_split_bytes(int64_t M, int64_t N, int64_t K, int64_t batch) {
  return v3_split_bytes(M, N, K, batch);
}

int zgemm_fp32(const ZGemm& z, int ta, int tb, int tc, hipStream_t s, void* mws, size_t mws_bytes);

int zgemm_bf16x6(const ZGemm& z, int ta, int tb, int tc, hipStream_t s, void* mws,
                 size_t mws_bytes) {
  if (!use_bf16_v3(z, mws)) return zgemm_fp32(z, ta, tb, tc, s, mws, mws_bytes);
  const int key = (ta << 2) | (tb << 1) | tc;
  switch (key) {
    case 1: return launch_bf16_v3<double2, double2, float2>(z, s);
    case 2: return launch_bf16_v3<double2, float2, double2>(z, s);
    case 3: return launch_bf16_v3<double2, float2, float2>(z, s);
    case 4: return launch_bf16_v3<float2, double2, double2>(z, s);
    case 5: return launch_bf16_v3<float2, double2, float2>(z, s);
    case 6: return launch_bf16_v3<float2, float2, double2>(z, s);
    case 7: return launch_bf16_v3<float2, float2, float2>(z, s);
  }
  return PTN_ERR_UNSUPPORTED;
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
