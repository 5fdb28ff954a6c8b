// Complex matrix products on CDNA4 MFMA through the real embedding
//   [Re C, Im C] = [Re A, Im A] . [[Re B, Im B], [-Im B, Re B]]
// i.e. a complex (M x N x K) product is a real (M x 2N x 2K) product whose A
// operand is the interleaved complex row itself and whose B operand is
// expanded from (Re B, Im B) when it is read from LDS. This keeps the
// interleaved complex layout of torch/NumPy in HBM, needs no extra flops
// (8 real flops per complex multiply-add, as four real products would) and
// uses one accumulator set per output tile.
//
// MFMA: v_mfma_f64_16x16x4f64 (FP64) and v_mfma_f32_16x16x4f32 (FP32) take
// A[i][k] in lane i + 16k and B[k][j] in lane j + 16k; the 16x16 result has
// column lane%16 and, in accumulator element v, row lane/16 + 4v (FP64) or
// row 4*(lane/16) + v (FP32).
#pragma once
#include <mutex>
#include <stdio.h>
#include <string.h>
#include <string>
#include <unordered_map>
#include "common.h"

namespace ptn {

__device__ __forceinline__ d4_t mfma16(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename Tc> struct acc4;
template <> struct acc4<double> {
  typedef d4_t type;
  static __device__ __forceinline__ int row(int group, int v) { return group + 4 * v; }
};
template <> struct acc4<float> {
  typedef f4_t type;
  static __device__ __forceinline__ int row(int group, int v) { return 4 * group + v; }
};

// Workgroup = 4 waves as 2 x 2; each wave owns (16*TM) rows x (8*TN) complex columns.
// BK complex k per LDS stage (2 BK real k = BK / 2 MFMA k-steps); LDS holds two
// stages (double buffer).
template <int TM, int TN, int BK_ = 8> struct TileCfg {
  static constexpr int WM = 16 * TM;
  static constexpr int WN = 8 * TN;
  static constexpr int BM = 2 * WM;
  static constexpr int BN = 2 * WN;
  static constexpr int BK = BK_;
  static constexpr int NA = BM * BK / 256;
  static constexpr int NB = BN * BK / 256;
};
// FP32 stages twice as deep as FP64 ones: same LDS bytes per stage
template <typename Tc> struct StageK { static constexpr int value = 8; };
template <> struct StageK<float> { static constexpr int value = 16; };

struct GemmKArgs {
  int64_t M, N, K;
  const void* A; int64_t a_m, a_k; int conjA;
  const void* B; int64_t b_k, b_n; int conjB;
  void* C; int64_t c_m, c_n;
  int accumulate;
  int64_t bz_a, bz_b, bz_c;
  int tiles_n;          // number of column tiles
  int tiles_m;
  // optional block-sparsity masks (batch 1 only): bit t of word w of row tile
  // tm (column tile tn) is set when the A rows of tm (B columns of tn) have a
  // nonzero entry in k-stage 64 w + t; stages where the two masks do not
  // overlap contribute exactly zero and are skipped
  const uint64_t* maskA;
  const uint64_t* maskB;
  int mwords;
  // predication: the workgroups of batch problem z return at once when
  // skip[z * skip_stride] != 0
  const int* skip;
  int64_t skip_stride;
};

// masked products keep their stage list in LDS: K / BK at most this many stages
constexpr int MAX_MASKED_STAGES = 2048;

// first k-stage >= from whose A and B masks overlap (every stage when unmasked)
__device__ __forceinline__ int next_stage(const GemmKArgs& g, int tm, int tn, int from, int nk) {
  if (!g.maskA) return from;
  for (int w = from >> 6; w < g.mwords; ++w) {
    uint64_t bits = g.maskA[(int64_t)tm * g.mwords + w] & g.maskB[(int64_t)tn * g.mwords + w];
    if (w == (from >> 6)) bits &= ~0ull << (from & 63);
    if (bits) {
      const int t = (w << 6) + __builtin_ctzll(bits);
      return t < nk ? t : nk;
    }
  }
  return nk;
}

// Tile masks of a GEMM operand. The operand has `ntile_items` items (A rows or
// B columns) split into tiles of `tile`, and K split into k-stages of `bk`;
// element (item, k) is at p[item * s_item + k * s_k] (+ batch 0 only).
// One workgroup per (tile, 64-stage word); threads walk the block along its
// contiguous dimension and OR the stage bits of nonzero entries in LDS.
template <typename T>
__global__ __launch_bounds__(256) void tile_mask_kernel(int64_t nitems, int64_t K, const T* __restrict__ p,
                                                        int64_t s_item, int64_t s_k, int tile, int bk,
                                                        int mwords, uint64_t* __restrict__ mask) {
  __shared__ unsigned int bits[2];
  const int tix = blockIdx.x, w = blockIdx.y;
  if (threadIdx.x < 2) bits[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i0 = (int64_t)tix * tile;
  const int64_t k0 = (int64_t)w * 64 * bk;
  const int64_t ni = (i0 + tile < nitems ? i0 + tile : nitems) - i0;
  const int64_t nkk = (k0 + 64 * bk < K ? k0 + 64 * bk : K) - k0;
  const bool kfast = (s_k == 1) || (s_item != 1);
  unsigned int lo = 0, hi = 0;
  if (ni > 0 && nkk > 0) {
    for (int64_t e = threadIdx.x; e < ni * nkk; e += 256) {
      const int64_t it = kfast ? e / nkk : e % ni;
      const int64_t kk = kfast ? e % nkk : e / ni;
      const T v = p[(i0 + it) * s_item + (k0 + kk) * s_k];
      if (v.x != 0 || v.y != 0) {
        const int st = (int)(kk / bk);
        if (st < 32) lo |= 1u << st; else hi |= 1u << (st - 32);
      }
    }
  }
  if (lo) atomicOr(&bits[0], lo);
  if (hi) atomicOr(&bits[1], hi);
  __syncthreads();
  if (threadIdx.x == 0)
    mask[(int64_t)tix * mwords + w] = ((uint64_t)bits[1] << 32) | bits[0];
}

// bytes of mask workspace for an M x N x K product at any tile choice
inline size_t gemm_mask_bytes(int64_t M, int64_t N, int64_t K) {
  const int64_t words = ((K + 7) / 8 + 63) / 64;
  const int64_t tiles = (M + 63) / 64 + (N + 31) / 32;
  return (size_t)(tiles * words * 8 + 255) & ~(size_t)255;
}

// map the linear workgroup id so that the workgroups that run on one XCD
// (hardware assigns id % 8 to XCD id % 8) walk a contiguous range of tiles,
// column-tile fastest within a group of 8 row tiles, for L2 reuse of A and B.
__device__ __forceinline__ void tile_coords(int wg, int nwg, int tiles_m, int tiles_n, int& tm,
                                            int& tn) {
  int logical = wg;
  if ((nwg & 7) == 0) logical = (wg & 7) * (nwg >> 3) + (wg >> 3);
  const int group = 8;
  int per_group = group * tiles_n;
  int g = logical / per_group;
  int first_m = g * group;
  int rows = min(tiles_m - first_m, group);
  int in = logical - g * per_group;
  tm = first_m + in % rows;
  tn = in / rows;
}

template <typename Tc, typename TA, typename TB, typename TC, int TM, int TN, bool A_KCONTIG,
          bool B_NCONTIG>
__global__ __launch_bounds__(256) void zgemm_kernel(GemmKArgs g) {
  typedef TileCfg<TM, TN, StageK<Tc>::value> T;
  constexpr int BM = T::BM, BN = T::BN, BK = T::BK, NA = T::NA, NB = T::NB;
  typedef typename acc4<Tc>::type acc_t;

  __shared__ Tc As[2][2 * BK][BM + 1];
  __shared__ Tc Bs[2][BK][2 * BN + 2];
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
  const Tc sa = g.conjA ? (Tc)-1 : (Tc)1;
  const Tc sb = g.conjB ? (Tc)-1 : (Tc)1;

  Tc ra[NA][2], rb[NB][2];
  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      int e = tid + 256 * i;
      int row, kc;
      if (A_KCONTIG) { row = e / BK; kc = e % BK; } else { kc = e / BM; row = e % BM; }
      int64_t m = m0 + row, k = k0 + kc;
      Tc re = 0, im = 0;
      if (m < g.M && k < g.K) load_c<Tc>(A + m * g.a_m + k * g.a_k, re, im);
      ra[i][0] = re; ra[i][1] = sa * im;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      int e = tid + 256 * i;
      int col, kc;
      if (B_NCONTIG) { kc = e / BN; col = e % BN; } else { col = e / BK; kc = e % BK; }
      int64_t n = n0 + col, k = k0 + kc;
      Tc re = 0, im = 0;
      if (n < g.N && k < g.K) load_c<Tc>(B + k * g.b_k + n * g.b_n, re, im);
      rb[i][0] = re; rb[i][1] = sb * im;
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      int e = tid + 256 * i;
      int row, kc;
      if (A_KCONTIG) { row = e / BK; kc = e % BK; } else { kc = e / BM; row = e % BM; }
      As[buf][2 * kc][row] = ra[i][0];
      As[buf][2 * kc + 1][row] = ra[i][1];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      int e = tid + 256 * i;
      int col, kc;
      if (B_NCONTIG) { kc = e / BN; col = e % BN; } else { col = e / BK; kc = e % BK; }
      Bs[buf][kc][2 * col] = rb[i][0];
      Bs[buf][kc][2 * col + 1] = rb[i][1];
    }
  };

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // per-lane selection of the expanded B operand: real k parity (a) and real
  // column parity (b): (0,0) Re, (0,1) Im, (1,0) -Im, (1,1) Re
  const int kk = lane >> 4;
  const int pa = kk & 1, pb = lane & 1;
  const bool take_im = (pa != pb);
  const Tc im_sign = (pa == 1) ? (Tc)-1 : (Tc)1;

  // k-stages to visit: all of them, or (with masks) the list of stages where
  // the A and B tile masks overlap, built once in LDS so that the main loop is
  // a plain counted loop (keeps the accumulators in place across iterations)
  const int nk = (int)((g.K + BK - 1) / BK);
  __shared__ unsigned short stages[MAX_MASKED_STAGES];
  __shared__ int nstage_s;
  int nst = nk;
  if (g.maskA) {
    if (tid == 0) {
      int c = 0;
      for (int t = next_stage(g, tm, tn, 0, nk); t < nk; t = next_stage(g, tm, tn, t + 1, nk))
        stages[c++] = (unsigned short)t;
      nstage_s = c;
    }
    __syncthreads();
    nst = nstage_s;
  }
  auto stage_of = [&](int it) -> int { return g.maskA ? (int)stages[it] : it; };
  // stage pipeline: global loads of the next stage into registers, MFMAs on
  // the current LDS buffer, registers into the other buffer, one barrier
  int buf = 0;
  if (nst > 0) {
    load_tiles((int64_t)stage_of(0) * BK);
    store_tiles(0);
  }
  __syncthreads();
  for (int it = 0; it < nst; ++it) {
    const bool more = it + 1 < nst;
    if (more) load_tiles((int64_t)stage_of(it + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK / 2; ++ks) {
      Tc av[TM], bv[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) av[i] = As[buf][4 * ks + kk][wm0 + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const Tc* p = &Bs[buf][2 * ks + (kk >> 1)][2 * (wn0 + 8 * j + ((lane & 15) >> 1))];
        Tc re = p[0], im = p[1];
        bv[j] = take_im ? im_sign * im : re;
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = mfma16(av[i], bv[j], acc[i][j]);
    }
    if (more) store_tiles(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // epilogue: lane holds the real (pb = 0) or imaginary (pb = 1) part of
  // complex column n for the rows of its accumulator elements v; the partner
  // lane (lane ^ 1) holds the other part. Even lanes store the rows of
  // elements 0, 1 and odd lanes those of elements 2, 3.
#pragma unroll
  for (int i = 0; i < TM; ++i) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      Tc mine[4], other[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        mine[v] = acc[i][j][v];
        other[v] = __shfl_xor(mine[v], 1);
      }
      const int64_t n = n0 + wn0 + 8 * j + ((lane & 15) >> 1);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        // constant register indices only (a lane-dependent index would force scratch)
        const int v = 2 * pb + h;
        const int64_t m = m0 + wm0 + 16 * i + acc4<Tc>::row(kk, v);
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

// algorithmic work of one product launch: 8 real flops per complex
// multiply-add; bytes = each operand once, C written (and read if accumulating)
inline double gemm_flop(const ZGemm& z) {
  return 8.0 * (double)z.M * (double)z.N * (double)z.K * (double)z.batch;
}
template <typename TA, typename TB, typename TC>
inline double gemm_bytes(const ZGemm& z) {
  return (double)z.batch * ((double)z.M * z.K * sizeof(TA) + (double)z.K * z.N * sizeof(TB) +
                            (double)z.M * z.N * sizeof(TC) * (z.accumulate ? 2 : 1));
}
// kernel instance name as a profiler lists it (template arguments, without
// the namespace and the argument list), e.g.
// zgemm_kernel<float, c64, c64, c128, 4, 8, true, true>
template <typename Tc, typename TA, typename TB, typename TC, int TM, int TN>
inline const char* gemm_kernel_name(bool ak, bool bn, const char* kernel = "zgemm_kernel") {
  static std::mutex mu;
  static std::unordered_map<std::string, std::string> names;
  char buf[160];
  // the split-bf16 kernel has no compute-type argument
  const bool with_tc = strcmp(kernel, "zgemm_kernel") == 0;
  snprintf(buf, sizeof(buf), "%s<%s%s%s, %s, %s, %d, %d, %s, %s>", kernel,
           with_tc ? type_tag<Tc>::s : "", with_tc ? ", " : "", type_tag<TA>::s,
           type_tag<TB>::s, type_tag<TC>::s, TM, TN, ak ? "true" : "false",
           bn ? "true" : "false");
  std::lock_guard<std::mutex> lk(mu);
  auto it = names.find(buf);
  if (it == names.end()) it = names.emplace(buf, buf).first;
  return it->second.c_str();
}

// masks for the chosen tile shape (BM rows of A, BN columns of B, BK-wide
// k-stages) in `mws` (gemm_mask_bytes bytes), batch 1 only
template <typename TA, typename TB>
inline int build_masks(const ZGemm& z, GemmKArgs& g, int BM, int BN, int BK, void* mws,
                       size_t mws_bytes, hipStream_t s) {
  g.maskA = nullptr; g.maskB = nullptr; g.mwords = 0;
  if (!mws || z.batch != 1 || z.K == 0) return PTN_OK;
  const int64_t nk = (z.K + BK - 1) / BK;
  if (nk > MAX_MASKED_STAGES) return PTN_OK;   // dense product (correct, no skipping)
  const int words = (int)((nk + 63) / 64);
  const size_t need = (size_t)((int64_t)g.tiles_m + g.tiles_n) * words * 8;
  if (mws_bytes < need) return PTN_ERR_WORKSPACE;
  uint64_t* ma = (uint64_t*)mws;
  uint64_t* mb = ma + (int64_t)g.tiles_m * words;
  tile_mask_kernel<TA><<<dim3((unsigned)g.tiles_m, (unsigned)words), 256, 0, s>>>(
      z.M, z.K, (const TA*)z.A, z.a_m, z.a_k, BM, BK, words, ma);
  tile_mask_kernel<TB><<<dim3((unsigned)g.tiles_n, (unsigned)words), 256, 0, s>>>(
      z.N, z.K, (const TB*)z.B, z.b_n, z.b_k, BN, BK, words, mb);
  PTN_LAUNCH_CHECK();
  g.maskA = ma; g.maskB = mb; g.mwords = words;
  return PTN_OK;
}

template <typename Tc, typename TA, typename TB, typename TC, int TM, int TN>
inline int launch_cfg(const ZGemm& z, hipStream_t s, void* mws = nullptr, size_t mws_bytes = 0) {
  typedef TileCfg<TM, TN, StageK<Tc>::value> T;
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
  // a K-contiguous A when its k stride is 1 (or K == 1), else M-contiguous
  const bool ak = (z.a_k == 1) || (z.a_m != 1);
  const bool bn = (z.b_n == 1) || (z.b_k != 1);
  const char* nm = gemm_kernel_name<Tc, TA, TB, TC, TM, TN>(ak, bn);
  const double fl = gemm_flop(z), by = gemm_bytes<TA, TB, TC>(z);
  if (ak && bn)
    tlaunch(nm, fl, by, zgemm_kernel<Tc, TA, TB, TC, TM, TN, true, true>, grid, dim3(256), s, g);
  else if (ak && !bn)
    tlaunch(nm, fl, by, zgemm_kernel<Tc, TA, TB, TC, TM, TN, true, false>, grid, dim3(256), s, g);
  else if (!ak && bn)
    tlaunch(nm, fl, by, zgemm_kernel<Tc, TA, TB, TC, TM, TN, false, true>, grid, dim3(256), s, g);
  else
    tlaunch(nm, fl, by, zgemm_kernel<Tc, TA, TB, TC, TM, TN, false, false>, grid, dim3(256), s, g);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// opt-in autotuning of the tile shape (ptn_set_autotune): set once per process
bool autotune_enabled();

// tile choice: the large tile (128 x 64 complex) when it still gives at least
// two workgroups per CU, else the small one (64 x 32)
template <typename Tc, typename TA, typename TB, typename TC>
inline int launch_heuristic(const ZGemm& z, hipStream_t s, void* mws, size_t mws_bytes) {
  // FP32: 128 x 128 complex tiles (half the operand bytes per flop of 128 x 64)
  // when they still give two workgroups per CU
  if constexpr (sizeof(Tc) == 4) {
    int64_t wide = ((z.M + 127) / 128) * ((z.N + 127) / 128) * z.batch;
    if (wide >= 512) return launch_cfg<Tc, TA, TB, TC, 4, 8>(z, s, mws, mws_bytes);
  }
  int64_t big = ((z.M + 127) / 128) * ((z.N + 63) / 64) * z.batch;
  if (big >= 512) return launch_cfg<Tc, TA, TB, TC, 4, 4>(z, s, mws, mws_bytes);
  return launch_cfg<Tc, TA, TB, TC, 2, 2>(z, s, mws, mws_bytes);
}

template <typename Tc, typename TA, typename TB, typename TC>
inline int launch_choice(int c, const ZGemm& z, hipStream_t s) {
  if (c == 0) return launch_cfg<Tc, TA, TB, TC, 2, 2>(z, s);
  if (c == 1) return launch_cfg<Tc, TA, TB, TC, 4, 4>(z, s);
  if constexpr (sizeof(Tc) == 4) return launch_cfg<Tc, TA, TB, TC, 4, 8>(z, s);
  return PTN_ERR_ARG;
}

// Tile autotuner: the first product of a given shape (M, N, K, batch and the
// contiguity of the operands) times every tile configuration (two timed
// launches each after one untimed, HIP events on the caller's stream) and
// keeps the fastest for that shape. Only products that overwrite C are tuned
// (a repeated launch then leaves the same result); the tuning launch of the
// chosen configuration is the product itself. It synchronises the stream
// once per new shape, hence opt-in.
template <typename Tc, typename TA, typename TB, typename TC>
inline int launch_tuned(const ZGemm& z, hipStream_t s) {
  struct Key {
    int64_t M, N, K, batch, flags;
    bool operator==(const Key& o) const {
      return M == o.M && N == o.N && K == o.K && batch == o.batch && flags == o.flags;
    }
  };
  struct Hash {
    size_t operator()(const Key& k) const {
      return std::hash<int64_t>()(k.M * 1000003 ^ k.N * 10007 ^ k.K * 101 ^ k.batch * 7 ^ k.flags);
    }
  };
  static std::mutex mu;
  static std::unordered_map<Key, int, Hash> best;
  const Key key{z.M, z.N, z.K, z.batch,
                (int64_t)(z.a_k == 1) | (int64_t)(z.a_m == 1) << 1 | (int64_t)(z.b_n == 1) << 2 |
                    (int64_t)(z.b_k == 1) << 3};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = best.find(key);
    if (it != best.end()) return launch_choice<Tc, TA, TB, TC>(it->second, z, s);
  }
  const int nc = sizeof(Tc) == 4 ? 3 : 2;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return PTN_ERR_HIP;
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return PTN_ERR_HIP; }
  int bestc = 0, st = PTN_OK;
  float bestms = 0.0f;
  for (int c = 0; c < nc && st == PTN_OK; ++c) {
    if ((st = launch_choice<Tc, TA, TB, TC>(c, z, s))) break;
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < 2 && st == PTN_OK; ++r) st = launch_choice<Tc, TA, TB, TC>(c, z, s);
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) { st = PTN_ERR_HIP; break; }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (c == 0 || ms < bestms) { bestms = ms; bestc = c; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  {
    std::lock_guard<std::mutex> lk(mu);
    best[key] = bestc;
  }
  // C holds the last configuration's (identical) result; write it once more
  // with the chosen one so that the output never depends on tuning order
  return launch_choice<Tc, TA, TB, TC>(bestc, z, s);
}

template <typename Tc, typename TA, typename TB, typename TC>
inline int launch_typed(const ZGemm& z, hipStream_t s, void* mws = nullptr, size_t mws_bytes = 0) {
  // tile-skipping products (mask workspace given): the small tile skips the
  // zero blocks of sector-structured operands at finer granularity, which
  // outweighs its lower efficiency per flop (U(1) two-site DMRG at L=128,
  // D=2048: 76 s against 82 s per sweep)
  if (mws && z.batch == 1) return launch_cfg<Tc, TA, TB, TC, 2, 2>(z, s, mws, mws_bytes);
  if (!z.accumulate && autotune_enabled() && !noalloc_mode()) {
    // tuning synchronises the stream, which a stream capture does not allow
    hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cst) == hipSuccess && cst == hipStreamCaptureStatusNone)
      return launch_tuned<Tc, TA, TB, TC>(z, s);
  }
  return launch_heuristic<Tc, TA, TB, TC>(z, s, mws, mws_bytes);
}

}  // namespace ptn
