// Sector-grouped H_eff for U(1)-symmetric tensors: the two products of
// H_eff computed over the admissible quantum-number blocks only, each as ONE
// launch over all of its blocks.
//
// The tensors keep the reference's dense layout with zero blocks
// (pytenet/mps.py:71, pytenet/block_sparse_util.py:56-64). With quantum
// numbers ql (left virtual bond), qs (physical), qr (right virtual bond), qml /
// qmr (MPO virtual bonds) and the sign conventions of the reference's
// sparsity checks (pytenet/mps.py:71, pytenet/mpo.py:71,154,
// pytenet/tdvp.py:62) the nonzero entries are
//   A[l, s, r]          ql[l] + qs[s] = qr[r]
//   W[m, s', s, m']     qml[m] + qs[s'] = qs[s] + qmr[m']
//   R[r, m', r']        qr[r] + qmr[m'] = qr[r']
//   L[l, m, lb]         ql[l] + qml[m] = ql[lb]
// so that
//   T1[(l, s), (m', r')] = sum_r A[l, s, r] R[r, m', r'] is nonzero only where
//     ql[l] + qs[s] + qmr[m'] = qr[r'], and for fixed (s, m') and right
//     sector Q of r it is the dense block product
//       rows l in sector Q - qs[s], k = r in sector Q, columns r' in sector Q + qmr[m'];
//   out[lb, (s', r')] = sum_{(l, m)} L[l, m, lb] T2[(l, m), (s', r')] is nonzero only where
//     ql[lb] + qs[s'] = qr[r'], and for left sector P of lb and fixed s' it is
//       rows lb in P, columns r' in sector P + qs[s'], k over the segments
//       (l in sector P - qml[m], m), one segment per MPO index m.
// Every output block belongs to exactly one product block (its sector fixes
// the k-range), so no two workgroups write the same entry and no block needs
// accumulation. The entries of the output outside the blocks are structurally
// zero and are written by one fill pass; the sums inside a block run over
// exactly the reference's nonzero terms. The MPO mode mix between the products
// (k_mix_sectors) reads only the T1 entries the first product writes and writes
// only the T2 entries the second product reads.
//
// Launch: the grid covers the tiles of all blocks; a workgroup finds its block
// by binary search over the blocks' first tiles (descriptors in device memory,
// built once per distinct set of quantum numbers and kept in a cache).
// Requires the two virtual bonds to be sector-contiguous (each label in one
// run of indices), which the reference's QR and SVD splits produce
// (pytenet/block_sparse_util.py:106-180, 244-319); otherwise the tile-mask
// products are used.
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>
#include "gemm_impl.h"
#include "ops.h"

namespace ptn {

constexpr int GMAXSEG = 16;   // k-segments per block (MPO bond dimension at most 16)

struct GBlock {
  int64_t M, N;                 // block rows / columns (complex elements)
  int64_t c_off;                // element offset of the block in C
  int64_t tile0;                // first linear tile of the block in the launch
  int32_t tiles_n, nseg;
  int64_t a_off[GMAXSEG], b_off[GMAXSEG], K[GMAXSEG];
};

struct GLaunch {
  const GBlock* blocks;
  int nblocks;
  int64_t a_m, a_k, b_k, b_n, c_m, c_n;   // strides shared by all blocks
  const void* A;
  const void* B;
  void* C;
  const int* skip;
  int conjB;                              // B read conjugated
};

// One workgroup = one BM x BN tile of one block; the MFMA tile body is the one
// of zgemm_kernel (real embedding of the complex product, double-buffered LDS
// stages), with k running over the block's segments.
template <typename Tc, typename TA, typename TB, typename TC, int TM, int TN, bool A_KCONTIG,
          bool B_NCONTIG>
__global__ __launch_bounds__(256) void zgemm_grouped_kernel(GLaunch g) {
  typedef TileCfg<TM, TN, StageK<Tc>::value> T;
  constexpr int BM = T::BM, BN = T::BN, BK = T::BK, NA = T::NA, NB = T::NB;
  typedef typename acc4<Tc>::type acc_t;
  __shared__ Tc As[2][2 * BK][BM + 1];
  __shared__ Tc Bs[2][BK][2 * BN + 2];
  if (g.skip && g.skip[0]) return;
  // the block of this workgroup: the last block whose first tile <= blockIdx.x
  const int64_t tile = blockIdx.x;
  int lo = 0, hi = g.nblocks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g.blocks[mid].tile0 <= tile) lo = mid; else hi = mid - 1;
  }
  const GBlock* __restrict__ bk = g.blocks + lo;
  const int64_t bM = bk->M, bN = bk->N;
  const int nseg = bk->nseg;
  const int64_t lt = tile - bk->tile0;
  const int64_t m0 = (lt / bk->tiles_n) * BM, n0 = (lt % bk->tiles_n) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * T::WM, wn0 = (wave & 1) * T::WN;
  const TA* __restrict__ A = (const TA*)g.A;
  const TB* __restrict__ B = (const TB*)g.B;
  TC* __restrict__ C = (TC*)g.C + bk->c_off;
  const Tc sb = g.conjB ? (Tc)-1 : (Tc)1;
  int nst = 0;
  for (int sg = 0; sg < nseg; ++sg) nst += (int)((bk->K[sg] + BK - 1) / BK);

  Tc ra[NA][2], rb[NB][2];
  // stage `it` of the concatenated segments: segment sg, k offset k0 in it
  auto load_tiles = [&](int it) {
    int sg = 0;
    int64_t k0 = (int64_t)it * BK;
    for (; sg < nseg - 1; ++sg) {
      const int64_t ks = ((bk->K[sg] + BK - 1) / BK) * BK;
      if (k0 < ks) break;
      k0 -= ks;
    }
    const TA* Ap = A + bk->a_off[sg];
    const TB* Bp = B + bk->b_off[sg];
    const int64_t K = bk->K[sg];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = tid + 256 * i;
      int row, kc;
      if (A_KCONTIG) { row = e / BK; kc = e % BK; } else { kc = e / BM; row = e % BM; }
      const int64_t m = m0 + row, k = k0 + kc;
      Tc re = 0, im = 0;
      if (m < bM && k < K) load_c<Tc>(Ap + m * g.a_m + k * g.a_k, re, im);
      ra[i][0] = re; ra[i][1] = im;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int e = tid + 256 * i;
      int col, kc;
      if (B_NCONTIG) { kc = e / BN; col = e % BN; } else { col = e / BK; kc = e % BK; }
      const int64_t n = n0 + col, k = k0 + kc;
      Tc re = 0, im = 0;
      if (n < bN && k < K) load_c<Tc>(Bp + k * g.b_k + n * g.b_n, re, im);
      rb[i][0] = re; rb[i][1] = sb * im;
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = tid + 256 * i;
      int row, kc;
      if (A_KCONTIG) { row = e / BK; kc = e % BK; } else { kc = e / BM; row = e % BM; }
      As[buf][2 * kc][row] = ra[i][0];
      As[buf][2 * kc + 1][row] = ra[i][1];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int e = tid + 256 * i;
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
  // expanded B operand per lane: real k parity (pa), real column parity (pb)
  const int kk = lane >> 4;
  const int pa = kk & 1, pb = lane & 1;
  const bool take_im = (pa != pb);
  const Tc im_sign = (pa == 1) ? (Tc)-1 : (Tc)1;
  int buf = 0;
  if (nst > 0) {
    load_tiles(0);
    store_tiles(0);
  }
  __syncthreads();
  for (int it = 0; it < nst; ++it) {
    const bool more = it + 1 < nst;
    if (more) load_tiles(it + 1);
#pragma unroll
    for (int ks = 0; ks < BK / 2; ++ks) {
      Tc av[TM], bv[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) av[i] = As[buf][4 * ks + kk][wm0 + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const Tc* p = &Bs[buf][2 * ks + (kk >> 1)][2 * (wn0 + 8 * j + ((lane & 15) >> 1))];
        const Tc re = p[0], im = p[1];
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
  // epilogue as zgemm_kernel: the partner lane (lane ^ 1) holds the other part
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
        const int v = 2 * pb + h;
        const int64_t m = m0 + wm0 + 16 * i + acc4<Tc>::row(kk, v);
        const double re = pb ? (double)other[2 + h] : (double)mine[h];
        const double im = pb ? (double)mine[2 + h] : (double)other[h];
        if (m < bM && n < bN) store_c(C + m * g.c_m + n * g.c_n, re, im);
      }
    }
  }
}

// zero fill of n complex elements (the structurally zero entries of the output)
template <typename TC>
__global__ void k_zero_fill(int64_t n, TC* __restrict__ p, const int* __restrict__ skip) {
  if (skip && skip[0]) return;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    store_c(p + t, 0.0, 0.0);
}

template <typename TC>
static int zero_fill(int64_t n, void* p, hipStream_t s) {
  if (n <= 0) return PTN_OK;
  int64_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  tlaunch("k_zero_fill", 0.0, (double)n * sizeof(TC), k_zero_fill<TC>, dim3((unsigned)b),
          dim3(256), s, n, (TC*)p, skip_flag().p);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// ---------------------------------------------------------------------------
// Sector-aware MPO mode mix T2[l, o, r'] = sum_i Wm[o, i] T1[l, i, r'] with
// i = (s, m'), o = (m, s'). For a left sector P of l and a right sector Q of
// r' (a "unit"), T1[l, i, r'] is nonzero only for the inputs i with
// qs[s] + qmr[m'] = Q - P whose k-sector P + qs[s] exists (exactly the entries
// the first grouped product writes), and the second product reads T2[l, o, r']
// only for the outputs o with qml[m] + qs[s'] = Q - P whose output sector
// P + qml[m] exists. The mix reads and writes just those: T1 needs no zero
// fill and T2 no zeros outside what is read. One wave per row l and 64-column
// chunk of a unit, the lanes over the columns r' (coalesced). The kernel is
// bound by the bytes in flight per CU: a lane's inputs are held in registers
// sized to the plan's longest input list (CAP = 8, 16 or 32), and complex64
// operands are mixed in FP32 arithmetic (A = float).
constexpr int MIX_LIST = 32;   // d * Mr and Ml * d are at most 32
enum { PLAN_HEFF = SECTORS_HEFF, PLAN_ENV_RIGHT = SECTORS_ENV_RIGHT, PLAN_ENV_LEFT = SECTORS_ENV_LEFT };

struct MixUnit {
  int64_t l0, nl, r0, nr;
  int64_t tile0;                // first workgroup of the unit
  int32_t ni, no;
  int32_t chunks;               // 64-column chunks per row: a wave per (row, chunk)
  int8_t il[MIX_LIST], ol[MIX_LIST];
};

template <typename TX, typename TY, typename A, int CAP>
__global__ __launch_bounds__(256) void k_mix_sectors(const MixUnit* __restrict__ units, int nunits,
                                                     int ni_all, int no_all, int64_t nr_all,
                                                     const double2* __restrict__ wm,
                                                     const TX* __restrict__ x, TY* __restrict__ y,
                                                     const int* __restrict__ skip, int wtrans) {
  __shared__ double2 W[MIX_LIST * MIX_LIST];
  if (skip && skip[0]) return;
  for (int t = threadIdx.x; t < ni_all * no_all; t += blockDim.x) W[t] = wm[t];
  __syncthreads();
  const int64_t tile = blockIdx.x;
  int lo = 0, hi = nunits - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (units[mid].tile0 <= tile) lo = mid; else hi = mid - 1;
  }
  const MixUnit* __restrict__ u = units + lo;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // one wave per (row, 64-column chunk) of the unit: short independent work
  // items keep many loads in flight
  const int64_t wi = (tile - u->tile0) * 4 + wave;
  const int64_t row = wi / u->chunks;
  if (row >= u->nl) return;
  const int64_t r = u->r0 + (wi % u->chunks) * 64 + lane;
  if (r >= u->r0 + u->nr) return;
  const int64_t l = u->l0 + row;
  const int ni = u->ni, no = u->no;
  const TX* __restrict__ xl = x + l * ni_all * nr_all;
  TY* __restrict__ yl = y + l * no_all * nr_all;
  {
    A xr[CAP], xi[CAP];
#pragma unroll
    for (int k = 0; k < CAP; ++k) {
      xr[k] = 0; xi[k] = 0;
      if (k < ni) load_c<A>(xl + (int64_t)u->il[k] * nr_all + r, xr[k], xi[k]);
    }
    for (int j = 0; j < no; ++j) {
      const int o = u->ol[j];
      A yr = 0, yi = 0;
#pragma unroll
      for (int k = 0; k < CAP; ++k) {
        if (k < ni) {
          // Wm[o, i], or stored transposed as W[i, o] (left environment)
          const double2 w = wtrans ? W[u->il[k] * no_all + o] : W[o * ni_all + u->il[k]];
          const A wr = (A)w.x, wi = (A)w.y;
          yr += wr * xr[k] - wi * xi[k];
          yi += wr * xi[k] + wi * xr[k];
        }
      }
      store_c(yl + (int64_t)o * nr_all + r, yr, yi);
    }
  }
}

// ---------------------------------------------------------------------------
// plans: the blocks of both products for one set of quantum numbers

struct Sector { int64_t start, len; };

// runs of equal labels of a bond; false if a label has two runs
static bool sectors_of(const int64_t* q, int64_t n, std::map<int64_t, Sector>& out) {
  out.clear();
  for (int64_t i = 0; i < n;) {
    int64_t j = i;
    while (j < n && q[j] == q[i]) ++j;
    if (out.count(q[i])) return false;
    out[q[i]] = Sector{i, j - i};
    i = j;
  }
  return true;
}

struct GProduct {
  std::vector<GBlock> blocks;
  int64_t tiles = 0;
  int big = 0;          // 1: 128 x 64 tiles (TileCfg<4, 4>), 0: 64 x 32 (TileCfg<2, 2>)
  double flop = 0.0;    // 8 M N K summed over the blocks
  double bytes = 0.0;   // operand blocks read once, output blocks written once
};

struct SectorPlan {
  SectorPlan() = default;
  SectorPlan(const SectorPlan&) = delete;
  SectorPlan& operator=(const SectorPlan&) = delete;
  // the descriptors are freed with the plan's last reference (hipFree waits
  // for the device, so no launch still reads them)
  ~SectorPlan() {
    if (dev) {
      int cur = 0;
      if (hipGetDevice(&cur) == hipSuccess && cur != device) (void)hipSetDevice(device);
      (void)hipFree(dev);
      if (cur != device) (void)hipSetDevice(cur);
    }
  }
  int kind = PLAN_HEFF;
  GProduct p1, p3;
  std::vector<MixUnit> mix;
  int64_t mix_tiles = 0;
  double mix_bytes = 0.0;   // T1 entries read and T2 entries written by the mix (elements)
  int mix_cap = MIX_LIST;   // register capacity for a unit's inputs (8, 16 or 32)
  GBlock* dev = nullptr;    // p1's blocks, then p3's, then the mix units
  MixUnit* dev_mix = nullptr;
  int device = 0;
  uint64_t stamp = 0;
};

// tile shape and first tiles of a product: the 128 x 64 tile when its padding
// costs at most 15 % more work than the 64 x 32 tile's and it still gives
// enough workgroups for the chip (>= 512)
static void finish_product(GProduct& P, size_t csz_a, size_t csz_b, size_t csz_c) {
  double w_small = 0.0, w_big = 0.0;
  int64_t t_big = 0;
  P.flop = 0.0;
  P.bytes = 0.0;
  for (const GBlock& b : P.blocks) {
    double k = 0.0;
    for (int i = 0; i < b.nseg; ++i) k += (double)b.K[i];
    P.flop += 8.0 * (double)b.M * (double)b.N * k;
    P.bytes += k * ((double)b.M * csz_a + (double)b.N * csz_b) + (double)b.M * b.N * csz_c;
    w_small += (double)((b.M + 63) / 64 * 64) * (double)((b.N + 31) / 32 * 32) * k;
    w_big += (double)((b.M + 127) / 128 * 128) * (double)((b.N + 63) / 64 * 64) * k;
    t_big += ((b.M + 127) / 128) * ((b.N + 63) / 64);
  }
  P.big = (w_big <= 1.15 * w_small && t_big >= 512) ? 1 : 0;
  const int64_t BM = P.big ? 128 : 64, BN = P.big ? 64 : 32;
  int64_t tiles = 0;
  for (GBlock& b : P.blocks) {
    b.tiles_n = (int32_t)((b.N + BN - 1) / BN);
    b.tile0 = tiles;
    tiles += ((b.M + BM - 1) / BM) * b.tiles_n;
  }
  P.tiles = tiles;
}

using SectorMap = std::map<int64_t, Sector>;

static GBlock block(int64_t M, int64_t N, int64_t c_off) {
  GBlock b{};
  b.M = M;
  b.N = N;
  b.c_off = c_off;
  return b;
}
static void add_segment(GBlock& b, int64_t a_off, int64_t b_off, int64_t K) {
  b.a_off[b.nseg] = a_off;
  b.b_off[b.nseg] = b_off;
  b.K[b.nseg] = K;
  ++b.nseg;
}

// T1[(l, s), (m', r')] = sum_r A[(l, s), r] R[r, (m', r')] (H_eff and the
// right environment step): block per (s, right sector Q of r, m')
static void plan_a_times_r(const Dims& D, const int64_t* qs, const int64_t* qmr,
                           const SectorMap& sl, const SectorMap& sr, GProduct& P) {
  for (int64_t s = 0; s < D.d; ++s)
    for (const auto& kq : sr) {
      const auto itl = sl.find(kq.first - qs[s]);
      if (itl == sl.end()) continue;
      for (int64_t mp = 0; mp < D.Mr; ++mp) {
        const auto itc = sr.find(kq.first + qmr[mp]);
        if (itc == sr.end()) continue;
        const int64_t l0 = itl->second.start, r0 = kq.second.start, c0 = itc->second.start;
        GBlock b = block(itl->second.len, itc->second.len,
                         (l0 * D.d + s) * (D.Mr * D.Dr) + mp * D.Dr + c0);
        add_segment(b, (l0 * D.d + s) * D.Dr + r0, r0 * (D.Mr * D.Dr) + mp * D.Dr + c0,
                    kq.second.len);
        P.blocks.push_back(b);
      }
    }
}

// H_eff: out[lb, (s', r')] = sum_{(l, m)} L[(l, m), lb] T2[(l, m), (s', r')]:
// block per (left sector P of lb, s'), one k-segment per MPO index m
static void plan_heff_out(const Dims& D, const int64_t* qs, const int64_t* qml,
                          const SectorMap& sl, const SectorMap& sr, GProduct& P) {
  for (const auto& pq : sl) {
    const int64_t lb0 = pq.second.start;
    for (int64_t s2 = 0; s2 < D.d; ++s2) {
      const auto itc = sr.find(pq.first + qs[s2]);
      if (itc == sr.end()) continue;
      const int64_t c0 = itc->second.start;
      GBlock b = block(pq.second.len, itc->second.len, lb0 * (D.d * D.Dr) + s2 * D.Dr + c0);
      for (int64_t m = 0; m < D.Ml; ++m) {
        const auto itl = sl.find(pq.first - qml[m]);
        if (itl == sl.end()) continue;
        const int64_t l0 = itl->second.start;
        add_segment(b, (l0 * D.Ml + m) * D.Dl + lb0, (l0 * D.Ml + m) * (D.d * D.Dr) + s2 * D.Dr + c0,
                    itl->second.len);
      }
      if (b.nseg) P.blocks.push_back(b);   // no segment: a structurally zero block
    }
  }
}

// right environment: out[(l, m), lb] = sum_{(s', r')} T2[(l, m), (s', r')] conj(B[lb, (s', r')]):
// block per (left sector P of l, m), one k-segment per physical index s'
static void plan_envr_out(const Dims& D, const int64_t* qs, const int64_t* qml,
                          const SectorMap& sl, const SectorMap& sr, GProduct& P) {
  for (const auto& pq : sl) {
    const int64_t l0 = pq.second.start;
    for (int64_t m = 0; m < D.Ml; ++m) {
      const auto itb = sl.find(pq.first + qml[m]);
      if (itb == sl.end()) continue;
      const int64_t lb0 = itb->second.start;
      GBlock b = block(pq.second.len, itb->second.len, (l0 * D.Ml + m) * D.Dl + lb0);
      for (int64_t s2 = 0; s2 < D.d; ++s2) {
        const auto itr = sr.find(itb->first + qs[s2]);
        if (itr == sr.end()) continue;
        const int64_t rb0 = itr->second.start;
        add_segment(b, (l0 * D.Ml + m) * (D.d * D.Dr) + s2 * D.Dr + rb0,
                    lb0 * (D.d * D.Dr) + s2 * D.Dr + rb0, itr->second.len);
      }
      if (b.nseg) P.blocks.push_back(b);
    }
  }
}

// left environment: T1[(l, m), (s', r')] = sum_lb L[(l, m), lb] conj(B[lb, (s', r')]):
// block per (left sector P of l, m, s')
static void plan_l_times_b(const Dims& D, const int64_t* qs, const int64_t* qml,
                           const SectorMap& sl, const SectorMap& sr, GProduct& P) {
  for (const auto& pq : sl) {
    const int64_t l0 = pq.second.start;
    for (int64_t m = 0; m < D.Ml; ++m) {
      const auto itb = sl.find(pq.first + qml[m]);
      if (itb == sl.end()) continue;
      const int64_t lb0 = itb->second.start;
      for (int64_t s2 = 0; s2 < D.d; ++s2) {
        const auto itr = sr.find(itb->first + qs[s2]);
        if (itr == sr.end()) continue;
        const int64_t rb0 = itr->second.start;
        GBlock b = block(pq.second.len, itr->second.len,
                         (l0 * D.Ml + m) * (D.d * D.Dr) + s2 * D.Dr + rb0);
        add_segment(b, (l0 * D.Ml + m) * D.Dl + lb0, lb0 * (D.d * D.Dr) + s2 * D.Dr + rb0,
                    itb->second.len);
        P.blocks.push_back(b);
      }
    }
  }
}

// left environment: out[r, (m', r')] = sum_{(l, s)} A[(l, s), r] T2[(l, s), (m', r')]:
// block per (right sector Q of r, m'), one k-segment per physical index s
static void plan_envl_out(const Dims& D, const int64_t* qs, const int64_t* qmr,
                          const SectorMap& sl, const SectorMap& sr, GProduct& P) {
  for (const auto& rq : sr) {
    const int64_t r0 = rq.second.start;
    for (int64_t mp = 0; mp < D.Mr; ++mp) {
      const auto itc = sr.find(rq.first + qmr[mp]);
      if (itc == sr.end()) continue;
      const int64_t c0 = itc->second.start;
      GBlock b = block(rq.second.len, itc->second.len, r0 * (D.Mr * D.Dr) + mp * D.Dr + c0);
      for (int64_t s = 0; s < D.d; ++s) {
        const auto itl = sl.find(rq.first - qs[s]);
        if (itl == sl.end()) continue;
        const int64_t l0 = itl->second.start;
        add_segment(b, (l0 * D.d + s) * D.Dr + r0, (l0 * D.d + s) * (D.Mr * D.Dr) + mp * D.Dr + c0,
                    itl->second.len);
      }
      if (b.nseg) P.blocks.push_back(b);
    }
  }
}

// mix units for a unit (left sector P, right sector Q, dq = Q - P):
// H_eff / right environment: inputs i = (s, m') with qs[s] + qmr[m'] = dq and
// k-sector P + qs[s] present, outputs o = (m, s') with qml[m] + qs[s'] = dq and
// P + qml[m] present; left environment: inputs i = (m, s') with
// qml[m] + qs[s'] = dq and P + qml[m] present, outputs o = (s, m') with
// qs[s] + qmr[m'] = dq and P + qs[s] present.
static void plan_mix(const Dims& D, bool left, const int64_t* qs, const int64_t* qml,
                     const int64_t* qmr, const SectorMap& sl, const SectorMap& sr,
                     SectorPlan& sp) {
  for (const auto& pq : sl)
    for (const auto& rq : sr) {
      const int64_t P = pq.first, dq = rq.first - P;
      MixUnit u{};
      int8_t* sm_list = left ? u.il : u.ol;    // (m, s') indices
      int8_t* ms_list = left ? u.ol : u.il;    // (s, m') indices
      int32_t& n_sm = left ? u.ni : u.no;
      int32_t& n_ms = left ? u.no : u.ni;
      for (int64_t m = 0; m < D.Ml; ++m)
        for (int64_t s2 = 0; s2 < D.d; ++s2)
          if (qml[m] + qs[s2] == dq && sl.count(P + qml[m])) sm_list[n_sm++] = (int8_t)(m * D.d + s2);
      for (int64_t s1 = 0; s1 < D.d; ++s1)
        for (int64_t mp = 0; mp < D.Mr; ++mp)
          if (qs[s1] + qmr[mp] == dq && sr.count(P + qs[s1])) ms_list[n_ms++] = (int8_t)(s1 * D.Mr + mp);
      if (u.no == 0) continue;   // nothing of this unit is read
      u.l0 = pq.second.start;
      u.nl = pq.second.len;
      u.r0 = rq.second.start;
      u.nr = rq.second.len;
      u.tile0 = sp.mix_tiles;
      u.chunks = (int32_t)((u.nr + 63) / 64);
      sp.mix_tiles += (u.nl * u.chunks + 3) / 4;
      sp.mix_bytes += (double)u.nl * (double)u.nr * (double)(u.ni + u.no);
      sp.mix.push_back(u);
    }
}

// blocks of both products and the mix units of a plan kind; false when a bond
// is not sector-contiguous, a segment or mix list would exceed its limit, or
// the launch would be too large
static bool build_plan(int kind, const Dims& D, const int64_t* ql, const int64_t* qs,
                       const int64_t* qr, const int64_t* qml, const int64_t* qmr, size_t csz,
                       SectorPlan& sp) {
  if (D.Dl <= 0 || D.Dr <= 0 || D.d <= 0 || D.Ml <= 0 || D.Mr <= 0 ||
      D.d * D.Mr > MIX_LIST || D.Ml * D.d > MIX_LIST)
    return false;
  if ((kind == PLAN_HEFF && D.Ml > GMAXSEG) || (kind != PLAN_HEFF && D.d > GMAXSEG)) return false;
  SectorMap sl, sr;
  if (!sectors_of(ql, D.Dl, sl) || !sectors_of(qr, D.Dr, sr)) return false;
  sp.kind = kind;
  if (kind == PLAN_ENV_LEFT) {
    plan_l_times_b(D, qs, qml, sl, sr, sp.p1);
    plan_mix(D, true, qs, qml, qmr, sl, sr, sp);
    plan_envl_out(D, qs, qmr, sl, sr, sp.p3);
    finish_product(sp.p1, csz, 16, csz);   // L in the compute type, B complex128
    finish_product(sp.p3, 16, csz, 16);    // A and out complex128
  } else {
    plan_a_times_r(D, qs, qmr, sl, sr, sp.p1);
    plan_mix(D, false, qs, qml, qmr, sl, sr, sp);
    if (kind == PLAN_HEFF) {
      plan_heff_out(D, qs, qml, sl, sr, sp.p3);
      finish_product(sp.p3, csz, csz, 16);
    } else {
      plan_envr_out(D, qs, qml, sl, sr, sp.p3);
      finish_product(sp.p3, csz, 16, 16);
    }
    finish_product(sp.p1, 16, csz, csz);   // A complex128, R and T1 in the compute type
  }
  int maxni = 0;
  for (const MixUnit& u : sp.mix) maxni = u.ni > maxni ? u.ni : maxni;
  sp.mix_cap = maxni <= 8 ? 8 : (maxni <= 16 ? 16 : MIX_LIST);
  return sp.p1.tiles > 0 && sp.p3.tiles > 0 && sp.p1.tiles < 0x7fffffff &&
         sp.p3.tiles < 0x7fffffff && !sp.mix.empty() && sp.mix_tiles < 0x7fffffff;
}

// arithmetic of the grouped products for a precision code: FP64 or FP32
// (the three-term split-bf16 mode has FP32 accuracy and takes the FP32
// products); -1 when they do not apply (two-term split-bf16: the tile-mask
// products)
static int grouped_prec(int prec) {
  const int bp = prec & ~PTN_BLOCK_SPARSE;
  if (bp == PTN_FP64) return PTN_FP64;
  if (bp == PTN_FP32 || bp == PTN_BF16X6) return PTN_FP32;
  return -1;
}

namespace {
std::mutex g_plan_mu;
// nullptr: not applicable. Allocated once and never destroyed, so that no
// plan is released (hipFree) during static destruction at process exit, when
// the HIP runtime may already be gone.
auto& g_plans = *new std::unordered_map<std::string, std::shared_ptr<SectorPlan>>();
uint64_t g_plan_clock = 0;
constexpr size_t PLAN_CACHE_MAX = 1024;
}  // namespace

static void append(std::string& k, const int64_t* p, int64_t n) {
  k.append((const char*)&n, sizeof(n));
  k.append((const char*)p, (size_t)n * sizeof(int64_t));
}

// The sector plan of kind PLAN_* for a site operator with quantum numbers
// (null when the operator has none, is not a site operator, or the grouped
// products do not apply). Builds and uploads it on first use (synchronous
// copy: call it outside stream capture); plans are cached per device, kind
// and set of quantum numbers.
int sector_plan_get(const ptn_local_op* op, int kind, SectorPlanRef* out) {
  out->reset();
  if (!op || op->kind != PTN_OP_HEFF || !(op->precision & PTN_BLOCK_SPARSE) || !op->q_l ||
      !op->q_s || !op->q_r || !op->q_ml || !op->q_mr)
    return PTN_OK;
  const int bp = grouped_prec(op->precision);
  if (bp < 0) return PTN_OK;
  if (noalloc_mode()) return PTN_OK;
  int dev = 0;
  PTN_HIP_CHECK(hipGetDevice(&dev));
  std::string key;
  const int64_t head[8] = {dev, bp, kind, op->Dl, op->d, op->Dr, op->Ml, op->Mr};
  key.append((const char*)head, sizeof(head));
  append(key, op->q_l, op->Dl);
  append(key, op->q_s, op->d);
  append(key, op->q_r, op->Dr);
  append(key, op->q_ml, op->Ml);
  append(key, op->q_mr, op->Mr);
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    if (it->second) it->second->stamp = ++g_plan_clock;
    *out = it->second;
    return PTN_OK;
  }
  if (g_plans.size() >= PLAN_CACHE_MAX) {
    // drop the older half of this device's plans from the cache (a plan still
    // held by a caller lives on until that caller releases it)
    std::vector<std::pair<uint64_t, std::string>> age;
    for (auto& kv : g_plans)
      if (!kv.second || kv.second->device == dev)
        age.emplace_back(kv.second ? kv.second->stamp : 0, kv.first);
    std::sort(age.begin(), age.end());
    for (size_t i = 0; i < age.size() / 2; ++i) g_plans.erase(age[i].second);
  }
  std::shared_ptr<SectorPlan> sp(new SectorPlan());
  const Dims D = {op->Dl, op->d, op->Dr, op->Ml, op->Mr};
  if (!build_plan(kind, D, op->q_l, op->q_s, op->q_r, op->q_ml, op->q_mr,
                  bp == PTN_FP64 ? 16 : 8, *sp)) {
    g_plans.emplace(key, nullptr);
    return PTN_OK;
  }
  std::vector<GBlock> all(sp->p1.blocks);
  all.insert(all.end(), sp->p3.blocks.begin(), sp->p3.blocks.end());
  const size_t bb = (all.size() * sizeof(GBlock) + 255) & ~(size_t)255;
  const size_t mb = sp->mix.size() * sizeof(MixUnit);
  hipError_t e = hipMalloc((void**)&sp->dev, bb + mb);
  if (e == hipSuccess) e = hipMemcpy(sp->dev, all.data(), all.size() * sizeof(GBlock),
                                     hipMemcpyHostToDevice);
  sp->dev_mix = (MixUnit*)((char*)sp->dev + bb);
  if (e == hipSuccess) e = hipMemcpy(sp->dev_mix, sp->mix.data(), mb, hipMemcpyHostToDevice);
  sp->device = dev;
  if (e != hipSuccess) {
    ptn_set_last_hip_error(e);
    return PTN_ERR_HIP;   // the plan (and its allocation) is released here
  }
  sp->stamp = ++g_plan_clock;
  *out = sp;
  g_plans.emplace(key, std::move(sp));
  return PTN_OK;
}

template <typename Tc, typename TA, typename TB, typename TC, bool AK, bool BNC>
static int grouped_launch(const GLaunch& g, const GProduct& P, hipStream_t s) {
  char name[112];
  snprintf(name, sizeof(name), "zgemm_grouped_kernel<%s, %s, %s, %s, %d, %d, %s, %s>",
           type_tag<Tc>::s, type_tag<TA>::s, type_tag<TB>::s, type_tag<TC>::s, P.big ? 4 : 2,
           P.big ? 4 : 2, AK ? "true" : "false", BNC ? "true" : "false");
  // launch records keep the name pointer: intern the names
  static std::mutex mu;
  static std::unordered_map<std::string, std::string> names;
  const char* nm;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = names.find(name);
    if (it == names.end()) it = names.emplace(name, name).first;
    nm = it->second.c_str();
  }
  if (P.big)
    tlaunch(nm, P.flop, P.bytes, zgemm_grouped_kernel<Tc, TA, TB, TC, 4, 4, AK, BNC>,
            dim3((unsigned)P.tiles), dim3(256), s, g);
  else
    tlaunch(nm, P.flop, P.bytes, zgemm_grouped_kernel<Tc, TA, TB, TC, 2, 2, AK, BNC>,
            dim3((unsigned)P.tiles), dim3(256), s, g);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// the sector mix of a plan: T2 (y) from T1 (x) with the MPO matrix w (complex128),
// complex128 operands in FP64 or complex64 in FP32 arithmetic; wtrans: w stored
// transposed (left environment)
template <int CAP>
static void mix_launch_cap(const SectorPlan* sp, bool fp64, int ni, int no, int64_t nr,
                           const void* w, const void* x, void* y, const int* skip, int wtrans,
                           hipStream_t s) {
  const dim3 grid((unsigned)sp->mix_tiles);
  const MixUnit* units = sp->dev_mix;
  const int nu = (int)sp->mix.size();
  if (fp64)
    tlaunch("k_mix_sectors<c128, c128>", 0.0, sp->mix_bytes * 16,
            k_mix_sectors<double2, double2, double, CAP>, grid, dim3(256), s, units, nu, ni, no,
            nr, (const double2*)w, (const double2*)x, (double2*)y, skip, wtrans);
  else
    tlaunch("k_mix_sectors<c64, c64>", 0.0, sp->mix_bytes * 8,
            k_mix_sectors<float2, float2, float, CAP>, grid, dim3(256), s, units, nu, ni, no, nr,
            (const double2*)w, (const float2*)x, (float2*)y, skip, wtrans);
}

static int mix_launch(const SectorPlan* sp, bool fp64, int ni, int no, int64_t nr, const void* w,
                      const void* x, void* y, const int* skip, int wtrans, hipStream_t s) {
  if (sp->mix_cap <= 8)
    mix_launch_cap<8>(sp, fp64, ni, no, nr, w, x, y, skip, wtrans, s);
  else if (sp->mix_cap <= 16)
    mix_launch_cap<16>(sp, fp64, ni, no, nr, w, x, y, skip, wtrans, s);
  else
    mix_launch_cap<MIX_LIST>(sp, fp64, ni, no, nr, w, x, y, skip, wtrans, s);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// H_eff through the grouped products: out = L^T . (W (x) (A . R)) as
// heff_apply computes it (same operand types: A and out complex128, L and R
// complex128 in FP64 and complex64 in FP32), only the blocks computed and out
// zero-filled outside them. Uses T1 and T2 of the heff_apply workspace.
int heff_apply_sectors(int prec, const Dims& D, const SectorPlan* sp, const void* l, int tl,
                       const void* w, const void* r, int tr, const void* a, void* out, void* ws,
                       size_t ws_bytes, hipStream_t s) {
  const int bp = grouped_prec(prec);
  if (!sp || bp < 0) return PTN_ERR_UNSUPPORTED;
  const int ct = bp == PTN_FP64 ? PTN_T_C128 : PTN_T_C64;
  if (tl != ct || tr != ct) return PTN_ERR_UNSUPPORTED;
  if (ws_bytes < heff_ws_bytes(prec, D)) return PTN_ERR_WORKSPACE;
  const size_t csz = bp == PTN_FP64 ? 16 : 8;
  const int64_t n1 = (D.Dl * D.d) * (D.Mr * D.Dr), nout = D.Dl * D.d * D.Dr;
  void* t1 = ws;   // written by the first product only inside its blocks (the mix reads no more)
  void* t2 = (char*)ws + (((size_t)n1 * csz + 255) & ~(size_t)255);
  const GBlock* b1 = sp->dev;
  const GBlock* b3 = sp->dev + sp->p1.blocks.size();
  const int* skip = skip_flag().p;
  int st;
  GLaunch g1{b1, (int)sp->p1.blocks.size(), D.d * D.Dr, 1, D.Mr * D.Dr, 1, D.d * D.Mr * D.Dr, 1,
             a, r, t1, skip, 0};
  const int ni = (int)(D.d * D.Mr), no = (int)(D.Ml * D.d);
  if (bp == PTN_FP64) {
    if ((st = grouped_launch<double, double2, double2, double2, true, true>(g1, sp->p1, s)))
      return st;
    if ((st = mix_launch(sp, true, ni, no, D.Dr, w, t1, t2, skip, 0, s))) return st;
  } else {
    if ((st = grouped_launch<float, double2, float2, float2, true, true>(g1, sp->p1, s)))
      return st;
    if ((st = mix_launch(sp, false, ni, no, D.Dr, w, t1, t2, skip, 0, s))) return st;
  }
  PTN_LAUNCH_CHECK();
  if ((st = zero_fill<double2>(nout, out, s))) return st;
  GLaunch g3{b3, (int)sp->p3.blocks.size(), 1, D.Ml * D.Dl, D.Ml * D.d * D.Dr, 1, D.d * D.Dr, 1,
             l, t2, out, skip, 0};
  if (bp == PTN_FP64)
    return grouped_launch<double, double2, double2, double2, false, true>(g3, sp->p3, s);
  return grouped_launch<float, float2, float2, double2, false, true>(g3, sp->p3, s);
}

// Right environment step through the grouped products (env_right's chain:
// T1 = A . R, the mode mix, out = T2 . B^H), only the blocks computed and out
// zero-filled outside them. R has type tr (complex128 in FP64, complex64 in
// FP32); A, B and out are complex128. Uses T1 and T2 of the env workspace.
int env_right_sectors(int prec, const Dims& D, const SectorPlan* sp, const void* a, const void* b,
                      const void* w, const void* r, int tr, void* out, void* ws, size_t ws_bytes,
                      hipStream_t s) {
  const int bp = grouped_prec(prec);
  if (!sp || sp->kind != PLAN_ENV_RIGHT || bp < 0) return PTN_ERR_UNSUPPORTED;
  const int ct = bp == PTN_FP64 ? PTN_T_C128 : PTN_T_C64;
  if (tr != ct) return PTN_ERR_UNSUPPORTED;
  if (ws_bytes < env_ws_bytes(prec, D)) return PTN_ERR_WORKSPACE;
  const size_t csz = bp == PTN_FP64 ? 16 : 8;
  const size_t big = std::max((size_t)(D.Dl * D.d) * (size_t)(D.Mr * D.Dr),
                              (size_t)(D.Dl * D.Ml) * (size_t)(D.d * D.Dr)) * csz;
  void* t1 = ws;
  void* t2 = (char*)ws + ((big + 255) & ~(size_t)255);
  const GBlock* b1 = sp->dev;
  const GBlock* b3 = sp->dev + sp->p1.blocks.size();
  const int* skip = skip_flag().p;
  const int ni = (int)(D.d * D.Mr), no = (int)(D.Ml * D.d);
  int st;
  GLaunch g1{b1, (int)sp->p1.blocks.size(), D.d * D.Dr, 1, D.Mr * D.Dr, 1, D.d * D.Mr * D.Dr, 1,
             a, r, t1, skip, 0};
  // out[(l, m), lb] = sum T2[(l, m), (s', r')] conj(B[lb, (s', r')])
  GLaunch g3{b3, (int)sp->p3.blocks.size(), D.Ml * D.d * D.Dr, 1, 1, D.d * D.Dr, D.Ml * D.Dl, 1,
             t2, b, out, skip, 1};
  if (bp == PTN_FP64) {
    if ((st = grouped_launch<double, double2, double2, double2, true, true>(g1, sp->p1, s)))
      return st;
    if ((st = mix_launch(sp, true, ni, no, D.Dr, w, t1, t2, skip, 0, s))) return st;
    if ((st = zero_fill<double2>(D.Dl * D.Ml * D.Dl, out, s))) return st;
    return grouped_launch<double, double2, double2, double2, true, false>(g3, sp->p3, s);
  }
  if ((st = grouped_launch<float, double2, float2, float2, true, true>(g1, sp->p1, s))) return st;
  if ((st = mix_launch(sp, false, ni, no, D.Dr, w, t1, t2, skip, 0, s))) return st;
  if ((st = zero_fill<double2>(D.Dl * D.Ml * D.Dl, out, s))) return st;
  return grouped_launch<float, float2, double2, double2, true, false>(g3, sp->p3, s);
}

// Left environment step through the grouped products (env_left's chain:
// T1 = L . B*, the mode mix with the transposed MPO tensor, out = A^T . T2).
// L has type tl; A, B and out are complex128.
int env_left_sectors(int prec, const Dims& D, const SectorPlan* sp, const void* a, const void* b,
                     const void* w, const void* l, int tl, void* out, void* ws, size_t ws_bytes,
                     hipStream_t s) {
  const int bp = grouped_prec(prec);
  if (!sp || sp->kind != PLAN_ENV_LEFT || bp < 0) return PTN_ERR_UNSUPPORTED;
  const int ct = bp == PTN_FP64 ? PTN_T_C128 : PTN_T_C64;
  if (tl != ct) return PTN_ERR_UNSUPPORTED;
  if (ws_bytes < env_ws_bytes(prec, D)) return PTN_ERR_WORKSPACE;
  const size_t csz = bp == PTN_FP64 ? 16 : 8;
  const size_t big = std::max((size_t)(D.Dl * D.d) * (size_t)(D.Mr * D.Dr),
                              (size_t)(D.Dl * D.Ml) * (size_t)(D.d * D.Dr)) * csz;
  void* t1 = ws;
  void* t2 = (char*)ws + ((big + 255) & ~(size_t)255);
  const GBlock* b1 = sp->dev;
  const GBlock* b3 = sp->dev + sp->p1.blocks.size();
  const int* skip = skip_flag().p;
  const int ni = (int)(D.Ml * D.d), no = (int)(D.d * D.Mr);
  int st;
  // T1[(l, m), (s', r')] = sum_lb L[(l, m), lb] conj(B[lb, (s', r')])
  GLaunch g1{b1, (int)sp->p1.blocks.size(), D.Ml * D.Dl, 1, D.d * D.Dr, 1, D.Ml * D.d * D.Dr, 1,
             l, b, t1, skip, 1};
  // out[r, (m', r')] = sum_{(l, s)} A[(l, s), r] T2[(l, s), (m', r')]
  GLaunch g3{b3, (int)sp->p3.blocks.size(), 1, D.d * D.Dr, D.d * D.Mr * D.Dr, 1, D.Mr * D.Dr, 1,
             a, t2, out, skip, 0};
  if (bp == PTN_FP64) {
    if ((st = grouped_launch<double, double2, double2, double2, true, true>(g1, sp->p1, s)))
      return st;
    if ((st = mix_launch(sp, true, ni, no, D.Dr, w, t1, t2, skip, 1, s))) return st;
    if ((st = zero_fill<double2>(D.Dr * D.Mr * D.Dr, out, s))) return st;
    return grouped_launch<double, double2, double2, double2, false, true>(g3, sp->p3, s);
  }
  if ((st = grouped_launch<float, float2, double2, float2, true, true>(g1, sp->p1, s))) return st;
  if ((st = mix_launch(sp, false, ni, no, D.Dr, w, t1, t2, skip, 1, s))) return st;
  if ((st = zero_fill<double2>(D.Dr * D.Mr * D.Dr, out, s))) return st;
  return grouped_launch<float, double2, float2, double2, false, true>(g3, sp->p3, s);
}

}  // namespace ptn
