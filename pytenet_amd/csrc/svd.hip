// Thin SVD of a complex128 matrix: QR preconditioning followed by blocked
// one-sided (Hestenes) Jacobi, all in FP64.
//
// With B = A (m >= n) or B = A^H (m < n), a p x q matrix with p >= q:
//   1. B = Q1 R by the blocked Householder QR of qr.hip (Q1 kept as reflectors);
//   2. R^H = Q2 R2 (second QR), and X = R2^H (q x q, column-major, padded with
//      zero columns to a multiple of 2*JB) is orthogonalised by one-sided
//      Jacobi, X V = U_X diag(sigma); starting from R2^H instead of B cuts the
//      number of sweeps several-fold for graded spectra (the two-sided QR
//      preconditioning of the Jacobi SVD of Drmac & Veselic);
//   3. B = Q1 R2^H Q2^H = (Q1 [U_X; 0]) diag(sigma) (Q2 V)^H: U_B and V_B are
//      formed by applying the reflectors of the two factorisations in compact
//      WY blocks.
// Jacobi: columns are grouped in blocks of JB; a round of the block tournament
// pairs every block with another and each block pair (2*JB columns) is
// orthogonalised by
//   k_jac_gram:  partial Gram matrices G = Y^H Y of the pair's columns over row
//                segments (FP64 MFMA);
//   k_jac_solve: one cyclic sweep of two-sided Jacobi on G in LDS (parallel
//                ordering, JB disjoint rotations per step), accumulating the unitary J;
//   k_jac_apply: Y <- Y J and V <- V J (FP64 MFMA).
// Sweeps repeat until a sweep in which no column pair needed a rotation by
// more than 1e-8 relative (|x_i^H x_j| <= 1e-8 |x_i| |x_j| when the pair was
// processed); quadratic convergence leaves couplings of order 1e-16 after it. Then sigma_j = |x_j| and
// U_X = X / sigma. One-sided Jacobi computes small singular values to high
// relative accuracy, which the discarded-weight truncation relies on.
//
// Plane rotation for columns (x, y) with a = |x|^2, b = |y|^2, c = x^H y = |c| e^{i phi}:
//   zeta = (b - a) / (2|c|), t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
//   cs = 1 / sqrt(1 + t^2), s = cs t e^{i phi},
//   x' = cs x - conj(s) y,   y' = s x + cs y.
#include <float.h>
#include "common.h"
#include "qr.h"

namespace ptn {

constexpr int JB = 16;          // columns per block
constexpr int JP = 2 * JB;      // columns per block pair
constexpr int GSEG = 128;       // rows per Gram workgroup (one partial Gram matrix each)
constexpr int GCH = 64;         // rows per LDS chunk in the Gram kernel
constexpr int AROWS = 64;       // rows per apply workgroup (a multiple of 64; 16-row tiles per wave)

__device__ double wg_sum3(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  double s = red[0];
  __syncthreads();
  return s;
}

// blockIdx.y selects the matrix of a batch (q x q each)
__global__ void k_eye(int64_t q, double2* __restrict__ V) {
  V += (int64_t)blockIdx.y * q * q;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < q * q;
       t += (int64_t)gridDim.x * blockDim.x)
    V[t] = make_double2((t / q) == (t % q) ? 1.0 : 0.0, 0.0);
}

// pair k of round r of the round-robin (circle) tournament over ne players
__device__ __forceinline__ void round_pair(int ne, int r, int k, int& p, int& q) {
  const int cyc = ne - 1;
  if (k == 0) { p = r; q = cyc; }
  else { p = (r + k) % cyc; q = (r - k + cyc) % cyc; }
  if (p > q) { int t = p; p = q; q = t; }
}

// column index (in X / V) of local column c (0..JP-1) of the pair (bp, bq)
__device__ __forceinline__ int64_t pair_col(int bp, int bq, int c) {
  return (c < JB) ? (int64_t)bp * JB + c : (int64_t)bq * JB + (c - JB);
}

// element types of the two Jacobi precisions: complex storage, MFMA accumulator
template <typename T> struct JacT;
template <> struct JacT<double> { using c2 = double2; using acc = d4_t; };
template <> struct JacT<float> { using c2 = float2; using acc = f4_t; };
__device__ __forceinline__ d4_t jmfma(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t jmfma(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// row of accumulator element v in lane group kk = lane / 16 of a 16 x 16 MFMA
// result: kk + 4 v for v_mfma_f64_16x16x4f64, 4 kk + v for v_mfma_f32_16x16x4f32
template <typename T> __device__ __forceinline__ int acc_row(int kk, int v) {
  return sizeof(T) == 8 ? kk + 4 * v : 4 * kk + v;
}

// partial Gram matrices Gp[pair][seg][i][j] = sum_{rows in seg} conj(Y[row][i]) Y[row][j]
// on MFMA (v_mfma_f64_16x16x4f64, or v_mfma_f32_16x16x4f32 for the FP32
// preconditioning sweeps) through the real embedding of the complex product
// with A = Y^H (real k = 2 row + part):
//   A_real[i][2r] = Re Y[r][i], A_real[i][2r+1] = -Im Y[r][i];
//   B_real[2r + q][2j + p] = (Re, Im; -Im, Re)[q][p] of Y[r][j].
// The 32 x 64 real result is 8 tiles of 16 x 16; wave w owns row tile w & 1 and
// column tiles 2 (w >> 1) and 2 (w >> 1) + 1. Rows are staged through LDS in
// chunks of GCH. blockIdx.z selects the problem of a batch (X at X + z sx).
static_assert(JP == 32, "the Gram and apply kernels are laid out for 32-column block pairs");
template <typename T>
__global__ __launch_bounds__(256) void k_jac_gram(int64_t p, int nblk, int r,
                                                  const typename JacT<T>::c2* __restrict__ X,
                                                  typename JacT<T>::c2* __restrict__ Gp, int nseg,
                                                  int64_t sx) {
  using C2 = typename JacT<T>::c2;
  using Acc = typename JacT<T>::acc;
  __shared__ C2 Y[GCH][JP + 1];
  X += (int64_t)blockIdx.z * sx;
  int bp, bq;
  round_pair(nblk, r, blockIdx.y, bp, bq);
  const int seg = blockIdx.x;
  const int64_t r0 = (int64_t)seg * GSEG;
  const int64_t r1 = (r0 + GSEG < p) ? r0 + GSEG : p;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, kk = lane >> 4;
  const int i0 = 16 * (wave & 1), jt0 = 2 * (wave >> 1);
  const int q = kk & 1, pb = l16 & 1;
  Acc acc[2] = {Acc{0, 0, 0, 0}, Acc{0, 0, 0, 0}};
  // the next chunk is loaded into registers while the current one is multiplied
  constexpr int NL = GCH * JP / 256;
  C2 pre[NL];
  auto load_chunk = [&](int64_t c0) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int e = t + 256 * u;
      const int col = e / GCH, row = e % GCH;
      const int64_t grow = c0 + row;
      pre[u] = (grow < r1) ? X[pair_col(bp, bq, col) * p + grow] : C2{0, 0};
    }
  };
  load_chunk(r0);
  for (int64_t c0 = r0; c0 < r1; c0 += GCH) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int e = t + 256 * u;
      Y[e % GCH][e / GCH] = pre[u];
    }
    __syncthreads();
    if (c0 + GCH < r1) load_chunk(c0 + GCH);
#pragma unroll 4
    for (int s = 0; s < GCH / 2; ++s) {
      const int row = 2 * s + (kk >> 1);
      const C2 ya = Y[row][i0 + l16];
      const T a = q ? -ya.y : ya.x;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const C2 yb = Y[row][8 * (jt0 + h) + (l16 >> 1)];
        const T b = (q == pb) ? yb.x : (q ? -yb.y : yb.y);
        acc[h] = jmfma(a, b, acc[h]);
      }
    }
  }
  // result element v of tile h: complex row i0 + acc_row(kk, v), real column l16
  // of the tile (complex column 8 (jt0 + h) + l16 / 2, real part for even l16);
  // the partner lane l16 ^ 1 holds the other part
  C2* out = Gp + (((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * nseg + seg) * JP * JP;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    T mine[4], other[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      mine[v] = acc[h][v];
      other[v] = __shfl_xor(mine[v], 1);
    }
    const int j = 8 * (jt0 + h) + (l16 >> 1);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      // even lanes store elements 0, 1 and odd lanes elements 2, 3 (constant indices)
      const int i = i0 + acc_row<T>(kk, 2 * pb + u);
      const T re = pb ? other[2 + u] : mine[u];
      const T im = pb ? mine[2 + u] : other[u];
      out[i * JP + j] = C2{re, im};
    }
  }
}

// rotation parameters of the index pair (i, j) of a Hermitian Gram matrix with
// diagonal entries a, b and off-diagonal entry c = G[i][j]: identity unless
// |c| > tol sqrt(a b) (compared in squares, the Gram entries already hold
// squared column scales); one reciprocal square root gives 1/|c|. Branch-free
// (a skipped rotation computes on a harmless stand-in and selects the identity).
__device__ __forceinline__ void jac_rotation(double a, double b, double2 c, double tol,
                                             double big, double& cs, double& sr, double& si,
                                             bool& rot, bool& rbig) {
  const double tol2 = tol * tol, big2 = big * big;
  const double ac2 = c.x * c.x + c.y * c.y;
  const double ab = fmax(a, 0.0) * fmax(b, 0.0);
  rot = ac2 > 0.0 && ac2 > tol2 * ab;
  rbig = rot && ac2 > big2 * ab;
  const double iac = rsqrt(rot ? ac2 : 1.0);
  const double zeta = (b - a) * (0.5 * iac);
  const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double c0 = rsqrt(1.0 + tt * tt);
  const double sn = c0 * tt;
  cs = rot ? c0 : 1.0;
  sr = rot ? sn * (c.x * iac) : 0.0;   // s = sn e^{i phi}
  si = rot ? sn * (c.y * iac) : 0.0;
}
// FP32 preconditioning sweeps: the same rotation from the hardware reciprocal,
// reciprocal square root and square root (about 1 ulp each; the sweeps only
// need rotations that are unitary to FP32 rounding). |c| is formed from c
// scaled by mc = max(|Re c|, |Im c|), so no square underflows into the
// denormals that the hardware reciprocal square root flushes to zero: Gram
// entries down to 1e-36 (columns scaled to unit maximum norm, so column norms
// down to about 1e-18) are rotated, smaller couplings are left to the FP64
// sweeps. A zeta whose square overflows gives t = 0 (no rotation; |c| is then
// negligible against |a - b|).
__device__ __forceinline__ void jac_rotation(float a, float b, float2 c, float tol, float big,
                                             float& cs, float& sr, float& si, bool& rot,
                                             bool& rbig) {
  const float mc0 = fmaxf(fabsf(c.x), fabsf(c.y));
  const bool nz = mc0 > 1e-36f;
  const float imc = __builtin_amdgcn_rcpf(nz ? mc0 : 1.0f);
  const float cx = c.x * imc, cy = c.y * imc;   // max(|cx|, |cy|) = 1 when nz
  const float r2 = nz ? cx * cx + cy * cy : 1.0f;
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float ac = mc0 * r2 * ir;   // |c|
  const float sab = __builtin_amdgcn_sqrtf(fmaxf(a, 0.0f)) * __builtin_amdgcn_sqrtf(fmaxf(b, 0.0f));
  rot = nz && ac > tol * sab;
  rbig = rot && ac > big * sab;
  const float iac = imc * ir;
  const float zeta = (b - a) * (0.5f * iac);
  const float tt = copysignf(1.0f, zeta) *
                   __builtin_amdgcn_rcpf(fabsf(zeta) + __builtin_amdgcn_sqrtf(1.0f + zeta * zeta));
  const float c0 = __builtin_amdgcn_rsqf(1.0f + tt * tt);
  const float sn = c0 * tt;
  cs = rot ? c0 : 1.0f;
  sr = rot ? sn * (cx * ir) : 0.0f;   // s = sn e^{i phi}, e^{i phi} = c / |c|
  si = rot ? sn * (cy * ir) : 0.0f;
}

// players of pair k at step 0 of the round-robin over 32 (round_pair with
// r = 0); from one step to the next both advance by one modulo 31, except the
// fixed player 31 of pair 0
__device__ __forceinline__ void rr_start(int k, int& p, int& q) {
  p = k;
  q = (k == 0) ? 31 : 31 - k;
}
__device__ __forceinline__ void rr_next(int k, int& p, int& q) {
  p = (p == 30) ? 0 : p + 1;
  if (k != 0) q = (q == 30) ? 0 : q + 1;
}

// one cyclic Jacobi sweep on the pair's Gram matrix (sum of the partials,
// fixed order); writes the accumulated unitary J and whether any rotation
// was needed (one workgroup per pair). A step of the sweep rotates JB
// disjoint index pairs (i_k, j_k) at once: with R_k = [[cs, s], [-conj(s), cs]]
// acting on columns (i_k, j_k), G <- R^H G R is done as independent 2 x 2
// block updates (rows of pair k, columns of pair l; one thread per (k, l))
// and J <- J R as independent column-pair updates. Lane l of a wave computes
// the rotation of pair l % JB (the column pair of its block) and takes that of
// its row pair from lane k of the wave by a shuffle; G is read from one buffer
// and the rotated G written to the other, so a step needs a single barrier.
// `counters` counts the pairs with a rotation above `big`.
static_assert(JB * JB == 256 && JB == 16, "the solve maps one thread to each (pair, pair) block of G");
template <typename T>
__global__ __launch_bounds__(256) void k_jac_solve(int nblk, const typename JacT<T>::c2* __restrict__ Gp,
                                                   int nseg, T tol, T big,
                                                   typename JacT<T>::c2* __restrict__ Jg,
                                                   int* __restrict__ rotflag,
                                                   int* __restrict__ counters) {
  using C2 = typename JacT<T>::c2;
  __shared__ C2 G[2][JP][JP + 1];
  __shared__ C2 J[JP][JP + 1];
  __shared__ int rot_any, rot_big;
  const int t = threadIdx.x;
  // pair blockIdx.x of problem blockIdx.y
  const int64_t pidx = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
  const C2* gsrc = Gp + pidx * nseg * JP * JP;
  {
    // sum of the partial Gram matrices in a fixed order; the four elements of a
    // thread and several partials are loaded before they are added
    constexpr int NE = JP * JP / 256;
    T sr[NE], si[NE];
#pragma unroll
    for (int u = 0; u < NE; ++u) { sr[u] = 0; si[u] = 0; }
#pragma unroll 8
    for (int s = 0; s < nseg; ++s) {
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        const C2 v = gsrc[(int64_t)s * JP * JP + t + 256 * u];
        sr[u] += v.x;
        si[u] += v.y;
      }
    }
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = t + 256 * u;
      G[0][e / JP][e % JP] = C2{sr[u], si[u]};
      J[e / JP][e % JP] = C2{(e / JP) == (e % JP) ? T(1) : T(0), T(0)};
    }
  }
  if (t == 0) { rot_any = 0; rot_big = 0; }
  __syncthreads();
  const int bk = t / JB, bl = t % JB;     // the (row pair, column pair) block of this thread
  int pk, qk, pl, ql;
  rr_start(bk, pk, qk);
  rr_start(bl, pl, ql);
  bool any = false, anybig = false;
  for (int step = 0; step < JP - 1; ++step) {
    const int cur = step & 1;
    const int ik = min(pk, qk), jk = max(pk, qk), il = min(pl, ql), jl = max(pl, ql);
    const C2 b00 = G[cur][ik][il], b01 = G[cur][ik][jl], b10 = G[cur][jk][il],
             b11 = G[cur][jk][jl];
    const T gaa = G[cur][il][il].x, gbb = G[cur][jl][jl].x;
    const C2 gab = G[cur][il][jl];
    const C2 x0 = J[2 * bk][il], y0 = J[2 * bk][jl];
    const C2 x1 = J[2 * bk + 1][il], y1 = J[2 * bk + 1][jl];
    T cl, slr, sli;
    bool rl, bgl;
    jac_rotation(gaa, gbb, gab, tol, big, cl, slr, sli, rl, bgl);
    any |= rl;
    anybig |= bgl;
    // the row pair's rotation, computed by lane bk (< 16) of this wave
    const T ck = __shfl(cl, bk), skr = __shfl(slr, bk), ski = __shfl(sli, bk);
    {
      // columns: x' = cl x - conj(sl) y, y' = sl x + cl y (x = column il, y = column jl)
      C2 c00, c01, c10, c11;
      c00 = C2{cl * b00.x - (slr * b01.x + sli * b01.y), cl * b00.y - (slr * b01.y - sli * b01.x)};
      c10 = C2{cl * b10.x - (slr * b11.x + sli * b11.y), cl * b10.y - (slr * b11.y - sli * b11.x)};
      c01 = C2{(slr * b00.x - sli * b00.y) + cl * b01.x, (slr * b00.y + sli * b00.x) + cl * b01.y};
      c11 = C2{(slr * b10.x - sli * b10.y) + cl * b11.x, (slr * b10.y + sli * b10.x) + cl * b11.y};
      // rows: row_i' = ck row_i - sk row_j, row_j' = conj(sk) row_i + ck row_j
      C2 (*Gn)[JP + 1] = G[cur ^ 1];
      Gn[ik][il] = C2{ck * c00.x - (skr * c10.x - ski * c10.y), ck * c00.y - (skr * c10.y + ski * c10.x)};
      Gn[ik][jl] = C2{ck * c01.x - (skr * c11.x - ski * c11.y), ck * c01.y - (skr * c11.y + ski * c11.x)};
      Gn[jk][il] = C2{(skr * c00.x + ski * c00.y) + ck * c10.x, (skr * c00.y - ski * c00.x) + ck * c10.y};
      Gn[jk][jl] = C2{(skr * c01.x + ski * c01.y) + ck * c11.x, (skr * c01.y - ski * c01.x) + ck * c11.y};
    }
    // J <- J R: column pair bl for rows 2*bk and 2*bk + 1
    J[2 * bk][il] = C2{cl * x0.x - (slr * y0.x + sli * y0.y), cl * x0.y - (slr * y0.y - sli * y0.x)};
    J[2 * bk][jl] = C2{(slr * x0.x - sli * x0.y) + cl * y0.x, (slr * x0.y + sli * x0.x) + cl * y0.y};
    J[2 * bk + 1][il] = C2{cl * x1.x - (slr * y1.x + sli * y1.y), cl * x1.y - (slr * y1.y - sli * y1.x)};
    J[2 * bk + 1][jl] = C2{(slr * x1.x - sli * x1.y) + cl * y1.x, (slr * x1.y + sli * x1.x) + cl * y1.y};
    rr_next(bk, pk, qk);
    rr_next(bl, pl, ql);
    __syncthreads();
  }
  if (any) rot_any = 1;
  if (anybig) rot_big = 1;
  __syncthreads();
  C2* jout = Jg + pidx * JP * JP;
  for (int e = t; e < JP * JP; e += 256) jout[e] = J[e / JP][e % JP];
  if (t == 0) {
    rotflag[pidx] = rot_any;
    if (rot_big) atomicAdd(counters, 1);
  }
}

// Y <- Y J (pass 0, rows of X) or V <- V J (pass 1) for AROWS rows per
// workgroup, 16 per wave, on MFMA: the rows' 32 complex entries are
// the A operand (real k = 2 column + part, the interleaved complex layout) and
// J's real embedding [[Re, Im], [-Im, Re]] the B operand, read from LDS.
// Every wave loads all A operands of its rows before it stores any result.
// nmat = 2: problem blockIdx.z >> 1 of a batch, pass blockIdx.z & 1 (X and V);
// nmat = 1: problem blockIdx.z, X only (the rotations are not accumulated).
// X (p x qp) and V (qp x qp) stacked per problem.
template <typename T>
__global__ __launch_bounds__(256) void k_jac_apply(int64_t p, int64_t qp, int nblk, int r,
                                                   typename JacT<T>::c2* __restrict__ X,
                                                   typename JacT<T>::c2* __restrict__ V,
                                                   const typename JacT<T>::c2* __restrict__ Jg,
                                                   const int* __restrict__ rotflag, int nmat) {
  using C2 = typename JacT<T>::c2;
  using Acc = typename JacT<T>::acc;
  __shared__ C2 J[JP][JP + 1];
  const int64_t prob = nmat == 2 ? (blockIdx.z >> 1) : blockIdx.z;
  const int64_t pidx = prob * gridDim.y + blockIdx.y;
  const int pass = nmat == 2 ? (blockIdx.z & 1) : 0;
  C2* M = pass ? V + prob * qp * qp : X + prob * p * qp;
  const int64_t ld = pass ? qp : p;
  const int64_t rowblk = (int64_t)blockIdx.x * AROWS;
  if (rowblk >= ld) return;
  const int t = threadIdx.x;
  // the rotation and the pair's flag are loaded together (J is in bounds
  // whether or not the pair rotated; it is used only if it did)
  const C2* jsrc = Jg + pidx * JP * JP;
  constexpr int NJ = JP * JP / 256;
  C2 jv[NJ];
#pragma unroll
  for (int u = 0; u < NJ; ++u) jv[u] = jsrc[t + 256 * u];
  if (!rotflag[pidx]) return;
  int bp, bq;
  round_pair(nblk, r, blockIdx.y, bp, bq);
#pragma unroll
  for (int u = 0; u < NJ; ++u) {
    const int e = t + 256 * u;
    J[e / JP][e % JP] = jv[u];
  }
  __syncthreads();
  const int lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, kk = lane >> 4;
  const int q = kk & 1, pb = l16 & 1;
  constexpr int RT = AROWS / 64;   // 16-row tiles per wave
#pragma unroll 1
  for (int rt = 0; rt < RT; ++rt) {
    const int64_t row0 = rowblk + 16 * (RT * wave + rt);
    if (row0 >= ld) break;
    const int64_t arow = row0 + l16;
    const bool aok = arow < ld;
    // each lane loads only the part (real or imaginary) it feeds to the MFMA;
    // the wave's two lanes of an element read its two adjacent halves
    const T* __restrict__ Mt = reinterpret_cast<const T*>(M);
    T a[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int kc = 2 * s + (kk >> 1);
      a[s] = aok ? Mt[2 * (pair_col(bp, bq, kc) * ld + arow) + q] : T(0);
    }
    Acc acc[4];
#pragma unroll
    for (int tile = 0; tile < 4; ++tile) acc[tile] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int kc = 2 * s + (kk >> 1);
#pragma unroll
      for (int tile = 0; tile < 4; ++tile) {
        const C2 jv = J[kc][8 * tile + (l16 >> 1)];
        const T b = (q == pb) ? jv.x : (q ? -jv.y : jv.y);
        acc[tile] = jmfma(a[s], b, acc[tile]);
      }
    }
#pragma unroll
    for (int tile = 0; tile < 4; ++tile) {
      T mine[4], other[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        mine[v] = acc[tile][v];
        other[v] = __shfl_xor(mine[v], 1);
      }
      C2* col = M + pair_col(bp, bq, 8 * tile + (l16 >> 1)) * ld;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int64_t row = row0 + acc_row<T>(kk, 2 * pb + u);
        const T re = pb ? other[2 + u] : mine[u];
        const T im = pb ? mine[2 + u] : other[u];
        if (row < ld) col[row] = C2{re, im};
      }
    }
  }
}

// column norms of X (problem blockIdx.y of a batch at X + y sx, nrm + y sn)
__global__ __launch_bounds__(256) void k_colnorms(int64_t pr, const double2* __restrict__ X,
                                                  double* __restrict__ nrm, int64_t sx,
                                                  int64_t sn) {
  __shared__ double red[256];
  const double2* x = X + blockIdx.y * sx + (int64_t)blockIdx.x * pr;
  nrm += blockIdx.y * sn;
  double s = 0;
  for (int64_t i = threadIdx.x; i < pr; i += 256) {
    double2 u = x[i];
    s += u.x * u.x + u.y * u.y;
  }
  s = wg_sum3(s, red);
  if (threadIdx.x == 0) nrm[blockIdx.x] = sqrt(s);
}

// descending order of the q column norms (stable: ties keep column order);
// rank of column j = #{i : s_i > s_j} + #{i < j : s_i == s_j}, one thread per
// column j (blockIdx.x selects a chunk of 256 columns, blockIdx.y the problem;
// nrm and perm with stride sn, s_out with stride ss). A non-finite norm is
// ordered as -1 (norms are >= 0) so that the ranks stay a permutation; the
// flag reports it. The norms are those of the input times scl[problem].
__global__ __launch_bounds__(256) void k_rank(int64_t q, const double* __restrict__ nrm,
                                              int64_t* __restrict__ perm, double* __restrict__ s_out,
                                              int* __restrict__ nonfinite, int64_t sn, int64_t ss,
                                              const double* __restrict__ scl) {
  __shared__ double key[256];
  nrm += blockIdx.y * sn;
  perm += blockIdx.y * sn;
  s_out += blockIdx.y * ss;
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const double sj = j < q ? nrm[j] : 0.0;
  if (j < q && !isfinite(sj)) atomicOr(nonfinite, 1);
  const double kj = isfinite(sj) ? sj : -1.0;
  int64_t rank = 0;
  for (int64_t c0 = 0; c0 < q; c0 += 256) {
    __syncthreads();
    const int64_t ic = c0 + threadIdx.x;
    key[threadIdx.x] = ic < q ? (isfinite(nrm[ic]) ? nrm[ic] : -1.0) : 0.0;
    __syncthreads();
    const int64_t n = (q - c0 < 256) ? q - c0 : 256;
    for (int64_t u = 0; u < n; ++u) {
      const double ki = key[u];
      rank += (ki > kj) || (ki == kj && c0 + u < j);
    }
  }
  if (j < q) {
    perm[rank] = j;
    s_out[rank] = sj / scl[blockIdx.y];   // undo the input scaling (a power of two)
  }
}

// row-major outputs, pc = perm[c], from U_B (column-major p x q) and V_B
// (column-major q x q) of B = U_B diag(sigma) V_B^H:
//   m >= n (B = A):    U[i][c] = UB[i][pc],  Vh[c][j] = conj(VB[j][pc])
//   m <  n (B = A^H):  U[i][c] = VB[i][pc],  Vh[c][j] = conj(UB[j][pc])
// Problem blockIdx.y of a batch: UB, VB, perm (stride sn) stacked in the
// workspace, u and vh with strides su and sv.
__global__ void k_svd_out(int64_t m, int64_t n, int64_t k, int adjoint, int64_t p, int64_t q,
                          const double2* __restrict__ UB, const double2* __restrict__ VB,
                          const int64_t* __restrict__ perm, double2* __restrict__ u, int64_t ldu,
                          double2* __restrict__ vh, int64_t ldvh, int64_t sn, int64_t su,
                          int64_t sv) {
  const int64_t b = blockIdx.y;
  UB += b * p * q;
  VB += b * q * q;
  perm += b * sn;
  u += b * su;
  vh += b * sv;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m * k) {
    const int64_t i = t / k, c = t % k, pc = perm[c];
    u[i * ldu + c] = adjoint ? VB[pc * q + i] : UB[pc * p + i];
  }
  if (t < k * n) {
    const int64_t c = t / n, j = t % n, pc = perm[c];
    const double2 v = adjoint ? UB[pc * p + j] : VB[pc * q + j];
    vh[c * ldvh + j] = make_double2(v.x, -v.y);
  }
}

// B = A^H (row-major n x m) from A (m x n, lda); problem blockIdx.y of a batch
// at a + y sa, b + y m n
__global__ void k_adjoint(int64_t m, int64_t n, const double2* __restrict__ a, int64_t lda,
                          double2* __restrict__ b, int64_t sa) {
  a += blockIdx.y * sa;
  b += blockIdx.y * m * n;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * n) return;
  const int64_t i = t / n, j = t % n;
  const double2 v = a[i * lda + j];
  b[j * m + i] = make_double2(v.x, -v.y);
}

// R^H of a factorisation (strict upper triangle of R in the column-major p x q
// W, diagonal in rdiag): element (i, c) of R^H is conj(R[c][i]). Written either
// column-major with leading dimension ld and zero padding columns up to ncols
// (colmajor = 1), or row-major q x q (colmajor = 0). Problem blockIdx.y of a
// batch: W, rdiag with strides sw, sr; out with stride ld ncols.
__global__ void k_rh(int64_t p, int64_t q, int64_t ld, int64_t ncols, int colmajor,
                     const double2* __restrict__ W, const double2* __restrict__ rdiag,
                     double2* __restrict__ out, int64_t sw, int64_t sr) {
  W += blockIdx.y * sw;
  rdiag += blockIdx.y * sr;
  out += blockIdx.y * ld * ncols;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ld * ncols;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = colmajor ? t / ld : t % ld, i = colmajor ? t % ld : t / ld;
    double2 v = make_double2(0.0, 0.0);
    if (c < q && i < q) v = (c < i) ? W[i * p + c] : (c == i ? rdiag[c] : v);
    out[t] = make_double2(v.x, -v.y);
  }
}

// UB (column-major p x q) = [X / sigma; 0] for X column-major q x q (sigma = its
// column norms; a zero column stays zero). Problem blockIdx.y of a batch: X
// with stride sx, nrm with stride sn, UB with stride p q.
__global__ void k_embed_u(int64_t p, int64_t q, const double2* __restrict__ X,
                          const double* __restrict__ nrm, double2* __restrict__ UB, int64_t sx,
                          int64_t sn) {
  X += blockIdx.y * sx;
  nrm += blockIdx.y * sn;
  UB += blockIdx.y * p * q;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < p * q;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / p, i = t % p;
    double2 v = make_double2(0.0, 0.0);
    const double sc = nrm[c];
    if (i < q && sc > 0) { const double2 x = X[c * q + i]; v = make_double2(x.x / sc, x.y / sc); }
    UB[t] = v;
  }
}

// VB (column-major q x q) = V(0:q, 0:q) for V column-major with leading dimension qp
// (problem blockIdx.y of a batch: V with stride qp qp, VB with stride q q)
__global__ void k_copy_v(int64_t q, int64_t qp, const double2* __restrict__ V,
                         double2* __restrict__ VB) {
  V += blockIdx.y * qp * qp;
  VB += blockIdx.y * q * q;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < q * q;
       t += (int64_t)gridDim.x * blockDim.x)
    VB[t] = V[(t / q) * qp + t % q];
}

__global__ void k_set_ints(int* p, int n, int v) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

__global__ void k_svd_info(int32_t* info, int sweeps, int notconv, const int* counters,
                           int sweeps32) {
  if (threadIdx.x == 0) {
    info[0] = sweeps;
    info[1] = notconv;
    info[2] = counters[1];
    info[3] = sweeps32;
  }
}

// scale[b] = largest column norm of problem b (1 when zero or not finite), one
// workgroup per problem
__global__ __launch_bounds__(256) void k_maxnorm(int64_t q, const double* __restrict__ nrm,
                                                 int64_t sn, double* __restrict__ scale) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  nrm += blockIdx.x * sn;
  double m = 0.0;
  for (int64_t i = t; i < q; i += 256) m = fmax(m, nrm[i]);
  red[t] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] = fmax(red[t], red[t + w]);
    __syncthreads();
  }
  if (t == 0) scale[blockIdx.x] = (red[0] > 0.0 && isfinite(red[0])) ? red[0] : 1.0;
}

// X32 = X / scale[b] rounded to complex64 (n elements per problem, problem blockIdx.y)
__global__ void k_narrow_scaled(int64_t n, const double2* __restrict__ X,
                                const double* __restrict__ scale, float2* __restrict__ X32) {
  X += blockIdx.y * n;
  X32 += blockIdx.y * n;
  const double inv = 1.0 / scale[blockIdx.y];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = X[t];
    X32[t] = make_float2((float)(x.x * inv), (float)(x.y * inv));
  }
}

// complex64 identity, q x q per problem (problem blockIdx.y)
__global__ void k_eye32(int64_t q, float2* __restrict__ V) {
  V += (int64_t)blockIdx.y * q * q;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < q * q;
       t += (int64_t)gridDim.x * blockDim.x)
    V[t] = make_float2((t / q) == (t % q) ? 1.0f : 0.0f, 0.0f);
}

__global__ void k_widen(int64_t n, const float2* __restrict__ x, double2* __restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    y[t] = make_double2((double)x[t].x, (double)x[t].y);
}

// Newton-Schulz step of the unitary polar factor: out = 1.5 Q - 0.5 Q (Q^H Q)
// with QG = Q (Q^H Q) (n elements)
__global__ void k_ns_update(int64_t n, const double2* __restrict__ Q, const double2* __restrict__ QG,
                            double2* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = Q[t], b = QG[t];
    out[t] = make_double2(1.5 * a.x - 0.5 * b.x, 1.5 * a.y - 0.5 * b.y);
  }
}

static inline size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

// pinned word for the per-sweep rotation count, allocated once per process
static int* pinned_counter() {
  static thread_local int* p = nullptr;
  if (!p && hipHostMalloc((void**)&p, sizeof(int), hipHostMallocDefault) != hipSuccess) p = nullptr;
  return p;
}

struct SvdLayout {
  int64_t p, q, qp, nseg, batch;
  int nblk;
  bool pre32;
  size_t bt, fact, fact2, rh, x, v, ub, vb, gp, jg, rflag, nrm, perm, cnt, x32, v32, t1, t2, t3,
      scale, total;
};

// FP32 preconditioning sweeps are used from this many columns up
constexpr int64_t PRE32_MIN_Q = 4 * JP;

// every per-problem region holds `batch` consecutive copies
static SvdLayout svd_layout(int64_t m, int64_t n, int64_t batch, bool precond32) {
  SvdLayout L;
  L.batch = batch;
  L.q = m < n ? m : n;
  L.p = m < n ? n : m;
  L.qp = ((L.q + JP - 1) / JP) * JP;
  L.nblk = (int)(L.qp / JB);
  L.nseg = (L.q + GSEG - 1) / GSEG;   // Jacobi runs on X = R2^H with q rows
  const int npairs = L.nblk / 2;
  const size_t B = (size_t)batch;
  size_t off = 0;
  L.bt = off; off += (m < n) ? al(B * L.p * L.q * 16) : 0;
  L.fact = off; off += al(qr_fact_bytes(L.p, L.q, L.q, batch));
  L.fact2 = off; off += al(qr_fact_bytes(L.q, L.q, L.q, batch));
  L.rh = off; off += al(B * L.q * L.q * 16);
  L.x = off; off += al(B * L.q * L.qp * 16);
  L.v = off; off += al(B * L.qp * L.qp * 16);
  L.ub = off; off += al(B * L.p * L.q * 16);
  L.vb = off; off += al(B * L.q * L.q * 16);
  L.gp = off; off += al(B * npairs * L.nseg * JP * JP * 16);
  L.jg = off; off += al(B * npairs * JP * JP * 16);
  L.rflag = off; off += al(B * npairs * 4);
  L.nrm = off; off += al(B * L.qp * 8);
  L.perm = off; off += al(B * L.qp * 8);
  L.cnt = off; off += al(16);
  L.pre32 = precond32 && L.q >= PRE32_MIN_Q;
  L.x32 = L.v32 = L.t1 = L.t2 = L.t3 = L.scale = off;
  if (L.pre32) {
    L.x32 = off; off += al(B * L.q * L.qp * 8);
    L.v32 = off; off += al(B * L.qp * L.qp * 8);
    L.t1 = off; off += al(B * L.qp * L.qp * 16);
    L.t2 = off; off += al(B * L.qp * L.qp * 16);
    L.t3 = off; off += al(B * L.qp * L.qp * 16);
    L.scale = off; off += al(B * 8);
  }
  L.total = off;
  return L;
}

size_t svd_ws_bytes(int64_t m, int64_t n, int64_t batch, bool precond32) {
  return svd_layout(m, n, batch, precond32).total;
}

// grid width of the grid-stride helper kernels for `total` elements per problem
static inline unsigned gs_blocks(int64_t total) {
  const int64_t b = (total + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

// Thin SVD of `batch` m x n matrices of identical shape (problem b at a + b sa,
// u + b su, s + b ss, vh + b svh). The Jacobi sweeps run on all problems at
// once and stop together, after the first sweep in which no problem needed a
// rotation above `big` (relative). info (if given) receives the number of
// sweeps, whether the sweep limit was hit, and whether any singular value was
// not finite.
// With `precond32` (and at least PRE32_MIN_Q columns) the Jacobi phase starts
// with FP32 sweeps: X32 = X / max column norm in complex64 is orthogonalised
// on FP32 MFMA until no pair needs a rotation above 1e-3 relative, with the
// rotations accumulated in V32. V32 is unitary only to FP32 rounding, so it is
// replaced by its unitary polar factor Q (two Newton-Schulz steps in FP64,
// Q <- 1.5 Q - 0.5 Q (Q^H Q), each squaring the distance from unitarity), and
// the FP64 sweeps continue from X Q with V = Q. The FP64 sweeps and their
// stopping rule are those of the plain path, so the result has the same FP64
// accuracy; the FP32 sweeps only remove most of the rotation work.
// Jacobi phase on X (q x qp column-major per problem, the preconditioned
// matrix R2^H). With want_v the rotations are accumulated into V (qp x qp per
// problem); without, V's buffer only serves the FP32 preconditioning (polar
// factor). On return X points at the orthogonalised columns (another buffer
// of the workspace after FP32 preconditioning).
struct JacobiResult {
  int sweeps = 0, sweeps32 = 0, converged = 0;
};
static int jacobi_core(const SvdLayout& L, char* base, double2*& X, double2* V, bool want_v,
                       hipStream_t st, JacobiResult& res) {
  const unsigned NB = (unsigned)L.batch;
  const int npairs = L.nblk / 2;
  const int64_t aseg = (L.qp + AROWS - 1) / AROWS;   // qp >= q rows
  double2* Gp = (double2*)(base + L.gp);
  double2* Jg = (double2*)(base + L.jg);
  int* rflag = (int*)(base + L.rflag);
  double* nrm = (double*)(base + L.nrm);
  int* counters = (int*)(base + L.cnt);   // [0] rotations in the current sweep, [1] non-finite
  const int64_t sx = L.q * L.qp;
  const int batch = (int)L.batch;
  int stt;
  int* host_rot = pinned_counter();
  if (!host_rot) return PTN_ERR_HIP;
  // algorithmic work per round (for the live launch timing): Gram of every
  // block pair over the q rows (8 flops per complex multiply-add), one
  // rotation sweep of JP(JP-1)/2 rotations per pair (each updating two rows
  // and two columns of the pair's Gram matrix and two columns of J: 12 JP
  // complex multiply-adds), the pair's columns times J in X (and V); bytes in
  // complex128 elements (the FP32 phase moves half): the pair's columns read
  // (and written back by the apply), partial Gram matrices written / read
  struct RoundWork { double gram_flop, gram_bytes, solve_flop, solve_bytes, apply_flop, apply_bytes; };
  const double np_b = (double)npairs * (double)L.batch;
  const RoundWork W = {np_b * 8.0 * (double)L.q * JP * JP,
                       np_b * 16.0 * ((double)L.q * JP + (double)L.nseg * JP * JP),
                       np_b * 8.0 * 12.0 * JP * (JP * (JP - 1) / 2),
                       np_b * 16.0 * ((double)L.nseg + 1) * JP * JP,
                       np_b * 8.0 * (double)L.q * JP * JP,
                       np_b * 16.0 * 2.0 * (double)L.q * JP};
  k_set_ints<<<1, 64, 0, st>>>(counters, 2, 0);
  if (L.pre32) {
    float2* X32 = (float2*)(base + L.x32);
    float2* V32 = (float2*)(base + L.v32);
    float2* Gp32 = (float2*)Gp;
    float2* Jg32 = (float2*)Jg;
    double* scale = (double*)(base + L.scale);
    k_colnorms<<<dim3((unsigned)L.q, NB), 256, 0, st>>>(L.q, X, nrm, sx, L.qp);
    k_maxnorm<<<NB, 256, 0, st>>>(L.q, nrm, L.qp, scale);
    k_narrow_scaled<<<dim3(gs_blocks(sx), NB), 256, 0, st>>>(sx, X, scale, X32);
    k_eye32<<<dim3(gs_blocks(L.qp * L.qp), NB), 256, 0, st>>>(L.qp, V32);
    PTN_LAUNCH_CHECK();
    const float tol32 = 32.0f * FLT_EPSILON, big32 = 1e-3f;
    const int max_sweeps32 = 20;
    while (res.sweeps32 < max_sweeps32) {
      k_set_ints<<<1, 64, 0, st>>>(counters, 1, 0);
      for (int r = 0; r < L.nblk - 1; ++r) {
        tlaunch("k_jac_gram<float>", W.gram_flop, W.gram_bytes / 2, k_jac_gram<float>,
                dim3((unsigned)L.nseg, (unsigned)npairs, NB), dim3(256), st, (int64_t)L.q, L.nblk,
                r, (const float2*)X32, Gp32, (int)L.nseg, sx);
        tlaunch("k_jac_solve<float>", W.solve_flop, W.solve_bytes / 2, k_jac_solve<float>,
                dim3((unsigned)npairs, NB), dim3(256), st, L.nblk, (const float2*)Gp32,
                (int)L.nseg, tol32, big32, Jg32, rflag, counters);
        tlaunch("k_jac_apply<float>", 2 * W.apply_flop, W.apply_bytes, k_jac_apply<float>,
                dim3((unsigned)aseg, (unsigned)npairs, 2 * NB), dim3(256), st, (int64_t)L.q,
                (int64_t)L.qp, L.nblk, r, X32, V32, (const float2*)Jg32, (const int*)rflag, 2);
      }
      PTN_LAUNCH_CHECK();
      PTN_HIP_CHECK(hipMemcpyAsync(host_rot, counters, sizeof(int), hipMemcpyDeviceToHost, st));
      PTN_HIP_CHECK(hipStreamSynchronize(st));
      ++res.sweeps32;
      if (*host_rot == 0) break;
    }
    // unitary polar factor of V32 (qp x qp column-major per problem) into V
    const int64_t nv = L.qp * L.qp, nvb = nv * batch;
    double2* Q = (double2*)(base + L.t1);
    double2* QhQ = (double2*)(base + L.t2);
    double2* QQ = (double2*)(base + L.t3);
    k_widen<<<gs_blocks(nvb), 256, 0, st>>>(nvb, V32, Q);
    PTN_LAUNCH_CHECK();
    for (int it = 0; it < 2; ++it) {
      ZGemm g1 = {L.qp, L.qp, L.qp, Q, L.qp, 1, 1, Q, 1, L.qp, 0, QhQ, 1, L.qp, 0,
                  L.batch, nv, nv, nv};
      if ((stt = zgemm(g1, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, st)))
        return stt;
      ZGemm g2 = {L.qp, L.qp, L.qp, Q, 1, L.qp, 0, QhQ, 1, L.qp, 0, QQ, 1, L.qp, 0,
                  L.batch, nv, nv, nv};
      if ((stt = zgemm(g2, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, st)))
        return stt;
      k_ns_update<<<gs_blocks(nvb), 256, 0, st>>>(nvb, Q, QQ, it == 1 ? V : Q);
      PTN_LAUNCH_CHECK();
    }
    // X <- X V (q x qp column-major per problem), into t3
    ZGemm g3 = {L.q, L.qp, L.qp, X, 1, L.q, 0, V, 1, L.qp, 0, QQ, 1, L.q, 0,
                L.batch, sx, nv, sx};
    if ((stt = zgemm(g3, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, st))) return stt;
    X = QQ;
  } else if (want_v) {
    k_eye<<<dim3(gs_blocks(L.qp * L.qp), NB), 256, 0, st>>>(L.qp, V);
    PTN_LAUNCH_CHECK();
  }
  const double tol = DBL_EPSILON * sqrt((double)L.q);
  // the iteration stops after a sweep in which no pair needed a rotation above
  // this relative size: convergence is quadratic, so the couplings left after
  // such a sweep are of order big^2 (n big^2 at most), and a further sweep
  // would only rotate by rounding-level angles
  const double big = 1e-8;
  const int max_sweeps = 60;
  const int nmat = want_v ? 2 : 1;
  res.converged = (L.q < 2);
  while (!res.converged && res.sweeps < max_sweeps) {
    k_set_ints<<<1, 64, 0, st>>>(counters, 1, 0);
    for (int r = 0; r < L.nblk - 1; ++r) {
      tlaunch("k_jac_gram<double>", W.gram_flop, W.gram_bytes, k_jac_gram<double>,
              dim3((unsigned)L.nseg, (unsigned)npairs, NB), dim3(256), st, (int64_t)L.q, L.nblk, r,
              (const double2*)X, Gp, (int)L.nseg, sx);
      tlaunch("k_jac_solve<double>", W.solve_flop, W.solve_bytes, k_jac_solve<double>,
              dim3((unsigned)npairs, NB), dim3(256), st, L.nblk, (const double2*)Gp, (int)L.nseg,
              tol, big, Jg, rflag, counters);
      tlaunch("k_jac_apply<double>", nmat * W.apply_flop, nmat * W.apply_bytes,
              k_jac_apply<double>, dim3((unsigned)aseg, (unsigned)npairs, nmat * NB), dim3(256), st,
              (int64_t)L.q, (int64_t)L.qp, L.nblk, r, X, V, (const double2*)Jg, (const int*)rflag,
              nmat);
    }
    PTN_LAUNCH_CHECK();
    PTN_HIP_CHECK(hipMemcpyAsync(host_rot, counters, sizeof(int), hipMemcpyDeviceToHost, st));
    PTN_HIP_CHECK(hipStreamSynchronize(st));
    ++res.sweeps;
    if (*host_rot == 0) res.converged = 1;
  }
  return PTN_OK;
}

// Thin SVD of `batch` m x n matrices of identical shape (problem b at a + b sa,
// u + b su, s + b ss, vh + b svh). The Jacobi sweeps run on all problems at
// once and stop together, after the first sweep in which no problem needed a
// rotation above `big` (relative). info (if given) receives the number of
// sweeps, whether the sweep limit was hit, and whether any singular value was
// not finite.
// With `precond32` (and at least PRE32_MIN_Q columns) the Jacobi phase starts
// with FP32 sweeps: X32 = X / max column norm in complex64 is orthogonalised
// on FP32 MFMA until no pair needs a rotation above 1e-3 relative, with the
// rotations accumulated in V32. V32 is unitary only to FP32 rounding, so it is
// replaced by its unitary polar factor Q (two Newton-Schulz steps in FP64,
// Q <- 1.5 Q - 0.5 Q (Q^H Q), each squaring the distance from unitarity), and
// the FP64 sweeps continue from X Q with V = Q. The FP64 sweeps and their
// stopping rule are those of the plain path, so the result has the same FP64
// accuracy; the FP32 sweeps only remove most of the rotation work.
int svd_run(int64_t m, int64_t n, const void* a, int64_t lda, void* u, int64_t ldu, double* s,
            void* vh, int64_t ldvh, int32_t* info, void* ws, size_t ws_bytes, hipStream_t st,
            int64_t batch, int64_t sa, int64_t su, int64_t ss, int64_t svh, bool precond32) {
  if (m <= 0 || n <= 0 || lda < n) return PTN_ERR_ARG;
  const int64_t k = m < n ? m : n;
  if (ldu < k || ldvh < n) return PTN_ERR_ARG;
  if (batch < 1 || batch > 32767) return PTN_ERR_ARG;
  const SvdLayout L = svd_layout(m, n, batch, precond32);
  if (ws_bytes < L.total) return PTN_ERR_WORKSPACE;
  const int npairs = L.nblk / 2;
  const int64_t aseg = (L.qp + AROWS - 1) / AROWS;   // qp >= q rows
  if (npairs > 65535 || aseg > 65535 || L.q > 65535) return PTN_ERR_ARG;
  const int adjoint = m < n;
  const unsigned NB = (unsigned)batch;
  char* base = (char*)ws;
  double2* X = (double2*)(base + L.x);
  double2* V = (double2*)(base + L.v);
  double2* UB = (double2*)(base + L.ub);
  double2* VB = (double2*)(base + L.vb);
  double* nrm = (double*)(base + L.nrm);
  int64_t* perm = (int64_t*)(base + L.perm);
  int* counters = (int*)(base + L.cnt);
  const int64_t sx = L.q * L.qp;          // X: q x qp column-major per problem
  int stt;

  // 1. B = Q R
  const void* B = a;
  int64_t ldb = lda, sb = sa;
  if (adjoint) {
    k_adjoint<<<dim3((unsigned)((m * n + 255) / 256), NB), 256, 0, st>>>(
        m, n, (const double2*)a, lda, (double2*)(base + L.bt), sa);
    PTN_LAUNCH_CHECK();
    B = base + L.bt;
    ldb = m;
    sb = m * n;
  }
  QrFact f = qr_fact_layout(L.p, L.q, L.q, base + L.fact, batch);
  if ((stt = qr_factor(f, (const double2*)B, ldb, st, sb))) return stt;
  // 2. R^H = Q2 R2, and one-sided Jacobi on X = R2^H
  QrFact f2 = qr_fact_layout(L.q, L.q, L.q, base + L.fact2, batch);
  double2* RH = (double2*)(base + L.rh);
  k_rh<<<dim3(gs_blocks(L.q * L.q), NB), 256, 0, st>>>(L.p, L.q, L.q, L.q, 0, f.W, f.rdiag, RH,
                                                       f.bs.W, f.bs.k);
  PTN_LAUNCH_CHECK();
  if ((stt = qr_factor(f2, RH, L.q, st, L.q * L.q))) return stt;
  k_rh<<<dim3(gs_blocks(L.q * L.qp), NB), 256, 0, st>>>(L.q, L.q, L.q, L.qp, 1, f2.W, f2.rdiag, X,
                                                        f2.bs.W, f2.bs.k);
  PTN_LAUNCH_CHECK();
  JacobiResult jr;
  if ((stt = jacobi_core(L, base, X, V, true, st, jr))) return stt;
  // 3. B = Q1 R2^H Q2^H = (Q1 [U_X; 0]) diag(sigma) (Q2 V)^H with U_X = X / sigma
  k_colnorms<<<dim3((unsigned)L.q, NB), 256, 0, st>>>(L.q, X, nrm, sx, L.qp);
  k_embed_u<<<dim3(gs_blocks(L.p * L.q), NB), 256, 0, st>>>(L.p, L.q, X, nrm, UB, sx, L.qp);
  k_copy_v<<<dim3(gs_blocks(L.q * L.q), NB), 256, 0, st>>>(L.q, L.qp, V, VB);
  PTN_LAUNCH_CHECK();
  if ((stt = qr_apply_q(f, UB, L.p, L.q, false, st, L.p * L.q))) return stt;
  if ((stt = qr_apply_q(f2, VB, L.q, L.q, false, st, L.q * L.q))) return stt;
  k_rank<<<dim3((unsigned)((L.q + 255) / 256), NB), 256, 0, st>>>(L.q, nrm, perm, s, counters + 1,
                                                                   L.qp, ss, f.scl);
  const int64_t tot = (m * k > k * n) ? m * k : k * n;
  k_svd_out<<<dim3((unsigned)((tot + 255) / 256), NB), 256, 0, st>>>(
      m, n, k, adjoint, L.p, L.q, UB, VB, perm, (double2*)u, ldu, (double2*)vh, ldvh, L.qp, su,
      svh);
  if (info)
    k_svd_info<<<1, 64, 0, st>>>(info, jr.sweeps, jr.converged ? 0 : 1, counters, jr.sweeps32);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// ---------------------------------------------------------------------------
// Truncating split (one sector): the factors of the two-site split without the
// unneeded half of the SVD.
//
// For sigma into the right factor ("right") the left factor is the isometry U_k
// of the kept left singular vectors and the right one U_k^H Theta (= Sigma_k
// V_k^H); for "left" the roles are mirrored on Theta^H: the right factor is
// V_k^H and the left one Theta V_k (= U_k Sigma_k). Either way only the left
// singular vectors of one matrix M (Theta or Theta^H) are needed: with M' = M
// (tall) or M' = R0^H from M^H = Q0 R0 (wide; same left singular vectors),
// M' = Q1 R1 and R1^H = Q2 R2 (the two QR preconditioners of svd_run), the
// one-sided Jacobi iteration on X = R2^H gives U_X (X V = U_X Sigma) without
// accumulating V, and U_M = Q1 [U_X; 0]. Only the kept columns of U_X go
// through Q1, and the other factor is one product with Theta, which also
// carries exactly the weight of the kept subspace (the projection of Theta).

// UB (column-major p x k) = [X(:, perm[c]) / sigma_perm[c]; 0] for c < k
__global__ void k_embed_u_perm(int64_t p, int64_t q, int64_t k, const double2* __restrict__ X,
                               const double* __restrict__ nrm, const int64_t* __restrict__ perm,
                               double2* __restrict__ UB) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < p * k;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / p, i = t % p, pc = perm[c];
    double2 v = make_double2(0.0, 0.0);
    const double sc = nrm[pc];
    if (i < q && sc > 0) { const double2 x = X[pc * q + i]; v = make_double2(x.x / sc, x.y / sc); }
    UB[t] = v;
  }
}

// out (row-major p x k, optionally conjugated) from UB (column-major p x k)
__global__ void k_colmajor_out(int64_t p, int64_t k, const double2* __restrict__ UB,
                               double2* __restrict__ out, int conj) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < p * k;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / k, c = t % k;
    const double2 v = UB[c * p + i];
    out[t] = make_double2(v.x, conj ? -v.y : v.y);
  }
}

// elementwise complex conjugate copy (n elements)
__global__ void k_conj_copy(int64_t n, const double2* __restrict__ x, double2* __restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    y[t] = make_double2(x[t].x, -x[t].y);
}

struct SplitLayout {
  int64_t pM, qM, pi, qi;   // M (pM x qM) and the tall inner matrix M' (pi x qi)
  bool wide;
  size_t adj, fact0, t0, inner, total;
  SvdLayout L;
};

// dst = product of the two input scale factors (one problem)
__global__ void k_mul_scale(const double* a, const double* b, double* dst) {
  if (threadIdx.x == 0) *dst = *a * *b;
}

static SplitLayout split_layout(int64_t m, int64_t n, int distr, bool precond32) {
  SplitLayout S;
  S.pM = distr == PTN_SPLIT_LEFT ? n : m;
  S.qM = distr == PTN_SPLIT_LEFT ? m : n;
  S.wide = S.pM < S.qM;
  S.pi = S.pM;
  S.qi = S.wide ? S.pM : S.qM;
  size_t off = 0;
  // Theta^H (the QR input for "left", or M^H for a wide "right" split)
  const bool need_adj = (distr == PTN_SPLIT_LEFT) != S.wide;
  S.adj = off; off += need_adj ? al((size_t)m * n * 16) : 0;
  S.fact0 = off; off += S.wide ? al(qr_fact_bytes(S.qM, S.pM, S.pM)) : 0;
  S.t0 = off; off += S.wide ? al((size_t)S.pM * S.pM * 16) : 0;
  off += al(16);   // combined input scale (the last 16 bytes before the inner layout)
  S.inner = off;
  S.L = svd_layout(S.pi, S.qi, 1, precond32);
  off += S.L.total;
  S.total = off;
  return S;
}

size_t svd_split_ws_bytes(int64_t m, int64_t n, int distr, bool precond32) {
  if (distr != PTN_SPLIT_LEFT && distr != PTN_SPLIT_RIGHT) return 0;
  return split_layout(m, n, distr, precond32).total;
}

// kept count of the discarded-weight rule (pytenet/bond_ops.py:23-38) on the
// descending singular values s[0..q): normalise by the 2-norm, square,
// accumulate from the smallest upwards, keep the values whose running sum
// exceeds tol; then the cap (the max_bond largest of them; <= 0: no cap)
static int64_t kept_count(const double* s, int64_t q, double tol, int64_t max_bond) {
  double w2 = 0.0;
  for (int64_t i = 0; i < q; ++i) w2 += s[i] * s[i];
  if (!(w2 > 0.0)) return 0;
  const double w = sqrt(w2);
  int64_t k = q;
  double acc = 0.0;
  for (int64_t i = q - 1; i >= 0; --i) {   // ascending weights
    const double pi = (s[i] / w) * (s[i] / w);
    acc += pi;
    if (acc > tol) break;
    k = i;
  }
  if (max_bond > 0 && k > max_bond) k = max_bond;
  return k;
}

int svd_split_run(int64_t m, int64_t n, const void* theta, int64_t lda, int distr, double tol,
                  int64_t max_bond, bool precond32, void* a0, void* a1, double* s_host,
                  int64_t* k_out, int32_t* info_host, void* ws, size_t ws_bytes, hipStream_t st) {
  if (m <= 0 || n <= 0 || lda < n || !k_out || !s_host) return PTN_ERR_ARG;
  if (distr != PTN_SPLIT_LEFT && distr != PTN_SPLIT_RIGHT) return PTN_ERR_ARG;
  const SplitLayout S = split_layout(m, n, distr, precond32);
  if (ws_bytes < S.total) return PTN_ERR_WORKSPACE;
  const SvdLayout& L = S.L;
  const int npairs = L.nblk / 2;
  const int64_t aseg = (L.qp + AROWS - 1) / AROWS;
  if (npairs > 65535 || aseg > 65535 || L.q > 65535) return PTN_ERR_ARG;
  char* wsb = (char*)ws;
  char* base = wsb + S.inner;
  int stt;
  const double2* th = (const double2*)theta;
  // M^H in adj when the QR input needs it
  double2* adj = (double2*)(wsb + S.adj);
  const bool need_adj = (distr == PTN_SPLIT_LEFT) != S.wide;
  if (need_adj) {
    k_adjoint<<<dim3((unsigned)((m * n + 255) / 256), 1), 256, 0, st>>>(m, n, th, lda, adj, 0);
    PTN_LAUNCH_CHECK();
  }
  // the matrix whose QR starts the factorisation: M' (tall) directly, or M^H (wide)
  const double2* qin;
  int64_t ldq;
  if (!S.wide) {
    qin = distr == PTN_SPLIT_LEFT ? adj : th;                // M = Theta^H / Theta
    ldq = distr == PTN_SPLIT_LEFT ? m : lda;
  } else {
    qin = distr == PTN_SPLIT_LEFT ? th : adj;                // M^H = Theta / Theta^H
    ldq = distr == PTN_SPLIT_LEFT ? lda : m;
  }
  QrFact f0;
  if (S.wide) {
    // M^H (qM x pM) = Q0 R0; M' = R0^H (pM x pM, row-major) has M's left singular vectors
    f0 = qr_fact_layout(S.qM, S.pM, S.pM, wsb + S.fact0, 1);
    if ((stt = qr_factor(f0, qin, ldq, st, 0))) return stt;
    double2* T0 = (double2*)(wsb + S.t0);
    k_rh<<<dim3(gs_blocks(S.pM * S.pM), 1), 256, 0, st>>>(S.qM, S.pM, S.pM, S.pM, 0, f0.W,
                                                          f0.rdiag, T0, f0.bs.W, f0.bs.k);
    PTN_LAUNCH_CHECK();
    qin = T0;
    ldq = S.pM;
  }
  double2* X = (double2*)(base + L.x);
  double2* V = (double2*)(base + L.v);
  double2* UB = (double2*)(base + L.ub);
  double* nrm = (double*)(base + L.nrm);
  int64_t* perm = (int64_t*)(base + L.perm);
  int* counters = (int*)(base + L.cnt);
  double* sdev = (double*)(base + L.vb);   // the singular values (V_B is not formed)
  const int64_t sx = L.q * L.qp;
  QrFact f = qr_fact_layout(L.p, L.q, L.q, base + L.fact, 1);
  if ((stt = qr_factor(f, qin, ldq, st, 0))) return stt;
  QrFact f2 = qr_fact_layout(L.q, L.q, L.q, base + L.fact2, 1);
  double2* RH = (double2*)(base + L.rh);
  k_rh<<<dim3(gs_blocks(L.q * L.q), 1), 256, 0, st>>>(L.p, L.q, L.q, L.q, 0, f.W, f.rdiag, RH,
                                                      f.bs.W, f.bs.k);
  PTN_LAUNCH_CHECK();
  if ((stt = qr_factor(f2, RH, L.q, st, L.q * L.q))) return stt;
  k_rh<<<dim3(gs_blocks(L.q * L.qp), 1), 256, 0, st>>>(L.q, L.q, L.q, L.qp, 1, f2.W, f2.rdiag, X,
                                                       f2.bs.W, f2.bs.k);
  PTN_LAUNCH_CHECK();
  JacobiResult jr;
  if ((stt = jacobi_core(L, base, X, V, false, st, jr))) return stt;
  // the singular values carry the power-of-two input scales of the QR(s)
  double* scl = f.scl;
  if (S.wide) {
    scl = (double*)(wsb + S.inner - 16);
    k_mul_scale<<<1, 64, 0, st>>>(f.scl, f0.scl, scl);
  }
  k_colnorms<<<dim3((unsigned)L.q, 1), 256, 0, st>>>(L.q, X, nrm, sx, L.qp);
  k_rank<<<dim3((unsigned)((L.q + 255) / 256), 1), 256, 0, st>>>(L.q, nrm, perm, sdev,
                                                                  counters + 1, L.qp, L.q, scl);
  PTN_LAUNCH_CHECK();
  // the singular values and the non-finite flag on the host (the kept count
  // sizes the outputs)
  int* host_rot = pinned_counter();
  if (!host_rot) return PTN_ERR_HIP;
  PTN_HIP_CHECK(hipMemcpyAsync(s_host, sdev, (size_t)L.q * sizeof(double), hipMemcpyDeviceToHost,
                               st));
  PTN_HIP_CHECK(hipMemcpyAsync(host_rot, counters + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  PTN_HIP_CHECK(hipStreamSynchronize(st));
  if (info_host) {
    info_host[0] = jr.sweeps;
    info_host[1] = jr.converged ? 0 : 1;
    info_host[2] = *host_rot;
    info_host[3] = jr.sweeps32;
  }
  if (*host_rot) { *k_out = 0; return PTN_ERR_NONFINITE; }
  const int64_t k = kept_count(s_host, L.q, tol, max_bond);
  *k_out = k;
  if (k == 0) return PTN_OK;
  // U_M(:, :k) = Q1 [U_X(:, perm[:k]); 0]   (pM x k column-major; f.scl does not enter U)
  k_embed_u_perm<<<gs_blocks(L.p * k), 256, 0, st>>>(L.p, L.q, k, X, nrm, perm, UB);
  PTN_LAUNCH_CHECK();
  if ((stt = qr_apply_q(f, UB, L.p, k, false, st, 0))) return stt;
  if (distr == PTN_SPLIT_RIGHT) {
    // a0 = U_k (m x k row-major); a1 = U_k^H Theta (k x n)
    k_colmajor_out<<<gs_blocks(m * k), 256, 0, st>>>(m, k, UB, (double2*)a0, 0);
    PTN_LAUNCH_CHECK();
    ZGemm g = {k, n, m, UB, m, 1, 1, theta, lda, 1, 0, a1, n, 1, 0, 1, 0, 0, 0};
    return zgemm(g, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, st);
  }
  // "left": V_k = U_M(:, :k) (n x k column-major); a1 = V_k^H (k x n row-major:
  // the conjugate of the column-major storage), a0 = Theta V_k (m x k)
  k_conj_copy<<<gs_blocks(n * k), 256, 0, st>>>(n * k, UB, (double2*)a1);
  PTN_LAUNCH_CHECK();
  ZGemm g = {m, k, n, theta, lda, 1, 0, UB, 1, n, 0, a0, k, 1, 0, 1, 0, 0, 0};
  return zgemm(g, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, st);
}

}  // namespace ptn
