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
// Sweeps repeat until no pair needs a rotation, i.e. every column pair
// satisfies |x_i^H x_j| <= tol |x_i| |x_j|. Then sigma_j = |x_j| and
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
constexpr int AROWS = 64;       // rows per apply workgroup (16 per wave)

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

// partial Gram matrices Gp[pair][seg][i][j] = sum_{rows in seg} conj(Y[row][i]) Y[row][j]
// on FP64 MFMA (v_mfma_f64_16x16x4f64) through the real embedding of the
// complex product with A = Y^H (real k = 2 row + part):
//   A_real[i][2r] = Re Y[r][i], A_real[i][2r+1] = -Im Y[r][i];
//   B_real[2r + q][2j + p] = (Re, Im; -Im, Re)[q][p] of Y[r][j].
// The 32 x 64 real result is 8 tiles of 16 x 16; wave w owns row tile w & 1 and
// column tiles 2 (w >> 1) and 2 (w >> 1) + 1. Rows are staged through LDS in
// chunks of GCH.
static_assert(JP == 32, "the Gram and apply kernels are laid out for 32-column block pairs");
// blockIdx.z selects the problem of a batch (X at X + z sx).
__global__ __launch_bounds__(256) void k_jac_gram(int64_t p, int nblk, int r,
                                                  const double2* __restrict__ X,
                                                  double2* __restrict__ Gp, int nseg, int64_t sx) {
  __shared__ double2 Y[GCH][JP + 1];
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
  d4_t acc[2] = {d4_t{0, 0, 0, 0}, d4_t{0, 0, 0, 0}};
  // the next chunk is loaded into registers while the current one is multiplied
  constexpr int NL = GCH * JP / 256;
  double2 pre[NL];
  auto load_chunk = [&](int64_t c0) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int e = t + 256 * u;
      const int col = e / GCH, row = e % GCH;
      const int64_t grow = c0 + row;
      pre[u] = (grow < r1) ? X[pair_col(bp, bq, col) * p + grow] : make_double2(0.0, 0.0);
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
      const double2 ya = Y[row][i0 + l16];
      const double a = q ? -ya.y : ya.x;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double2 yb = Y[row][8 * (jt0 + h) + (l16 >> 1)];
        const double b = (q == pb) ? yb.x : (q ? -yb.y : yb.y);
        acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[h], 0, 0, 0);
      }
    }
  }
  // result element v of tile h: complex row i0 + kk + 4v, real column l16 of the
  // tile (complex column 8 (jt0 + h) + l16 / 2, real part for even l16); the
  // partner lane l16 ^ 1 holds the other part
  double2* out = Gp + (((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * nseg + seg) * JP * JP;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    double mine[4], other[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      mine[v] = acc[h][v];
      other[v] = __shfl_xor(mine[v], 1);
    }
    const int j = 8 * (jt0 + h) + (l16 >> 1);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      // even lanes store elements 0, 1 and odd lanes elements 2, 3 (constant indices)
      const int i = i0 + kk + 4 * (2 * pb + u);
      const double re = pb ? other[2 + u] : mine[u];
      const double im = pb ? mine[2 + u] : other[u];
      out[i * JP + j] = make_double2(re, im);
    }
  }
}

// one cyclic Jacobi sweep on the pair's Gram matrix (sum of the partials,
// fixed order); writes the accumulated unitary J and whether any rotation
// was needed (one workgroup per pair). A step of the sweep rotates JB
// disjoint index pairs (i_k, j_k) at once: with R_k = [[cs, s], [-conj(s), cs]]
// acting on columns (i_k, j_k), G <- R^H G R is done as independent 2 x 2
// block updates (rows of pair k, columns of pair l; one thread per (k, l))
// and J <- J R as independent column-pair updates, so a step needs one
// barrier after the rotation parameters and one after the updates.
static_assert(JB * JB == 256, "the solve maps one thread to each (pair, pair) block of G");
__global__ __launch_bounds__(256) void k_jac_solve(int nblk, const double2* __restrict__ Gp,
                                                   int nseg, double tol, double big,
                                                   double2* __restrict__ Jg,
                                                   int* __restrict__ rotflag,
                                                   int* __restrict__ counters) {
  __shared__ double2 G[JP][JP + 1];
  __shared__ double2 J[JP][JP + 1];
  __shared__ double rc[JB], rsr[JB], rsi[JB];
  __shared__ int rot_any, rot_big;
  const int t = threadIdx.x;
  // pair blockIdx.x of problem blockIdx.y
  const int64_t pidx = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
  const double2* gsrc = Gp + pidx * nseg * JP * JP;
  {
    // sum of the partial Gram matrices in a fixed order; the four elements of a
    // thread and several partials are loaded before they are added
    constexpr int NE = JP * JP / 256;
    double sr[NE], si[NE];
#pragma unroll
    for (int u = 0; u < NE; ++u) { sr[u] = 0; si[u] = 0; }
#pragma unroll 4
    for (int s = 0; s < nseg; ++s) {
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        const double2 v = gsrc[(int64_t)s * JP * JP + t + 256 * u];
        sr[u] += v.x;
        si[u] += v.y;
      }
    }
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = t + 256 * u;
      G[e / JP][e % JP] = make_double2(sr[u], si[u]);
      J[e / JP][e % JP] = make_double2((e / JP) == (e % JP) ? 1.0 : 0.0, 0.0);
    }
  }
  if (t == 0) { rot_any = 0; rot_big = 0; }
  __syncthreads();
  const int bk = t / JB, bl = t % JB;     // the (row pair, column pair) block of this thread
  for (int step = 0; step < JP - 1; ++step) {
    if (t < JB) {
      int i, j;
      round_pair(JP, step, t, i, j);
      const double a = G[i][i].x, b = G[j][j].x;
      const double2 c = G[i][j];
      // |c| > tol sqrt(a b) compared in squares (Gram entries already hold
      // squared column scales); one reciprocal square root gives 1/|c|
      const double ac2 = c.x * c.x + c.y * c.y;
      const double ab = fmax(a, 0.0) * fmax(b, 0.0);
      double cs = 1.0, sr = 0.0, si = 0.0;
      if (ac2 > 0.0 && ac2 > (tol * tol) * ab) {
        const double iac = rsqrt(ac2);
        const double zeta = (b - a) * (0.5 * iac);
        const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        cs = rsqrt(1.0 + tt * tt);
        const double sn = cs * tt;
        sr = sn * (c.x * iac);   // s = sn e^{i phi}
        si = sn * (c.y * iac);
        rot_any = 1;
        if (ac2 > (big * big) * ab) rot_big = 1;
      }
      rc[t] = cs; rsr[t] = sr; rsi[t] = si;
    }
    __syncthreads();
    {
      int ik, jk, il, jl;
      round_pair(JP, step, bk, ik, jk);
      round_pair(JP, step, bl, il, jl);
      const double ck = rc[bk], skr = rsr[bk], ski = rsi[bk];
      const double cl = rc[bl], slr = rsr[bl], sli = rsi[bl];
      const double2 b00 = G[ik][il], b01 = G[ik][jl], b10 = G[jk][il], b11 = G[jk][jl];
      // columns: x' = cl x - conj(sl) y, y' = sl x + cl y (x = column il, y = column jl)
      double2 c00, c01, c10, c11;
      c00 = make_double2(cl * b00.x - (slr * b01.x + sli * b01.y), cl * b00.y - (slr * b01.y - sli * b01.x));
      c10 = make_double2(cl * b10.x - (slr * b11.x + sli * b11.y), cl * b10.y - (slr * b11.y - sli * b11.x));
      c01 = make_double2((slr * b00.x - sli * b00.y) + cl * b01.x, (slr * b00.y + sli * b00.x) + cl * b01.y);
      c11 = make_double2((slr * b10.x - sli * b10.y) + cl * b11.x, (slr * b10.y + sli * b10.x) + cl * b11.y);
      // rows: row_i' = ck row_i - sk row_j, row_j' = conj(sk) row_i + ck row_j
      G[ik][il] = make_double2(ck * c00.x - (skr * c10.x - ski * c10.y), ck * c00.y - (skr * c10.y + ski * c10.x));
      G[ik][jl] = make_double2(ck * c01.x - (skr * c11.x - ski * c11.y), ck * c01.y - (skr * c11.y + ski * c11.x));
      G[jk][il] = make_double2((skr * c00.x + ski * c00.y) + ck * c10.x, (skr * c00.y - ski * c00.x) + ck * c10.y);
      G[jk][jl] = make_double2((skr * c01.x + ski * c01.y) + ck * c11.x, (skr * c01.y - ski * c01.x) + ck * c11.y);
    }
    // J <- J R: column pair bl for rows 2*bk and 2*bk + 1
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int row = 2 * bk + h;
      int i, j;
      round_pair(JP, step, bl, i, j);
      const double cs = rc[bl], sr = rsr[bl], si = rsi[bl];
      const double2 x = J[row][i], y = J[row][j];
      J[row][i] = make_double2(cs * x.x - (sr * y.x + si * y.y), cs * x.y - (sr * y.y - si * y.x));
      J[row][j] = make_double2((sr * x.x - si * x.y) + cs * y.x, (sr * x.y + si * x.x) + cs * y.y);
    }
    __syncthreads();
  }
  double2* jout = Jg + pidx * JP * JP;
  for (int e = t; e < JP * JP; e += 256) jout[e] = J[e / JP][e % JP];
  if (t == 0) {
    rotflag[pidx] = rot_any;
    if (rot_big) atomicAdd(counters, 1);
  }
}

// Y <- Y J (pass blockIdx.z & 1 = 0, rows of X) or V <- V J (pass 1) for AROWS rows
// per workgroup, 16 per wave, on FP64 MFMA: the rows' 32 complex entries are
// the A operand (real k = 2 column + part, the interleaved complex layout) and
// J's real embedding [[Re, Im], [-Im, Re]] the B operand, read from LDS.
// Every wave loads all A operands of its rows before it stores any result.
// Problem blockIdx.z >> 1 of a batch: X (p x qp) and V (qp x qp) stacked.
__global__ __launch_bounds__(256) void k_jac_apply(int64_t p, int64_t qp, int nblk, int r,
                                                   double2* __restrict__ X,
                                                   double2* __restrict__ V,
                                                   const double2* __restrict__ Jg,
                                                   const int* __restrict__ rotflag) {
  __shared__ double2 J[JP][JP + 1];
  const int64_t prob = blockIdx.z >> 1;
  const int64_t pidx = prob * gridDim.y + blockIdx.y;
  if (!rotflag[pidx]) return;
  const int pass = blockIdx.z & 1;
  double2* M = pass ? V + prob * qp * qp : X + prob * p * qp;
  const int64_t ld = pass ? qp : p;
  const int64_t rowblk = (int64_t)blockIdx.x * AROWS;
  if (rowblk >= ld) return;
  int bp, bq;
  round_pair(nblk, r, blockIdx.y, bp, bq);
  const int t = threadIdx.x;
  const double2* jsrc = Jg + pidx * JP * JP;
  for (int e = t; e < JP * JP; e += 256) J[e / JP][e % JP] = jsrc[e];
  __syncthreads();
  const int lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, kk = lane >> 4;
  const int q = kk & 1, pb = l16 & 1;
  const int64_t row0 = rowblk + 16 * wave;
  const int64_t arow = row0 + l16;
  const bool aok = arow < ld;
  double a[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int kc = 2 * s + (kk >> 1);
    const double2 y = aok ? M[pair_col(bp, bq, kc) * ld + arow] : make_double2(0.0, 0.0);
    a[s] = q ? y.y : y.x;
  }
  d4_t acc[4];
#pragma unroll
  for (int tile = 0; tile < 4; ++tile) acc[tile] = d4_t{0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int kc = 2 * s + (kk >> 1);
#pragma unroll
    for (int tile = 0; tile < 4; ++tile) {
      const double2 jv = J[kc][8 * tile + (l16 >> 1)];
      const double b = (q == pb) ? jv.x : (q ? -jv.y : jv.y);
      acc[tile] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b, acc[tile], 0, 0, 0);
    }
  }
#pragma unroll
  for (int tile = 0; tile < 4; ++tile) {
    double mine[4], other[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      mine[v] = acc[tile][v];
      other[v] = __shfl_xor(mine[v], 1);
    }
    double2* col = M + pair_col(bp, bq, 8 * tile + (l16 >> 1)) * ld;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int64_t row = row0 + kk + 4 * (2 * pb + u);
      const double re = pb ? other[2 + u] : mine[u];
      const double im = pb ? mine[2 + u] : other[u];
      if (row < ld) col[row] = make_double2(re, im);
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

// descending order of the q column norms (stable: ties keep column order),
// one workgroup; rank of column j = #{i : s_i > s_j} + #{i < j : s_i == s_j}.
// A non-finite norm is ordered as -1 (norms are >= 0) so that the ranks stay a
// permutation; the flag reports it. One workgroup per problem of a batch
// (nrm and perm with stride sn, s_out with stride ss).
__global__ __launch_bounds__(256) void k_rank(int64_t q, const double* __restrict__ nrm,
                                              int64_t* __restrict__ perm, double* __restrict__ s_out,
                                              int* __restrict__ nonfinite, int64_t sn, int64_t ss) {
  nrm += blockIdx.x * sn;
  perm += blockIdx.x * sn;
  s_out += blockIdx.x * ss;
  for (int64_t j = threadIdx.x; j < q; j += 256) {
    const double sj = nrm[j];
    if (!isfinite(sj)) atomicOr(nonfinite, 1);
    const double kj = isfinite(sj) ? sj : -1.0;
    int64_t rank = 0;
    for (int64_t i = 0; i < q; ++i) {
      const double ki = isfinite(nrm[i]) ? nrm[i] : -1.0;
      rank += (ki > kj) || (ki == kj && i < j);
    }
    perm[rank] = j;
    s_out[rank] = sj;
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

__global__ void k_svd_info(int32_t* info, int sweeps, int notconv, const int* counters) {
  if (threadIdx.x == 0) {
    info[0] = sweeps;
    info[1] = notconv;
    info[2] = counters[1];
    info[3] = 0;
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
  size_t bt, fact, fact2, rh, x, v, ub, vb, gp, jg, rflag, nrm, perm, cnt, total;
};

// every per-problem region holds `batch` consecutive copies
static SvdLayout svd_layout(int64_t m, int64_t n, int64_t batch) {
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
  L.total = off;
  return L;
}

size_t svd_ws_bytes(int64_t m, int64_t n, int64_t batch) { return svd_layout(m, n, batch).total; }

// grid width of the grid-stride helper kernels for `total` elements per problem
static inline unsigned gs_blocks(int64_t total) {
  const int64_t b = (total + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

// Thin SVD of `batch` m x n matrices of identical shape (problem b at a + b sa,
// u + b su, s + b ss, vh + b svh). The Jacobi sweeps run on all problems at
// once and stop together: a sweep in which no problem needed a rotation above
// `big` is followed by one final sweep. info (if given) receives the number of
// sweeps, whether the sweep limit was hit, and whether any singular value was
// not finite.
int svd_run(int64_t m, int64_t n, const void* a, int64_t lda, void* u, int64_t ldu, double* s,
            void* vh, int64_t ldvh, int32_t* info, void* ws, size_t ws_bytes, hipStream_t st,
            int64_t batch, int64_t sa, int64_t su, int64_t ss, int64_t svh) {
  if (m <= 0 || n <= 0 || lda < n) return PTN_ERR_ARG;
  const int64_t k = m < n ? m : n;
  if (ldu < k || ldvh < n) return PTN_ERR_ARG;
  if (batch < 1 || batch > 32767) return PTN_ERR_ARG;
  const SvdLayout L = svd_layout(m, n, batch);
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
  double2* Gp = (double2*)(base + L.gp);
  double2* Jg = (double2*)(base + L.jg);
  int* rflag = (int*)(base + L.rflag);
  double* nrm = (double*)(base + L.nrm);
  int64_t* perm = (int64_t*)(base + L.perm);
  int* counters = (int*)(base + L.cnt);   // [0] rotations in the current sweep, [1] non-finite
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
  k_eye<<<dim3(gs_blocks(L.qp * L.qp), NB), 256, 0, st>>>(L.qp, V);
  PTN_LAUNCH_CHECK();
  int* host_rot = pinned_counter();
  if (!host_rot) return PTN_ERR_HIP;
  const double tol = DBL_EPSILON * sqrt((double)L.q);
  // a sweep without rotations above this relative size is followed by one
  // final sweep (quadratic convergence takes it to tol) instead of repeating
  // until rounding noise stops triggering tiny rotations
  const double big = 1e-7;
  const int max_sweeps = 60;
  int sweeps = 0, converged = (L.q < 2), final_sweep = 0;
  k_set_ints<<<1, 64, 0, st>>>(counters, 2, 0);
  while (!converged && sweeps < max_sweeps) {
    k_set_ints<<<1, 64, 0, st>>>(counters, 1, 0);
    for (int r = 0; r < L.nblk - 1; ++r) {
      k_jac_gram<<<dim3((unsigned)L.nseg, (unsigned)npairs, NB), 256, 0, st>>>(
          L.q, L.nblk, r, X, Gp, (int)L.nseg, sx);
      k_jac_solve<<<dim3((unsigned)npairs, NB), 256, 0, st>>>(L.nblk, Gp, (int)L.nseg, tol, big,
                                                              Jg, rflag, counters);
      k_jac_apply<<<dim3((unsigned)aseg, (unsigned)npairs, 2 * NB), 256, 0, st>>>(
          L.q, L.qp, L.nblk, r, X, V, Jg, rflag);
    }
    PTN_LAUNCH_CHECK();
    PTN_HIP_CHECK(hipMemcpyAsync(host_rot, counters, sizeof(int), hipMemcpyDeviceToHost, st));
    PTN_HIP_CHECK(hipStreamSynchronize(st));
    ++sweeps;
    if (final_sweep) converged = 1;
    else if (*host_rot == 0) final_sweep = 1;
  }
  // 3. B = Q1 R2^H Q2^H = (Q1 [U_X; 0]) diag(sigma) (Q2 V)^H with U_X = X / sigma
  k_colnorms<<<dim3((unsigned)L.q, NB), 256, 0, st>>>(L.q, X, nrm, sx, L.qp);
  k_embed_u<<<dim3(gs_blocks(L.p * L.q), NB), 256, 0, st>>>(L.p, L.q, X, nrm, UB, sx, L.qp);
  k_copy_v<<<dim3(gs_blocks(L.q * L.q), NB), 256, 0, st>>>(L.q, L.qp, V, VB);
  PTN_LAUNCH_CHECK();
  if ((stt = qr_apply_q(f, UB, L.p, L.q, false, st, L.p * L.q))) return stt;
  if ((stt = qr_apply_q(f2, VB, L.q, L.q, false, st, L.q * L.q))) return stt;
  k_rank<<<NB, 256, 0, st>>>(L.q, nrm, perm, s, counters + 1, L.qp, ss);
  const int64_t tot = (m * k > k * n) ? m * k : k * n;
  k_svd_out<<<dim3((unsigned)((tot + 255) / 256), NB), 256, 0, st>>>(
      m, n, k, adjoint, L.p, L.q, UB, VB, perm, (double2*)u, ldu, (double2*)vh, ldvh, L.qp, su,
      svh);
  if (info) k_svd_info<<<1, 64, 0, st>>>(info, sweeps, converged ? 0 : 1, counters);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

}  // namespace ptn
