// Householder QR of a complex128 matrix with LAPACK's conventions, so that Q
// and R agree with numpy.linalg.qr (zgeqrf + zungqr) to rounding, including
// the completion of Q for rank-deficient inputs:
//   reflector j from alpha = A[j, j], x = A[j+1:, j]:
//     xnorm = |x|; if xnorm == 0 and Im(alpha) == 0: tau = 0 (H = I)
//     else beta = -sign(Re alpha) |(alpha, x)|, tau = (beta - alpha) / beta,
//          v = [1, x / (alpha - beta)], R[j, j] = beta (real)
//   A <- H^H A with H = I - tau v v^H;   Q = H_0 H_1 ... H_{k-1} [I_k; 0].
//
// Blocked right-looking algorithm on a column-major copy (panels of QR_NB
// columns):
//   panel:    for each pair of columns j, j+1 one launch (k_qr_col2) computes
//             reflectors j and j+1 in every workgroup (the same deterministic
//             reductions) and applies them to one remaining column of the panel
//             per workgroup (k_qr_col for a last odd column);
//   WY form:  T from V^H V (k_qr_vtv, k_qr_tmat), H_j0 ... = I - V T V^H;
//   trailing: A <- A - V (T^H (V^H A)): V^H A as a split-K complex product on
//             MFMA, the sum of the K chunks and the small T^H product in
//             k_qr_reduce, and one accumulating MFMA product.
// Q (or Q times a given matrix) is formed panel by panel in reverse order
// with the same three steps and T in place of T^H.
#include <float.h>
#include "qr.h"

namespace ptn {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// sums of a and b over a 256-thread workgroup, returned to every thread; the
// order of the additions is fixed, so every workgroup that reduces the same
// data gets bitwise the same result
__device__ void block_sum2(double& a, double& b) {
  __shared__ double red[2][4];
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][w] = a; red[1][w] = b; }
  __syncthreads();
  a = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
  b = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  __syncthreads();
}

// sums of N values over a 256-thread workgroup (fixed order, as block_sum2)
template <int N> __device__ void block_sum(double (&v)[N]) {
  __shared__ double red[N][4];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    v[k] = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) red[k][w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
  __syncthreads();
}

// Householder reflector from alpha and |x| (LAPACK zlarfg conventions):
// H = I when x = 0 and alpha is real (tau = 0, R entry alpha, v = e_j);
// else beta = -sign(Re alpha) |(alpha, x)|, tau = (beta - alpha) / beta and
// v = [1, x / (alpha - beta)], held as the complex factor sc = 1 / (alpha - beta)
struct Refl {
  double2 tau, rd, sc;
};
__device__ __forceinline__ Refl make_refl(double2 alpha, double xnorm) {
  Refl h;
  if (xnorm == 0.0 && alpha.y == 0.0) {
    h.tau = make_double2(0.0, 0.0);
    h.rd = alpha;
    h.sc = make_double2(0.0, 0.0);
    return h;
  }
  const double ar = fabs(alpha.x), ai = fabs(alpha.y);
  const double mx = fmax(fmax(ar, ai), xnorm);
  const double nrm = mx * sqrt((ar / mx) * (ar / mx) + (ai / mx) * (ai / mx) +
                               (xnorm / mx) * (xnorm / mx));
  const double beta = -copysign(nrm, alpha.x);
  const double dr = alpha.x - beta, di = alpha.y;
  const double den = dr * dr + di * di;
  h.sc = make_double2(dr / den, -di / den);
  h.tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
  h.rd = make_double2(beta, 0.0);
  return h;
}
__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

__global__ void k_zero_u64(int64_t n, unsigned long long* __restrict__ p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

// mxb[b] = largest |Re|, |Im| of matrix b (m x n, lda, at a + b sa): the bit
// patterns of non-negative doubles order like the doubles, so an integer
// atomic max (one per workgroup) gives the float max (NaN entries are skipped
// by fmax)
__global__ __launch_bounds__(256) void k_maxabs(int64_t m, int64_t n, const double2* __restrict__ a,
                                                int64_t lda, int64_t sa,
                                                unsigned long long* __restrict__ mxb) {
  a += blockIdx.y * sa;
  double v = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m * n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = a[(t / n) * lda + t % n];
    v = fmax(v, fmax(fabs(x.x), fabs(x.y)));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  __shared__ double wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    if (v > 0.0) atomicMax(mxb + blockIdx.y, (unsigned long long)__double_as_longlong(v));
  }
}

// the exact power of two that brings the largest entry into [0.5, 1) when it
// lies outside [2^-256, 2^256], else 1 (no change in the usual range)
__device__ __forceinline__ double qr_scale_of(unsigned long long bits) {
  const double mx = __longlong_as_double((long long)bits);
  if (!(mx > 0.0) || !isfinite(mx) || (mx >= 0x1p-256 && mx <= 0x1p256)) return 1.0;
  int e;
  frexp(mx, &e);
  return ldexp(1.0, -e);
}

// column-major copy times the matrix's scale: W[j*m + i] = scl A[i*lda + j]
__global__ void k_to_colmajor(int64_t m, int64_t n, const double2* __restrict__ a, int64_t lda,
                              double2* __restrict__ w, int64_t sa, int64_t sw,
                              const unsigned long long* __restrict__ mxb, double* __restrict__ scl) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double sc = qr_scale_of(mxb[blockIdx.y]);
  if (t == 0) scl[blockIdx.y] = sc;
  if (t >= m * n) return;
  a += blockIdx.y * sa;
  w += blockIdx.y * sw;
  int64_t i = t / n, j = t % n;
  const double2 x = a[i * lda + j];
  w[j * m + i] = make_double2(x.x * sc, x.y * sc);
}

// reflector j, applied to column c = j + blockIdx.x (< pe) of the panel.
// Workgroup 0 stores v_j (with the unit and the zeros above), tau_j and R[j, j];
// column j of W itself is left as it is (rows > j are not read again).
__global__ __launch_bounds__(256) void k_qr_col(int64_t m, int64_t j, double2* __restrict__ W,
                                                double2* __restrict__ Vf,
                                                double2* __restrict__ tau,
                                                double2* __restrict__ rdiag, QrStrides bs) {
  W += blockIdx.y * bs.W;
  Vf += blockIdx.y * bs.V;
  tau += blockIdx.y * bs.k;
  rdiag += blockIdx.y * bs.k;
  const double2* xj = W + j * m;
  const double2 alpha = xj[j];
  double s = 0;
  for (int64_t i = j + 1 + threadIdx.x; i < m; i += 256) {
    const double2 x = xj[i];
    s += x.x * x.x + x.y * x.y;
  }
  double unused = 0;
  block_sum2(s, unused);
  const double xnorm = sqrt(s);
  double2* vj = Vf + j * m;
  if (xnorm == 0.0 && alpha.y == 0.0) {
    // H_j = I: the other columns are unchanged
    if (blockIdx.x == 0) {
      for (int64_t i = threadIdx.x; i < m; i += 256)
        vj[i] = make_double2(i == j ? 1.0 : 0.0, 0.0);
      if (threadIdx.x == 0) {
        tau[j] = make_double2(0.0, 0.0);
        rdiag[j] = alpha;
      }
    }
    return;
  }
  // |(alpha_r, alpha_i, xnorm)| with scaling against overflow
  const double ar = fabs(alpha.x), ai = fabs(alpha.y);
  const double mx = fmax(fmax(ar, ai), xnorm);
  const double nrm = mx * sqrt((ar / mx) * (ar / mx) + (ai / mx) * (ai / mx) +
                               (xnorm / mx) * (xnorm / mx));
  const double beta = -copysign(nrm, alpha.x);
  // v_i = x_i / (alpha - beta) for i > j
  const double dr = alpha.x - beta, di = alpha.y;
  const double den = dr * dr + di * di;
  const double sr = dr / den, si = -di / den;
  const double2 t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
  if (blockIdx.x == 0) {
    for (int64_t i = threadIdx.x; i < m; i += 256) {
      double2 v;
      if (i < j) v = make_double2(0.0, 0.0);
      else if (i == j) v = make_double2(1.0, 0.0);
      else { const double2 x = xj[i]; v = make_double2(x.x * sr - x.y * si, x.x * si + x.y * sr); }
      vj[i] = v;
    }
    if (threadIdx.x == 0) {
      tau[j] = t;
      rdiag[j] = make_double2(beta, 0.0);
    }
    return;
  }
  // column c: w = v^H x, x <- x - conj(tau) w v (rows >= j)
  double2* xc = W + (j + blockIdx.x) * m;
  double wr = 0, wi = 0;
  for (int64_t i = j + threadIdx.x; i < m; i += 256) {
    double2 v;
    if (i == j) v = make_double2(1.0, 0.0);
    else { const double2 x = xj[i]; v = make_double2(x.x * sr - x.y * si, x.x * si + x.y * sr); }
    const double2 xi = xc[i];
    wr += v.x * xi.x + v.y * xi.y;
    wi += v.x * xi.y - v.y * xi.x;
  }
  block_sum2(wr, wi);
  const double fr = t.x * wr + t.y * wi, fi = t.x * wi - t.y * wr;   // conj(tau) w
  for (int64_t i = j + threadIdx.x; i < m; i += 256) {
    double2 v;
    if (i == j) v = make_double2(1.0, 0.0);
    else { const double2 x = xj[i]; v = make_double2(x.x * sr - x.y * si, x.x * si + x.y * sr); }
    double2 xi = xc[i];
    xi.x -= v.x * fr - v.y * fi;
    xi.y -= v.x * fi + v.y * fr;
    xc[i] = xi;
  }
}

// reflectors j and j+1 in one launch; workgroup b handles column c = j + b.
// Every workgroup forms reflector j from column j, applies it to column j+1
// on the fly (y = H_j^H x_{j+1}, never stored) and forms reflector j+1 from y,
// all with the same deterministic reductions. Workgroup 0 stores v_j, tau_j
// and R[j][j]; workgroup 1 stores v_{j+1}, tau_{j+1}, R[j+1][j+1] and R[j][j+1]
// (into rsup: columns j and j+1 of W are read by every workgroup during the
// launch and are not written); workgroups b >= 2 apply H_j^H then H_{j+1}^H to
// their column.
__global__ __launch_bounds__(256) void k_qr_col2(int64_t m, int64_t j, double2* __restrict__ W,
                                                 double2* __restrict__ Vf,
                                                 double2* __restrict__ tau,
                                                 double2* __restrict__ rdiag,
                                                 double2* __restrict__ rsup, QrStrides bs) {
  W += blockIdx.y * bs.W;
  Vf += blockIdx.y * bs.V;
  tau += blockIdx.y * bs.k;
  rdiag += blockIdx.y * bs.k;
  rsup += blockIdx.y * bs.k;
  const double2* xj = W + j * m;
  const double2* xk = W + (j + 1) * m;
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  double n1[1] = {0};
  for (int64_t i = j + 1 + t; i < m; i += 256) {
    const double2 x = xj[i];
    n1[0] += x.x * x.x + x.y * x.y;
  }
  block_sum<1>(n1);
  const Refl h1 = make_refl(xj[j], sqrt(n1[0]));
  auto v1 = [&](int64_t i) -> double2 {
    return i == j ? make_double2(1.0, 0.0) : cmul2(xj[i], h1.sc);   // i >= j
  };
  if (b == 0) {
    double2* vj = Vf + j * m;
    for (int64_t i = t; i < m; i += 256) vj[i] = (i < j) ? make_double2(0.0, 0.0) : v1(i);
    if (t == 0) { tau[j] = h1.tau; rdiag[j] = h1.rd; }
    return;
  }
  double2* xc = W + (j + b) * m;
  const bool own = b >= 2;
  // w1 = v1^H x_{j+1}, wc = v1^H x_c (rows >= j)
  double d2[4] = {0, 0, 0, 0};
  for (int64_t i = j + t; i < m; i += 256) {
    const double2 v = v1(i), x = xk[i];
    d2[0] += v.x * x.x + v.y * x.y;
    d2[1] += v.x * x.y - v.y * x.x;
    if (own) {
      const double2 z = xc[i];
      d2[2] += v.x * z.x + v.y * z.y;
      d2[3] += v.x * z.y - v.y * z.x;
    }
  }
  block_sum<4>(d2);
  const double2 f1 = cmul2(conj2(h1.tau), make_double2(d2[0], d2[1]));
  const double2 fc = cmul2(conj2(h1.tau), make_double2(d2[2], d2[3]));
  auto y = [&](int64_t i) -> double2 {   // H_j^H x_{j+1}, rows >= j
    const double2 v = v1(i), g = cmul2(f1, v), x = xk[i];
    return make_double2(x.x - g.x, x.y - g.y);
  };
  auto z = [&](int64_t i) -> double2 {   // H_j^H x_c, rows >= j
    const double2 v = v1(i), g = cmul2(fc, v), x = xc[i];
    return make_double2(x.x - g.x, x.y - g.y);
  };
  // |y| below row j+1, and y^H z below row j+1
  double d3[3] = {0, 0, 0};
  for (int64_t i = j + 2 + t; i < m; i += 256) {
    const double2 yi = y(i);
    d3[0] += yi.x * yi.x + yi.y * yi.y;
    if (own) {
      const double2 zi = z(i);
      d3[1] += yi.x * zi.x + yi.y * zi.y;
      d3[2] += yi.x * zi.y - yi.y * zi.x;
    }
  }
  block_sum<3>(d3);
  const double2 alpha2 = y(j + 1);
  const Refl h2 = make_refl(alpha2, sqrt(d3[0]));
  auto v2 = [&](int64_t i) -> double2 {   // i >= j+1
    return i == j + 1 ? make_double2(1.0, 0.0) : cmul2(y(i), h2.sc);
  };
  if (b == 1) {
    double2* vk = Vf + (j + 1) * m;
    for (int64_t i = t; i < m; i += 256) vk[i] = (i <= j) ? make_double2(0.0, 0.0) : v2(i);
    if (t == 0) { tau[j + 1] = h2.tau; rdiag[j + 1] = h2.rd; rsup[j] = y(j); }
    return;
  }
  // w2 = v2^H z = z_{j+1} + conj(sc2) (y^H z)_{> j+1};  x_c <- z - conj(tau2) w2 v2
  const double2 zk = (j + 1 < m) ? z(j + 1) : make_double2(0.0, 0.0);
  const double2 yz = cmul2(conj2(h2.sc), make_double2(d3[1], d3[2]));
  const double2 f2 = cmul2(conj2(h2.tau), make_double2(zk.x + yz.x, zk.y + yz.y));
  __syncthreads();   // every thread has read row j+1 of x_c (zk) before it is overwritten
  // every row is read into registers before the row is written (same thread)
  for (int64_t i = j + t; i < m; i += 256) {
    double2 o = z(i);
    if (i >= j + 1) {
      const double2 g = cmul2(f2, v2(i));
      o.x -= g.x;
      o.y -= g.y;
    }
    xc[i] = o;
  }
}

// k_qr_col2 with the rows j .. m-1 of the three columns held in registers
// (row j + t + 256 u in thread t, u < R; requires m - j <= 256 R): one load of
// each column instead of one per pass. Every workgroup of a launch runs the
// same reductions in the same order, so the reflectors agree bitwise across
// workgroups (the grouping of rows per thread differs from k_qr_col2, so the
// two kernels agree to rounding).
template <int R>
__global__ __launch_bounds__(256) void k_qr_col2r(int64_t m, int64_t j, double2* __restrict__ W,
                                                  double2* __restrict__ Vf,
                                                  double2* __restrict__ tau,
                                                  double2* __restrict__ rdiag,
                                                  double2* __restrict__ rsup, QrStrides bs) {
  W += blockIdx.y * bs.W;
  Vf += blockIdx.y * bs.V;
  tau += blockIdx.y * bs.k;
  rdiag += blockIdx.y * bs.k;
  rsup += blockIdx.y * bs.k;
  const double2* xj = W + j * m;
  const double2* xk = W + (j + 1) * m;
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  double2* xc = W + (j + b) * m;
  const bool own = b >= 2;
  double2 rj[R], rk[R], rc[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int64_t i = j + t + 256 * u;
    const bool in = i < m;
    rj[u] = in ? xj[i] : make_double2(0.0, 0.0);
    rk[u] = (in && b >= 1) ? xk[i] : make_double2(0.0, 0.0);
    rc[u] = (in && own) ? xc[i] : make_double2(0.0, 0.0);
  }
  const double2 aj = xj[j];
  double n1[1] = {0};
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int64_t i = j + t + 256 * u;
    if (i > j && i < m) n1[0] += rj[u].x * rj[u].x + rj[u].y * rj[u].y;
  }
  block_sum<1>(n1);
  const Refl h1 = make_refl(aj, sqrt(n1[0]));
  auto v1r = [&](int u, int64_t i) -> double2 {
    return i == j ? make_double2(1.0, 0.0) : cmul2(rj[u], h1.sc);
  };
  if (b == 0) {
    double2* vj = Vf + j * m;
    for (int64_t i = t; i < j; i += 256) vj[i] = make_double2(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int64_t i = j + t + 256 * u;
      if (i < m) vj[i] = v1r(u, i);
    }
    if (t == 0) { tau[j] = h1.tau; rdiag[j] = h1.rd; }
    return;
  }
  double d2[4] = {0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int64_t i = j + t + 256 * u;
    if (i < m) {
      const double2 v = v1r(u, i), x = rk[u];
      d2[0] += v.x * x.x + v.y * x.y;
      d2[1] += v.x * x.y - v.y * x.x;
      if (own) {
        const double2 z = rc[u];
        d2[2] += v.x * z.x + v.y * z.y;
        d2[3] += v.x * z.y - v.y * z.x;
      }
    }
  }
  block_sum<4>(d2);
  const double2 f1 = cmul2(conj2(h1.tau), make_double2(d2[0], d2[1]));
  const double2 fc = cmul2(conj2(h1.tau), make_double2(d2[2], d2[3]));
  double2 y[R], z[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int64_t i = j + t + 256 * u;
    const double2 v = v1r(u, i), g = cmul2(f1, v), gc = cmul2(fc, v);
    y[u] = make_double2(rk[u].x - g.x, rk[u].y - g.y);
    z[u] = make_double2(rc[u].x - gc.x, rc[u].y - gc.y);
  }
  double d3[3] = {0, 0, 0};
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int64_t i = j + t + 256 * u;
    if (i >= j + 2 && i < m) {
      d3[0] += y[u].x * y[u].x + y[u].y * y[u].y;
      if (own) {
        d3[1] += y[u].x * z[u].x + y[u].y * z[u].y;
        d3[2] += y[u].x * z[u].y - y[u].y * z[u].x;
      }
    }
  }
  block_sum<3>(d3);
  // rows j and j+1 of y and z from memory (every thread needs them)
  auto yg = [&](int64_t i) -> double2 {
    const double2 v = i == j ? make_double2(1.0, 0.0) : cmul2(xj[i], h1.sc), g = cmul2(f1, v),
                  x = xk[i];
    return make_double2(x.x - g.x, x.y - g.y);
  };
  const double2 alpha2 = yg(j + 1);
  const Refl h2 = make_refl(alpha2, sqrt(d3[0]));
  if (b == 1) {
    double2* vk = Vf + (j + 1) * m;
    for (int64_t i = t; i <= j; i += 256) vk[i] = make_double2(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int64_t i = j + t + 256 * u;
      if (i > j && i < m) vk[i] = i == j + 1 ? make_double2(1.0, 0.0) : cmul2(y[u], h2.sc);
    }
    if (t == 0) { tau[j + 1] = h2.tau; rdiag[j + 1] = h2.rd; rsup[j] = yg(j); }
    return;
  }
  // w2 = v2^H z = z_{j+1} + conj(sc2) (y^H z)_{> j+1};  x_c <- z - conj(tau2) w2 v2
  double2 zk = make_double2(0.0, 0.0);
  if (j + 1 < m) {
    const double2 v = cmul2(xj[j + 1], h1.sc), g = cmul2(fc, v), x = xc[j + 1];
    zk = make_double2(x.x - g.x, x.y - g.y);
  }
  const double2 yz = cmul2(conj2(h2.sc), make_double2(d3[1], d3[2]));
  const double2 f2 = cmul2(conj2(h2.tau), make_double2(zk.x + yz.x, zk.y + yz.y));
  __syncthreads();   // every thread has read row j+1 of x_c (zk) before it is overwritten
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int64_t i = j + t + 256 * u;
    if (i < m) {
      double2 o = z[u];
      if (i >= j + 1) {
        const double2 v2 = i == j + 1 ? make_double2(1.0, 0.0) : cmul2(y[u], h2.sc);
        const double2 g = cmul2(f2, v2);
        o.x -= g.x;
        o.y -= g.y;
      }
      xc[i] = o;
    }
  }
}

// R[jj][jj+1] of the two-reflector launches of a panel, written after the panel
__global__ void k_qr_fixsup(int64_t m, int64_t j0, int64_t pe, const double2* __restrict__ rsup,
                            double2* __restrict__ W, QrStrides bs) {
  rsup += blockIdx.y * bs.k;
  W += blockIdx.y * bs.W;
  const int64_t jj = j0 + 2 * (int64_t)threadIdx.x;
  if (jj + 1 < pe) W[(jj + 1) * m + jj] = rsup[jj];
}

// G[a][b] = v_a^H v_b for the panel's reflectors (rows j0 .. m-1), one workgroup per entry
__global__ __launch_bounds__(256) void k_qr_vtv(int64_t m, int64_t j0, int nbp,
                                                const double2* __restrict__ Vf,
                                                double2* __restrict__ G, QrStrides bs) {
  Vf += blockIdx.y * bs.V;
  G += blockIdx.y * bs.G;
  const int a = blockIdx.x / nbp, b = blockIdx.x % nbp;
  if (b <= a) return;   // only the strict upper triangle is used
  const double2* va = Vf + (j0 + a) * m;
  const double2* vb = Vf + (j0 + b) * m;
  double sr = 0, si = 0;
  for (int64_t i = j0 + threadIdx.x; i < m; i += 256) {
    const double2 x = va[i], y = vb[i];
    sr += x.x * y.x + x.y * y.y;
    si += x.x * y.y - x.y * y.x;
  }
  block_sum2(sr, si);
  if (threadIdx.x == 0) G[a * QR_NB + b] = make_double2(sr, si);
}

// compact WY factor (forward, columnwise): T[i][i] = tau_i,
// T[0:i, i] = -tau_i T[0:i, 0:i] (V[:, 0:i]^H v_i); zero padded to QR_NB x QR_NB.
// Thread r builds row r of T in registers (it only needs its own row and G),
// so the column recurrence needs no barriers.
__global__ __launch_bounds__(64) void k_qr_tmat(int nbp, const double2* __restrict__ G,
                                                const double2* __restrict__ tau,
                                                double2* __restrict__ T, QrStrides bs) {
  G += blockIdx.x * bs.G;
  tau += blockIdx.x * bs.k;
  T += blockIdx.x * bs.T;
  __shared__ double2 Gs[QR_NB][QR_NB + 1];
  __shared__ double2 taus[QR_NB];
  const int r = threadIdx.x;
  for (int e = r; e < QR_NB * QR_NB; e += 64) {
    const int a = e / QR_NB, b = e % QR_NB;
    Gs[a][b] = (a < b && b < nbp) ? G[e] : make_double2(0.0, 0.0);
  }
  if (r < QR_NB) taus[r] = (r < nbp) ? tau[r] : make_double2(0.0, 0.0);
  __syncthreads();
  if (r >= QR_NB) return;
  double2 trow[QR_NB];
#pragma unroll
  for (int c = 0; c < QR_NB; ++c) trow[c] = make_double2(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < QR_NB; ++i) {
    if (i < nbp) {
      const double2 ti = taus[i];
      double ar = 0, ai = 0;
#pragma unroll
      for (int s = 0; s < i; ++s) {
        // T[r][s] is zero for s < r, so the full range gives the same sum
        const double2 a = trow[s], g = Gs[s][i];
        ar += a.x * g.x - a.y * g.y;
        ai += a.x * g.y + a.y * g.x;
      }
      if (r < i) trow[i] = make_double2(-(ti.x * ar - ti.y * ai), -(ti.x * ai + ti.y * ar));
      else if (r == i) trow[i] = ti;
    }
  }
#pragma unroll
  for (int c = 0; c < QR_NB; ++c) T[r * QR_NB + c] = trow[c];
}

// Y = sum of the nparts partial products P[s] (QR_NB x nc, rows < nbp used);
// Z = -T^H Y (adjoint = 1) or Z = -T Y (adjoint = 0). A workgroup handles 8
// columns; thread (r, cl) sums row r of column cl, then forms row r of Z.
static_assert(QR_NB * 8 == 256, "k_qr_reduce maps its 256 threads onto QR_NB rows x 8 columns");
__global__ __launch_bounds__(256) void k_qr_reduce(int nbp, int64_t nc, int nparts,
                                                   const double2* __restrict__ P,
                                                   const double2* __restrict__ T, int adjoint,
                                                   double2* __restrict__ Z, QrStrides bs) {
  P += blockIdx.y * bs.P;
  T += blockIdx.y * bs.T;
  Z += blockIdx.y * bs.Z;
  __shared__ double2 Ts[QR_NB][QR_NB + 1];
  __shared__ double2 Ys[8][QR_NB + 1];
  const int t = threadIdx.x;
  for (int e = t; e < QR_NB * QR_NB; e += 256) Ts[e / QR_NB][e % QR_NB] = T[e];
  const int cl = t & 7, r = t >> 3;
  const int64_t c = (int64_t)blockIdx.x * 8 + cl;
  const bool ok = c < nc && r < nbp;
  double sr = 0, si = 0;
  if (ok)
    for (int s = 0; s < nparts; ++s) {
      const double2 v = P[((int64_t)s * QR_NB + r) * nc + c];
      sr += v.x;
      si += v.y;
    }
  Ys[cl][r] = make_double2(sr, si);
  __syncthreads();
  if (!ok) return;
  double zr = 0, zi = 0;
  if (adjoint) {
    // (T^H y)_r = sum_{q <= r} conj(T[q][r]) y_q
    for (int q = 0; q <= r; ++q) {
      const double2 a = Ts[q][r], y = Ys[cl][q];
      zr += a.x * y.x + a.y * y.y;
      zi += a.x * y.y - a.y * y.x;
    }
  } else {
    // (T y)_r = sum_{q >= r} T[r][q] y_q
    for (int q = r; q < nbp; ++q) {
      const double2 a = Ts[r][q], y = Ys[cl][q];
      zr += a.x * y.x - a.y * y.y;
      zi += a.x * y.y + a.y * y.x;
    }
  }
  Z[(int64_t)r * nc + c] = make_double2(-zr, -zi);
}

static inline size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

static QrStrides qr_strides(int64_t m, int64_t n, int64_t ncmax) {
  const int64_t k = m < n ? m : n;
  const int64_t npan = (k + QR_NB - 1) / QR_NB;
  const int64_t nc = ncmax > n ? ncmax : n;
  QrStrides b;
  b.W = m * n;
  b.V = m * k;
  b.k = k + 1;
  b.T = (npan + 1) * QR_NB * QR_NB;
  b.G = QR_NB * QR_NB;
  b.P = (QR_SPLITK + 1) * QR_NB * nc;
  b.Z = QR_NB * nc;
  return b;
}

size_t qr_fact_bytes(int64_t m, int64_t n, int64_t ncmax, int64_t batch) {
  const QrStrides b = qr_strides(m, n, ncmax);
  const size_t B = (size_t)batch;
  return al(B * b.W * 16) + al(B * b.V * 16) + 3 * al(B * b.k * 16) + al(B * b.T * 16) +
         al(B * b.G * 16) + al(B * b.P * 16) + al(B * b.Z * 16) + 2 * al(B * 8);
}

QrFact qr_fact_layout(int64_t m, int64_t n, int64_t ncmax, void* ws, int64_t batch) {
  QrFact f;
  f.m = m; f.n = n; f.k = m < n ? m : n;
  f.npan = (f.k + QR_NB - 1) / QR_NB;
  f.ncmax = ncmax > n ? ncmax : n;
  f.batch = batch;
  f.bs = qr_strides(m, n, ncmax);
  const size_t B = (size_t)batch;
  char* p = (char*)ws;
  f.W = (double2*)p; p += al(B * f.bs.W * 16);
  f.Vf = (double2*)p; p += al(B * f.bs.V * 16);
  f.tau = (double2*)p; p += al(B * f.bs.k * 16);
  f.rdiag = (double2*)p; p += al(B * f.bs.k * 16);
  f.rsup = (double2*)p; p += al(B * f.bs.k * 16);
  f.T = (double2*)p; p += al(B * f.bs.T * 16);
  f.G = (double2*)p; p += al(B * f.bs.G * 16);
  f.P = (double2*)p; p += al(B * f.bs.P * 16);
  f.Z = (double2*)p; p += al(B * f.bs.Z * 16);
  f.scl = (double*)p; p += al(B * 8);
  f.mxb = (unsigned long long*)p;
  return f;
}

// C[j0:m, c0:c0+nc] += V_panel Z with Z = -T^(H) V_panel^H C[j0:m, c0:c0+nc]
// (C column-major with leading dimension ldc; panel columns j0 .. j0+nbp-1;
// matrix b of a batch at C + b * sc). A single matrix splits the K = m - j0
// rows of V^H C into chunks (batched product) for parallelism; a batch of
// matrices is batched over the matrices instead.
static int apply_block(const QrFact& f, int64_t j0, int nbp, const double2* Tp, int adjoint,
                       double2* C, int64_t ldc, int64_t c0, int64_t nc, hipStream_t s,
                       int64_t sc) {
  const int64_t m = f.m, mr = m - j0;
  if (nc <= 0 || mr <= 0) return PTN_OK;
  const double2* V = f.Vf + j0 * m + j0;
  double2* Cs = C + c0 * ldc + j0;
  int st;
  int nparts;
  if (f.batch == 1) {
    // partials P[s] = V^H C over K chunks of `chunk` rows
    int64_t chunk = (mr + QR_SPLITK - 1) / QR_SPLITK;
    chunk = ((chunk + 7) / 8) * 8;
    const int64_t nfull = mr / chunk, rem = mr - nfull * chunk;
    if (nfull > 0) {
      ZGemm g = {nbp, nc, chunk, V, m, 1, 1, Cs, 1, ldc, 0, f.P, nc, 1, 0,
                 nfull, chunk, chunk, (int64_t)QR_NB * nc};
      if ((st = zgemm(g, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, s))) return st;
    }
    if (rem > 0) {
      ZGemm g = {nbp, nc, rem, V + nfull * chunk, m, 1, 1, Cs + nfull * chunk, 1, ldc, 0,
                 f.P + nfull * QR_NB * nc, nc, 1, 0, 1, 0, 0, 0};
      if ((st = zgemm(g, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, s))) return st;
    }
    nparts = (int)(nfull + (rem > 0 ? 1 : 0));
  } else {
    ZGemm g = {nbp, nc, mr, V, m, 1, 1, Cs, 1, ldc, 0, f.P, nc, 1, 0,
               f.batch, f.bs.V, sc, f.bs.P};
    if ((st = zgemm(g, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, s))) return st;
    nparts = 1;
  }
  k_qr_reduce<<<dim3((unsigned)((nc + 7) / 8), (unsigned)f.batch), 256, 0, s>>>(
      nbp, nc, nparts, f.P, Tp, adjoint, f.Z, f.bs);
  PTN_LAUNCH_CHECK();
  // C += V Z
  ZGemm g = {mr, nc, nbp, V, 1, m, 0, f.Z, nc, 1, 0, Cs, 1, ldc, 1, f.batch, f.bs.V, f.bs.Z, sc};
  return zgemm(g, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, s);
}

int qr_factor(QrFact& f, const double2* a, int64_t lda, hipStream_t s, int64_t sa) {
  const int64_t m = f.m, n = f.n, k = f.k;
  if (m == 0 || n == 0) return PTN_OK;
  if (f.batch < 1 || f.batch > 65535) return PTN_ERR_ARG;
  const unsigned B = (unsigned)f.batch;
  k_zero_u64<<<(unsigned)((f.batch + 255) / 256), 256, 0, s>>>(f.batch, f.mxb);
  {
    int64_t nb = (m * n + 255) / 256;
    const int64_t cap = f.batch > 16 ? 16 : 256;   // workgroups per matrix (one atomic each)
    if (nb > cap) nb = cap;
    k_maxabs<<<dim3((unsigned)nb, B), 256, 0, s>>>(m, n, a, lda, sa, f.mxb);
  }
  k_to_colmajor<<<dim3((unsigned)((m * n + 255) / 256), B), 256, 0, s>>>(m, n, a, lda, f.W, sa,
                                                                           f.bs.W, f.mxb, f.scl);
  PTN_LAUNCH_CHECK();
  for (int64_t p = 0; p < f.npan; ++p) {
    const int64_t j0 = p * QR_NB;
    const int64_t pe = (j0 + QR_NB < k) ? j0 + QR_NB : k;
    const int nbp = (int)(pe - j0);
    int64_t j = j0;
    for (; j + 1 < pe; j += 2) {
      const dim3 g((unsigned)(pe - j), B);
      const int64_t rows = m - j;
      if (rows <= 256 * 4)
        k_qr_col2r<4><<<g, 256, 0, s>>>(m, j, f.W, f.Vf, f.tau, f.rdiag, f.rsup, f.bs);
      else if (rows <= 256 * 8)
        k_qr_col2r<8><<<g, 256, 0, s>>>(m, j, f.W, f.Vf, f.tau, f.rdiag, f.rsup, f.bs);
      else if (rows <= 256 * 16)
        k_qr_col2r<16><<<g, 256, 0, s>>>(m, j, f.W, f.Vf, f.tau, f.rdiag, f.rsup, f.bs);
      else
        k_qr_col2<<<g, 256, 0, s>>>(m, j, f.W, f.Vf, f.tau, f.rdiag, f.rsup, f.bs);
    }
    if (j < pe)
      k_qr_col<<<dim3((unsigned)(pe - j), B), 256, 0, s>>>(m, j, f.W, f.Vf, f.tau, f.rdiag, f.bs);
    if (pe - j0 >= 2) k_qr_fixsup<<<dim3(1, B), QR_NB / 2, 0, s>>>(m, j0, pe, f.rsup, f.W, f.bs);
    PTN_LAUNCH_CHECK();
    double2* Tp = f.T + p * QR_NB * QR_NB;
    k_qr_vtv<<<dim3((unsigned)(nbp * nbp), B), 256, 0, s>>>(m, j0, nbp, f.Vf, f.G, f.bs);
    k_qr_tmat<<<B, 64, 0, s>>>(nbp, f.G, f.tau + j0, Tp, f.bs);
    PTN_LAUNCH_CHECK();
    int st;
    if (pe < n && (st = apply_block(f, j0, nbp, Tp, 1, f.W, m, pe, n - pe, s, f.bs.W))) return st;
  }
  return PTN_OK;
}

int qr_apply_q(const QrFact& f, double2* C, int64_t ldc, int64_t nc, bool identity_tail,
               hipStream_t s, int64_t sc) {
  if (nc > f.ncmax) return PTN_ERR_ARG;
  for (int64_t p = f.npan - 1; p >= 0; --p) {
    const int64_t j0 = p * QR_NB;
    const int64_t pe = (j0 + QR_NB < f.k) ? j0 + QR_NB : f.k;
    const int64_t c0 = identity_tail ? j0 : 0;
    int st;
    if ((st = apply_block(f, j0, (int)(pe - j0), f.T + p * QR_NB * QR_NB, 0, C, ldc, c0, nc - c0,
                          s, sc)))
      return st;
  }
  return PTN_OK;
}

// column-major m x k identity columns (one matrix per blockIdx.y, stride m * k)
__global__ void k_q_init(int64_t m, int64_t k, double2* __restrict__ Qc) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * k) return;
  Qc += blockIdx.y * m * k;
  int64_t c = t / m, i = t % m;
  Qc[t] = make_double2(i == c ? 1.0 : 0.0, 0.0);
}

// Q (row-major, ldq) from Qc; R (row-major k x n, ldr) from the upper triangle of W
// with the diagonal from rdiag; matrix b of a batch at q + b * sq, r + b * sr
__global__ void k_qr_out(int64_t m, int64_t n, int64_t k, const double2* __restrict__ W,
                         const double2* __restrict__ rdiag, const double2* __restrict__ Qc,
                         double2* __restrict__ q, int64_t ldq, double2* __restrict__ r,
                         int64_t ldr, QrStrides bs, int64_t sq, int64_t sr,
                         const double* __restrict__ scl) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  const double inv = 1.0 / scl[b];   // exact: scl is a power of two
  W += b * bs.W;
  rdiag += b * bs.k;
  Qc += b * m * k;
  q += b * sq;
  r += b * sr;
  if (t < m * k) {
    int64_t i = t / k, c = t % k;
    q[i * ldq + c] = Qc[c * m + i];
  }
  if (t < k * n) {
    int64_t i = t / n, c = t % n;
    const double2 v = (i < c) ? W[c * m + i] : (i == c ? rdiag[i] : make_double2(0.0, 0.0));
    r[i * ldr + c] = make_double2(v.x * inv, v.y * inv);
  }
}

size_t qr_ws_bytes(int64_t m, int64_t n, int64_t batch) {
  int64_t k = m < n ? m : n;
  return al(qr_fact_bytes(m, n, n, batch)) + al((size_t)batch * m * k * 16);
}

// `batch` row-major m x n matrices stacked contiguously (a: lda = n; q: m x k,
// ldq = k; r: k x n, ldr = n) when batch > 1
int qr_run(int64_t m, int64_t n, const void* a, int64_t lda, void* q, int64_t ldq, void* r,
           int64_t ldr, void* ws, size_t ws_bytes, hipStream_t s, int64_t batch) {
  if (m < 0 || n < 0 || lda < n || ldq < (m < n ? m : n) || ldr < n) return PTN_ERR_ARG;
  if (batch < 1 || batch > 65535) return PTN_ERR_ARG;
  if (ws_bytes < qr_ws_bytes(m, n, batch)) return PTN_ERR_WORKSPACE;
  const int64_t k = m < n ? m : n;
  if (m == 0 || n == 0) return PTN_OK;
  QrFact f = qr_fact_layout(m, n, n, ws, batch);
  double2* Qc = (double2*)((char*)ws + al(qr_fact_bytes(m, n, n, batch)));
  int st;
  if ((st = qr_factor(f, (const double2*)a, lda, s, m * lda))) return st;
  const unsigned B = (unsigned)batch;
  k_q_init<<<dim3((unsigned)((m * k + 255) / 256), B), 256, 0, s>>>(m, k, Qc);
  PTN_LAUNCH_CHECK();
  if ((st = qr_apply_q(f, Qc, m, k, true, s, m * k))) return st;
  const int64_t tot = (m * k > k * n) ? m * k : k * n;
  k_qr_out<<<dim3((unsigned)((tot + 255) / 256), B), 256, 0, s>>>(
      m, n, k, f.W, f.rdiag, Qc, (double2*)q, ldq, (double2*)r, ldr, f.bs, m * ldq, k * ldr,
      f.scl);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

}  // namespace ptn
