// Householder QR of a complex128 matrix with LAPACK's conventions, so that Q
// and R agree with numpy.linalg.qr (zgeqrf + zungqr) to rounding, including
// the completion of Q for rank-deficient inputs:
//   reflector j from alpha = A[j, j], x = A[j+1:, j]:
//     xnorm = |x|; if xnorm == 0 and Im(alpha) == 0: tau = 0 (H = I)
//     else beta = -sign(Re alpha) |(alpha, x)|, tau = (beta - alpha) / beta,
//          v = [1, x / (alpha - beta)], R[j, j] = beta (real)
//   A <- H^H A with H = I - tau v v^H;   Q = H_0 H_1 ... H_{k-1} [I_k; 0].
// The matrix is factorised in a column-major copy so that every column
// operation is a contiguous vector operation. Column j costs two launches
// (reflector, trailing update); forming Q one launch per reflector.
#include <float.h>
#include "common.h"

namespace ptn {

__device__ double wg_sum(double v) {
  __shared__ double red[256];
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

// column-major copy: W[j*m + i] = A[i*lda + j]
__global__ void k_to_colmajor(int64_t m, int64_t n, const double2* __restrict__ a, int64_t lda,
                              double2* __restrict__ w) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * n) return;
  int64_t i = t / n, j = t % n;
  w[j * m + i] = a[i * lda + j];
}

// reflector for column j (one workgroup); tau[j] = (re, im)
__global__ __launch_bounds__(256) void k_qr_reflector(int64_t m, int64_t j, double2* __restrict__ W,
                                                      double2* __restrict__ tau) {
  double2* col = W + j * m;
  // read alpha before the barrier inside wg_sum: thread 0 overwrites it below
  const double2 alpha = col[j];
  double s = 0;
  for (int64_t i = j + 1 + threadIdx.x; i < m; i += 256) {
    double2 x = col[i];
    s += x.x * x.x + x.y * x.y;
  }
  const double xnorm = sqrt(wg_sum(s));
  if (xnorm == 0.0 && alpha.y == 0.0) {
    if (threadIdx.x == 0) tau[j] = make_double2(0.0, 0.0);
    return;
  }
  // |(alpha_r, alpha_i, xnorm)| with scaling against overflow
  const double ar = fabs(alpha.x), ai = fabs(alpha.y);
  const double mx = fmax(fmax(ar, ai), xnorm);
  const double nrm = mx * sqrt((ar / mx) * (ar / mx) + (ai / mx) * (ai / mx) +
                               (xnorm / mx) * (xnorm / mx));
  const double beta = -copysign(nrm, alpha.x);
  // scale factor 1 / (alpha - beta)
  const double dr = alpha.x - beta, di = alpha.y;
  const double den = dr * dr + di * di;
  const double sr = dr / den, si = -di / den;
  for (int64_t i = j + 1 + threadIdx.x; i < m; i += 256) {
    double2 x = col[i];
    col[i] = make_double2(x.x * sr - x.y * si, x.x * si + x.y * sr);
  }
  if (threadIdx.x == 0) {
    tau[j] = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    col[j] = make_double2(beta, 0.0);
  }
}

// apply H_j^H = I - conj(tau_j) v v^H to columns c = j+1 .. n-1 (one workgroup per column)
__global__ __launch_bounds__(256) void k_qr_apply(int64_t m, int64_t n, int64_t j,
                                                  double2* __restrict__ W,
                                                  const double2* __restrict__ tau) {
  const int64_t c = j + 1 + blockIdx.x;
  if (c >= n) return;
  const double2 t = tau[j];
  if (t.x == 0.0 && t.y == 0.0) return;
  const double2* v = W + j * m;
  double2* x = W + c * m;
  // w = v^H x with v_j = 1
  double wr = 0, wi = 0;
  for (int64_t i = j + threadIdx.x; i < m; i += 256) {
    double2 vi = (i == j) ? make_double2(1.0, 0.0) : v[i];
    double2 xi = x[i];
    wr += vi.x * xi.x + vi.y * xi.y;
    wi += vi.x * xi.y - vi.y * xi.x;
  }
  wr = wg_sum(wr);
  wi = wg_sum(wi);
  // f = conj(tau) * w
  const double fr = t.x * wr + t.y * wi, fi = t.x * wi - t.y * wr;
  for (int64_t i = j + threadIdx.x; i < m; i += 256) {
    double2 vi = (i == j) ? make_double2(1.0, 0.0) : v[i];
    double2 xi = x[i];
    xi.x -= vi.x * fr - vi.y * fi;
    xi.y -= vi.x * fi + vi.y * fr;
    x[i] = xi;
  }
}

// Q (column-major m x k in Qc) := identity columns
__global__ void k_q_init(int64_t m, int64_t k, double2* __restrict__ Qc) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * k) return;
  int64_t c = t / m, i = t % m;
  Qc[t] = make_double2(i == c ? 1.0 : 0.0, 0.0);
}

// Q[:, c] <- H_j Q[:, c] for c = j .. k-1, H_j = I - tau_j v v^H
__global__ __launch_bounds__(256) void k_q_apply(int64_t m, int64_t k, int64_t j,
                                                 const double2* __restrict__ W,
                                                 const double2* __restrict__ tau,
                                                 double2* __restrict__ Qc) {
  const int64_t c = j + blockIdx.x;
  if (c >= k) return;
  const double2 t = tau[j];
  if (t.x == 0.0 && t.y == 0.0) return;
  const double2* v = W + j * m;
  double2* x = Qc + c * m;
  double wr = 0, wi = 0;
  for (int64_t i = j + threadIdx.x; i < m; i += 256) {
    double2 vi = (i == j) ? make_double2(1.0, 0.0) : v[i];
    double2 xi = x[i];
    wr += vi.x * xi.x + vi.y * xi.y;
    wi += vi.x * xi.y - vi.y * xi.x;
  }
  wr = wg_sum(wr);
  wi = wg_sum(wi);
  const double fr = t.x * wr - t.y * wi, fi = t.x * wi + t.y * wr;  // tau * w
  for (int64_t i = j + threadIdx.x; i < m; i += 256) {
    double2 vi = (i == j) ? make_double2(1.0, 0.0) : v[i];
    double2 xi = x[i];
    xi.x -= vi.x * fr - vi.y * fi;
    xi.y -= vi.x * fi + vi.y * fr;
    x[i] = xi;
  }
}

// Q (row-major, ldq) from Qc; R (row-major k x n, ldr) = upper triangle of W
__global__ void k_qr_out(int64_t m, int64_t n, int64_t k, const double2* __restrict__ W,
                         const double2* __restrict__ Qc, double2* __restrict__ q, int64_t ldq,
                         double2* __restrict__ r, int64_t ldr) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m * k) {
    int64_t i = t / k, c = t % k;
    q[i * ldq + c] = Qc[c * m + i];
  }
  if (t < k * n) {
    int64_t i = t / n, c = t % n;
    r[i * ldr + c] = (i <= c) ? W[c * m + i] : make_double2(0.0, 0.0);
  }
}

static inline size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

size_t qr_ws_bytes(int64_t m, int64_t n) {
  int64_t k = m < n ? m : n;
  return al((size_t)m * n * 16) + al((size_t)m * k * 16) + al((size_t)(k + 1) * 16);
}

int qr_run(int64_t m, int64_t n, const void* a, int64_t lda, void* q, int64_t ldq, void* r,
           int64_t ldr, void* ws, size_t ws_bytes, hipStream_t s) {
  if (m < 0 || n < 0 || lda < n || ldq < (m < n ? m : n) || ldr < n) return PTN_ERR_ARG;
  if (ws_bytes < qr_ws_bytes(m, n)) return PTN_ERR_WORKSPACE;
  const int64_t k = m < n ? m : n;
  if (m == 0 || n == 0) return PTN_OK;
  char* p = (char*)ws;
  double2* W = (double2*)p;
  p += al((size_t)m * n * 16);
  double2* Qc = (double2*)p;
  p += al((size_t)m * k * 16);
  double2* tau = (double2*)p;
  k_to_colmajor<<<(unsigned)((m * n + 255) / 256), 256, 0, s>>>(m, n, (const double2*)a, lda, W);
  PTN_LAUNCH_CHECK();
  for (int64_t j = 0; j < k; ++j) {
    k_qr_reflector<<<1, 256, 0, s>>>(m, j, W, tau);
    if (j + 1 < n) k_qr_apply<<<(unsigned)(n - j - 1), 256, 0, s>>>(m, n, j, W, tau);
  }
  PTN_LAUNCH_CHECK();
  k_q_init<<<(unsigned)((m * k + 255) / 256), 256, 0, s>>>(m, k, Qc);
  for (int64_t j = k - 1; j >= 0; --j)
    k_q_apply<<<(unsigned)(k - j), 256, 0, s>>>(m, k, j, W, tau, Qc);
  const int64_t tot = (m * k > k * n) ? m * k : k * n;
  k_qr_out<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(m, n, k, W, Qc, (double2*)q, ldq,
                                                          (double2*)r, ldr);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

}  // namespace ptn
