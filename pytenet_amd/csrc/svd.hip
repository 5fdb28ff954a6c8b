// Thin SVD of a complex128 matrix by one-sided (Hestenes) Jacobi in FP64.
//
// X = A (m >= n) or X = A^H (m < n) is held column-major; pairs of columns
// are orthogonalised by plane rotations until every pair satisfies
// |x_p^H x_q| <= tol * |x_p| |x_q|. Then sigma_j = |x_j|, U_j = x_j / sigma_j
// and V accumulates the rotations: X V = U diag(sigma). One-sided Jacobi
// computes small singular values to high relative accuracy, which the
// discarded-weight truncation relies on.
//
// Column pairs follow the round-robin (circle) tournament, so each round is
// q/2 independent pair rotations, one workgroup each. The host checks the
// rotation count once per sweep.
//
// Rotation for a pair (x, y) with a = |x|^2, b = |y|^2, c = x^H y = |c| e^{i phi}:
//   zeta = (b - a) / (2|c|), t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
//   cs = 1 / sqrt(1 + t^2), sn = cs t,
//   x' = cs x - sn e^{-i phi} y,   y' = sn e^{i phi} x + cs y.
#include <float.h>
#include "common.h"

namespace ptn {

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

// X (column-major p x q): from A (m x n, lda) directly or as A^H
__global__ void k_svd_load(int64_t m, int64_t n, const double2* __restrict__ a, int64_t lda,
                           int adjoint, double2* __restrict__ X) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * n) return;
  int64_t i = t / n, j = t % n;
  double2 v = a[i * lda + j];
  if (!adjoint) X[j * m + i] = v;                              // X = A, p = m
  else X[i * n + j] = make_double2(v.x, -v.y);                 // X = A^H, p = n, X[j][i]
}

__global__ void k_eye(int64_t q, double2* __restrict__ V) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= q * q) return;
  V[t] = make_double2((t / q) == (t % q) ? 1.0 : 0.0, 0.0);
}

__device__ __forceinline__ void round_pair(int64_t qe, int64_t r, int64_t k, int64_t& p,
                                           int64_t& q) {
  const int64_t cyc = qe - 1;
  if (k == 0) { p = r; q = cyc; }
  else { p = (r + k) % cyc; q = (r - k + cyc) % cyc; }
  if (p > q) { int64_t t = p; p = q; q = t; }
}

__global__ __launch_bounds__(256) void k_jacobi_round(int64_t pr, int64_t qn, int64_t qe,
                                                      int64_t r, double2* __restrict__ X,
                                                      double2* __restrict__ V, double tol,
                                                      int* __restrict__ rotations) {
  __shared__ double red[256];
  int64_t p, q;
  round_pair(qe, r, blockIdx.x, p, q);
  if (q >= qn) return;   // padding column
  double2* x = X + p * pr;
  double2* y = X + q * pr;
  double a = 0, b = 0, cr = 0, ci = 0;
  for (int64_t i = threadIdx.x; i < pr; i += 256) {
    double2 u = x[i], v = y[i];
    a += u.x * u.x + u.y * u.y;
    b += v.x * v.x + v.y * v.y;
    cr += u.x * v.x + u.y * v.y;     // conj(u) v
    ci += u.x * v.y - u.y * v.x;
  }
  a = wg_sum3(a, red);
  b = wg_sum3(b, red);
  cr = wg_sum3(cr, red);
  ci = wg_sum3(ci, red);
  const double ac = hypot(cr, ci);
  if (!(ac > tol * sqrt(a) * sqrt(b))) return;   // also skips zero columns
  const double zeta = (b - a) / (2.0 * ac);
  const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double cs = 1.0 / sqrt(1.0 + t * t);
  const double sn = cs * t;
  const double er = cr / ac, ei = ci / ac;   // e^{i phi}
  // x' = cs x - sn conj(e) y ; y' = sn e x + cs y
  const double s1r = sn * er, s1i = -sn * ei;   // sn e^{-i phi}
  const double s2r = sn * er, s2i = sn * ei;    // sn e^{i phi}
  for (int64_t i = threadIdx.x; i < pr; i += 256) {
    double2 u = x[i], v = y[i];
    double2 nu, nv;
    nu.x = cs * u.x - (s1r * v.x - s1i * v.y);
    nu.y = cs * u.y - (s1r * v.y + s1i * v.x);
    nv.x = (s2r * u.x - s2i * u.y) + cs * v.x;
    nv.y = (s2r * u.y + s2i * u.x) + cs * v.y;
    x[i] = nu;
    y[i] = nv;
  }
  double2* vx = V + p * qn;
  double2* vy = V + q * qn;
  for (int64_t i = threadIdx.x; i < qn; i += 256) {
    double2 u = vx[i], v = vy[i];
    double2 nu, nv;
    nu.x = cs * u.x - (s1r * v.x - s1i * v.y);
    nu.y = cs * u.y - (s1r * v.y + s1i * v.x);
    nv.x = (s2r * u.x - s2i * u.y) + cs * v.x;
    nv.y = (s2r * u.y + s2i * u.x) + cs * v.y;
    vx[i] = nu;
    vy[i] = nv;
  }
  if (threadIdx.x == 0) atomicAdd(rotations, 1);
}

// column norms of X
__global__ __launch_bounds__(256) void k_colnorms(int64_t pr, const double2* __restrict__ X,
                                                  double* __restrict__ nrm) {
  __shared__ double red[256];
  const double2* x = X + (int64_t)blockIdx.x * pr;
  double s = 0;
  for (int64_t i = threadIdx.x; i < pr; i += 256) {
    double2 u = x[i];
    s += u.x * u.x + u.y * u.y;
  }
  s = wg_sum3(s, red);
  if (threadIdx.x == 0) nrm[blockIdx.x] = sqrt(s);
}

// descending order of the q column norms (stable: ties keep column order),
// one workgroup; rank of column j = #{i : s_i > s_j} + #{i < j : s_i == s_j}
__global__ __launch_bounds__(256) void k_rank(int64_t q, const double* __restrict__ nrm,
                                              int64_t* __restrict__ perm, double* __restrict__ s_out,
                                              int* __restrict__ nonfinite) {
  // a non-finite norm is ordered as -1 (norms are >= 0) so that the ranks
  // stay a permutation; the flag reports it
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

// outputs in row-major: for X = A: U[i][c] = X[perm c][i] / s_c, Vh[c][j] = conj(V[j][perm c]);
// for X = A^H: U[i][c] = V[i][perm c], Vh[c][j] = conj(X[j][perm c]) / s_c
__global__ void k_svd_out(int64_t m, int64_t n, int64_t k, int adjoint, const double2* __restrict__ X,
                          const double2* __restrict__ V, const int64_t* __restrict__ perm,
                          const double* __restrict__ s, double2* __restrict__ u, int64_t ldu,
                          double2* __restrict__ vh, int64_t ldvh) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m * k) {
    int64_t i = t / k, c = t % k;
    int64_t pc = perm[c];
    double2 val;
    if (!adjoint) {
      double2 x = X[pc * m + i];
      double sc = s[c];
      val = sc > 0 ? make_double2(x.x / sc, x.y / sc) : make_double2(0.0, 0.0);
    } else {
      val = V[pc * m + i];     // V is m x m here (q = m)
    }
    u[i * ldu + c] = val;
  }
  if (t < k * n) {
    int64_t c = t / n, j = t % n;
    int64_t pc = perm[c];
    double2 val;
    if (!adjoint) {
      double2 v = V[pc * n + j];   // V is n x n (q = n), column pc
      val = make_double2(v.x, -v.y);
    } else {
      double2 x = X[pc * n + j];   // X is n x m column-major (p = n)
      double sc = s[c];
      val = sc > 0 ? make_double2(x.x / sc, -x.y / sc) : make_double2(0.0, 0.0);
    }
    vh[c * ldvh + j] = val;
  }
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

size_t svd_ws_bytes(int64_t m, int64_t n) {
  const int64_t q = m < n ? m : n, p = m < n ? n : m;
  return al((size_t)p * q * 16) + al((size_t)q * q * 16) + al((size_t)q * 8) * 2 +
         al((size_t)q * 8) + al(16);
}

int svd_run(int64_t m, int64_t n, const void* a, int64_t lda, void* u, int64_t ldu, double* s,
            void* vh, int64_t ldvh, int32_t* info, void* ws, size_t ws_bytes, hipStream_t st) {
  if (m <= 0 || n <= 0 || lda < n) return PTN_ERR_ARG;
  const int64_t k = m < n ? m : n;
  if (ldu < k || ldvh < n) return PTN_ERR_ARG;
  if (ws_bytes < svd_ws_bytes(m, n)) return PTN_ERR_WORKSPACE;
  const int adjoint = m < n;
  const int64_t pr = adjoint ? n : m;    // rows of X
  const int64_t qn = k;                  // columns of X
  char* p = (char*)ws;
  double2* X = (double2*)p; p += al((size_t)pr * qn * 16);
  double2* V = (double2*)p; p += al((size_t)qn * qn * 16);
  double* nrm = (double*)p; p += al((size_t)qn * 8);
  double* sbuf = (double*)p; p += al((size_t)qn * 8);
  int64_t* perm = (int64_t*)p; p += al((size_t)qn * 8);
  int* counters = (int*)p;   // [0] rotations in the current sweep, [1] non-finite flag

  k_svd_load<<<(unsigned)((m * n + 255) / 256), 256, 0, st>>>(m, n, (const double2*)a, lda,
                                                               adjoint, X);
  k_eye<<<(unsigned)((qn * qn + 255) / 256), 256, 0, st>>>(qn, V);
  PTN_LAUNCH_CHECK();
  int* host_rot = pinned_counter();
  if (!host_rot) return PTN_ERR_HIP;
  const double tol = DBL_EPSILON * sqrt((double)pr);
  const int64_t qe = (qn % 2) ? qn + 1 : qn;
  const int max_sweeps = 60;
  int sweeps = 0, converged = (qn < 2);
  k_set_ints<<<1, 64, 0, st>>>(counters, 2, 0);
  while (!converged && sweeps < max_sweeps) {
    k_set_ints<<<1, 64, 0, st>>>(counters, 1, 0);
    for (int64_t r = 0; r < qe - 1; ++r)
      k_jacobi_round<<<(unsigned)(qe / 2), 256, 0, st>>>(pr, qn, qe, r, X, V, tol, counters);
    PTN_LAUNCH_CHECK();
    PTN_HIP_CHECK(hipMemcpyAsync(host_rot, counters, sizeof(int), hipMemcpyDeviceToHost, st));
    PTN_HIP_CHECK(hipStreamSynchronize(st));
    ++sweeps;
    converged = (*host_rot == 0);
  }
  k_colnorms<<<(unsigned)qn, 256, 0, st>>>(pr, X, nrm);
  k_rank<<<1, 256, 0, st>>>(qn, nrm, perm, s, counters + 1);
  const int64_t tot = (m * k > k * n) ? m * k : k * n;
  k_svd_out<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(m, n, k, adjoint, X, V, perm, s,
                                                            (double2*)u, ldu, (double2*)vh, ldvh);
  PTN_LAUNCH_CHECK();
  if (info) k_svd_info<<<1, 64, 0, st>>>(info, sweeps, converged ? 0 : 1, counters);
  PTN_LAUNCH_CHECK();
  (void)sbuf;
  return PTN_OK;
}

}  // namespace ptn
